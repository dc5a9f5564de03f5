// capi.hip -- C ABI of the stage functions: validation and error mapping of
// the reference's free functions, then the HIP kernels.
//
// Each entry validates its arguments as the reference function it replaces
// does (same conditions, same exception class -> status, same message),
// runs on the context's stream and synchronises before it returns, as the
// reference's wrappers do.
#include "nbh_internal.hpp"

#include <cmath>
#include <cstring>
#include <limits>

namespace nbh {

namespace {

void need(bool ok, int status, const char* message)
{
    if (!ok)
        fail(status, message);
}

int read_error(nbh_context* ctx, int word)
{
    int h[ERR_WORD_COUNT];
    sync_read(ctx, error_words(ctx), h, ERR_WORD_COUNT);
    return h[word];
}

void clear_errors(nbh_context* ctx)
{
    launch_fill_i32(ctx->stream, error_words(ctx), 0, 64);
}

const char* plan_message(int code)
{
    switch (code) {
    case 1: return "A Morton key uses bits outside the configured spatial layout";
    case 2: return "A radix node has an invalid parent";
    case 3: return "A radix child has an octree level smaller than its parent";
    case 4: return "A radix internal node has an invalid prefix length";
    default: return "Unknown radix-to-octree planning error";
    }
}

const char* octree_message(int code)
{
    switch (code) {
    case 1: return "Invalid radix edge metadata";
    case 2: return "Invalid radix ancestor while constructing octree";
    case 3: return "Could not find the correct octree parent level";
    case 4: return "Two octree nodes attempted to occupy the same parent octant";
    case 5: return "A radix leaf terminates at an invalid octree level";
    case 6: return "Generated octree node ID is outside the allocated range";
    case 7: return "Invalid representative Morton key";
    default: return "Unknown sparse-octree construction error";
    }
}

const char* geometry_message(int code)
{
    switch (code) {
    case 1: return "Octree node has an invalid level";
    case 2: return "Octree node has an invalid Morton prefix";
    default: return "Unknown octree geometry error";
    }
}

const char* physics_message(int code)
{
    switch (code) {
    case 1: return "Octree contains an invalid child ID";
    case 2: return "Occupied octree leaf has invalid particle range";
    case 3: return "Sorted particle index is outside the particle array";
    case 4: return "Occupied octree leaf has children";
    case 5: return "Invalid octree parent during bottom-up propagation";
    case 6: return "Invalid octree child count";
    case 7: return "An octree parent was signalled too many times";
    default: return "Unknown octree physics error";
    }
}

} // namespace

const char* traversal_message(int code)
{
    switch (code) {
    case 1: return "Barnes-Hut traversal found an invalid child node";
    case 2: return "Barnes-Hut traversal found an invalid leaf particle range";
    case 3: return "Sorted particle index is outside the particle array";
    case 4: return "Barnes-Hut traversal found inconsistent parent-child topology";
    case 5: return "Barnes-Hut traversal found an empty terminal node";
    case 6: return "Singular gravitational interaction; use positive softening";
    case 7: return "Occupied octree leaf unexpectedly has children";
    case 8: return "Octree node has invalid geometry";
    case 9: return "Octree leaf holds 2^23 or more particles on one Morton key, more than "
                   "the mixed-precision walk records can count; use the FP64 pipeline";
    default: return "Unknown Barnes-Hut traversal error";
    }
}

const char* stage_message(int word, int code)
{
    switch (word) {
    case ERR_PLAN: return plan_message(code);
    case ERR_OCTREE: return octree_message(code);
    case ERR_GEOMETRY: return geometry_message(code);
    case ERR_PHYSICS: return physics_message(code);
    case ERR_TRAVERSAL: return traversal_message(code);
    default: return "Particle positions must be finite to compute a bounding box.";
    }
}

int passes_for_bits(int key_bits)
{
    int p = (key_bits + 7) / 8;
    if (p % 2)
        ++p;
    return p < 2 ? 2 : p;
}

// Sort (key, value) pairs whose digit histograms are already in `hist`.
template <typename K>
void sort_with_hist(nbh_context* ctx, K* keys, uint32_t* vals, size_t n,
                    int passes, uint32_t* hist)
{
    K* kalt = static_cast<K*>(ctx->ws.get(SLOT_SORT_KEYS_ALT, n * sizeof(K)));
    uint32_t* valt =
        static_cast<uint32_t*>(ctx->ws.get(SLOT_SORT_VALS_ALT, n * 4));
    const size_t words = sort_status_words(n, passes);
    uint32_t* status =
        static_cast<uint32_t*>(ctx->ws.get(SLOT_SORT_STATUS, words * 4));
    launch_fill_u32(ctx->stream, status, 0u, words);
    launch_radix_passes<K>(ctx->stream, keys, vals, kalt, valt, n, passes,
                           hist, status);
}

template <int BITS>
void morton_keys(nbh_context* ctx, const double* x, const double* y,
                 const double* z, size_t n, const double* box, void* keys,
                 uint32_t* idx)
{
    using K = typename Layout<BITS>::Key;
    constexpr int P = Layout<BITS>::sort_passes;
    uint32_t* hist =
        static_cast<uint32_t*>(ctx->ws.get(SLOT_SORT_HIST, P * 256 * 4));
    launch_fill_u32(ctx->stream, hist, 0u, P * 256);
    launch_encode<BITS>(ctx->stream, x, y, z, n, box, static_cast<K*>(keys),
                        idx, hist);
    sort_with_hist<K>(ctx, static_cast<K*>(keys), idx, n, P, hist);
}

} // namespace nbh

using namespace nbh;

extern "C" {

int nbh_bbox_compute(nbh_context* ctx, const double* d_x, const double* d_y,
                     const double* d_z, size_t count, double* d_box,
                     double* h_box)
{
    return guard([&] {
        need(ctx != nullptr, NBH_INVALID_ARGUMENT, "context must not be null");
        need(count != 0, NBH_INVALID_ARGUMENT,
             "Cannot compute a bounding box for an empty particle set.");
        need(d_x && d_y && d_z && d_box, NBH_INVALID_ARGUMENT,
             "Bounding-box pointers must not be null");
        double* partials = static_cast<double*>(
            ctx->ws.get(SLOT_MISC, bbox_partial_doubles() * 8));
        clear_errors(ctx);
        launch_bbox(ctx, d_x, d_y, d_z, count, d_box, partials,
                    error_words(ctx) + ERR_BBOX);
        if (read_error(ctx, ERR_BBOX) != 0)
            fail(NBH_DOMAIN_ERROR,
                 "Particle positions must be finite to compute a bounding box.");
        if (h_box) {
            NBH_HIP(hipMemcpyAsync(h_box, d_box, 4 * sizeof(double),
                                   hipMemcpyDeviceToHost, ctx->stream));
            NBH_HIP(hipStreamSynchronize(ctx->stream));
        }
    });
}

int nbh_morton_keys(nbh_context* ctx, const double* d_x, const double* d_y,
                    const double* d_z, size_t count, const double* d_box,
                    int bits_per_axis, void* d_keys, uint32_t* d_indices)
{
    return guard([&] {
        need(ctx != nullptr, NBH_INVALID_ARGUMENT, "context must not be null");
        need(count != 0, NBH_INVALID_ARGUMENT,
             "Cannot compute Morton keys for an empty particle set");
        need(d_x && d_y && d_z, NBH_INVALID_ARGUMENT,
             "Morton coordinate device pointers must not be null");
        need(count <= std::numeric_limits<uint32_t>::max(),
             NBH_OVERFLOW_ERROR,
             "Particle count exceeds the Morton index representation");
        need(bits_per_axis == 10 || bits_per_axis == 21,
             NBH_INVALID_ARGUMENT, "bits_per_axis must be 10 or 21");
        need(d_box && d_keys && d_indices, NBH_INVALID_ARGUMENT,
             "Morton box and output pointers must not be null");
        double box[4];
        NBH_HIP(hipMemcpyAsync(box, d_box, sizeof(box), hipMemcpyDeviceToHost,
                               ctx->stream));
        NBH_HIP(hipStreamSynchronize(ctx->stream));
        need(std::isfinite(box[0]) && std::isfinite(box[1]) &&
                 std::isfinite(box[2]) && std::isfinite(box[3]) &&
                 box[3] > 0.0,
             NBH_INVALID_ARGUMENT, "Morton bounding box is not valid");
        if (bits_per_axis == 10)
            morton_keys<10>(ctx, d_x, d_y, d_z, count, d_box, d_keys,
                            d_indices);
        else
            morton_keys<21>(ctx, d_x, d_y, d_z, count, d_box, d_keys,
                            d_indices);
        NBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int nbh_radix_sort_pairs(nbh_context* ctx, void* d_keys, uint32_t* d_values,
                         size_t count, int key_bytes, int key_bits)
{
    return guard([&] {
        need(ctx != nullptr, NBH_INVALID_ARGUMENT, "context must not be null");
        need(key_bytes == 4 || key_bytes == 8, NBH_INVALID_ARGUMENT,
             "key_bytes must be 4 or 8");
        need(key_bits >= 1 && key_bits <= key_bytes * 8, NBH_INVALID_ARGUMENT,
             "key_bits out of range");
        if (count == 0)
            return;
        need(d_keys && d_values, NBH_INVALID_ARGUMENT,
             "sort pointers must not be null");
        const int passes = passes_for_bits(key_bits);
        uint32_t* hist =
            static_cast<uint32_t*>(ctx->ws.get(SLOT_SORT_HIST, passes * 256 * 4));
        launch_fill_u32(ctx->stream, hist, 0u, passes * 256);
        if (key_bytes == 4) {
            launch_histogram<uint32_t>(ctx->stream,
                                       static_cast<uint32_t*>(d_keys), count,
                                       passes, hist);
            sort_with_hist<uint32_t>(ctx, static_cast<uint32_t*>(d_keys),
                                     d_values, count, passes, hist);
        } else {
            launch_histogram<uint64_t>(ctx->stream,
                                       static_cast<uint64_t*>(d_keys), count,
                                       passes, hist);
            sort_with_hist<uint64_t>(ctx, static_cast<uint64_t*>(d_keys),
                                     d_values, count, passes, hist);
        }
        NBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int nbh_leaf_groups_build(nbh_context* ctx, const void* d_sorted_keys,
                          int key_bytes, int n_particles, int capacity,
                          void* d_unique_keys, int* d_first_particle,
                          int* d_particle_count, int* h_n_groups)
{
    return guard([&] {
        need(ctx != nullptr, NBH_INVALID_ARGUMENT, "context must not be null");
        need(n_particles > 0, NBH_INVALID_ARGUMENT,
             "buildMortonLeafGroups requires nParticles >= 1");
        need(n_particles <= capacity, NBH_INVALID_ARGUMENT,
             "nParticles exceeds MortonLeafGroups capacity");
        need(d_sorted_keys != nullptr, NBH_INVALID_ARGUMENT,
             "d_sortedKeys must not be null");
        need(d_unique_keys && d_first_particle && d_particle_count,
             NBH_LOGIC_ERROR, "MortonLeafGroups memory has not been allocated");
        need(key_bytes == 4 || key_bytes == 8, NBH_INVALID_ARGUMENT,
             "key_bytes must be 4 or 8");
        const size_t words = scan_status_words(n_particles);
        char* misc = static_cast<char*>(ctx->ws.get(SLOT_SCAN_STATUS, words * 8 + 256));
        auto* status = reinterpret_cast<unsigned long long*>(misc);
        int* d_ng = reinterpret_cast<int*>(misc + words * 8);
        launch_fill_u64(ctx->stream, status, 0ull, words);
        if (key_bytes == 4)
            launch_leaf_groups<uint32_t>(
                ctx->stream, static_cast<const uint32_t*>(d_sorted_keys),
                n_particles, static_cast<uint32_t*>(d_unique_keys),
                d_first_particle, d_particle_count, d_ng, status);
        else
            launch_leaf_groups<uint64_t>(
                ctx->stream, static_cast<const uint64_t*>(d_sorted_keys),
                n_particles, static_cast<uint64_t*>(d_unique_keys),
                d_first_particle, d_particle_count, d_ng, status);
        int g = 0;
        sync_read(ctx, d_ng, &g, 1);
        if (g <= 0 || g > n_particles)
            fail(NBH_RUNTIME_ERROR,
                 "DeviceRunLengthEncode produced an invalid group count");
        if (h_n_groups)
            *h_n_groups = g;
    });
}

int nbh_tree_build(nbh_context* ctx, const nbh_tree_view* tree,
                   const void* d_sorted_keys, int key_bytes, int n_leaves)
{
    return guard([&] {
        need(ctx && tree, NBH_INVALID_ARGUMENT, "context and tree must not be null");
        need(n_leaves > 0, NBH_INVALID_ARGUMENT, "buildTree requires N >= 1");
        need(tree->n_leaves == n_leaves, NBH_INVALID_ARGUMENT,
             "buildTree N does not match allocated tree size");
        need(tree->parent != nullptr, NBH_LOGIC_ERROR,
             "Tree memory has not been allocated");
        need(n_leaves == 1 ||
                 (tree->left && tree->right && tree->range_first &&
                  tree->range_last && tree->prefix_length),
             NBH_LOGIC_ERROR,
             "Tree topology and metadata arrays have not been allocated");
        need(d_sorted_keys != nullptr, NBH_INVALID_ARGUMENT,
             "d_sortedKeys must not be null");
        need(key_bytes == 4 || key_bytes == 8, NBH_INVALID_ARGUMENT,
             "key_bytes must be 4 or 8");
        launch_fill_i32(ctx->stream, tree->parent, -1, 2 * static_cast<size_t>(n_leaves) - 1);
        if (n_leaves > 1) {
            if (key_bytes == 4)
                launch_karras<uint32_t>(ctx->stream,
                                        static_cast<const uint32_t*>(d_sorted_keys),
                                        nullptr, n_leaves, tree->left, tree->right,
                                        tree->parent, tree->range_first,
                                        tree->range_last, tree->prefix_length);
            else
                launch_karras<uint64_t>(ctx->stream,
                                        static_cast<const uint64_t*>(d_sorted_keys),
                                        nullptr, n_leaves, tree->left, tree->right,
                                        tree->parent, tree->range_first,
                                        tree->range_last, tree->prefix_length);
        }
        NBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int nbh_tree_center_of_mass(nbh_context* ctx, const nbh_tree_view* tree,
                            const int* d_first_particle,
                            const int* d_particle_count,
                            const uint32_t* d_sorted_indices,
                            const double* d_mass, const double* d_x,
                            const double* d_y, const double* d_z)
{
    return guard([&] {
        need(ctx && tree, NBH_INVALID_ARGUMENT, "context and tree must not be null");
        need(tree->n_leaves > 0, NBH_INVALID_ARGUMENT,
             "computeCentersOfMass requires N >= 1");
        need((d_first_particle == nullptr) == (d_particle_count == nullptr),
             NBH_LOGIC_ERROR, "MortonLeafGroups arrays are not allocated");
        need(d_sorted_indices != nullptr, NBH_INVALID_ARGUMENT,
             "d_sortedIndices must not be null");
        need(d_mass && d_x && d_y && d_z, NBH_INVALID_ARGUMENT,
             "Particle data arrays must not be null");
        need(tree->parent && tree->mass && tree->com_x && tree->com_y &&
                 tree->com_z,
             NBH_LOGIC_ERROR, "Tree memory has not been fully allocated");
        need(tree->n_leaves == 1 ||
                 (tree->left && tree->right && tree->visit_count),
             NBH_LOGIC_ERROR, "Tree internal-node memory has not been allocated");
        launch_tree_com(ctx->stream, *tree, d_first_particle, d_particle_count,
                        d_sorted_indices, d_mass, d_x, d_y, d_z);
        NBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int nbh_radix_to_octree_plan(nbh_context* ctx, const nbh_tree_view* tree,
                             const void* d_unique_keys, int bits_per_axis,
                             int* d_radix_level, int* d_edge_node_count,
                             int* d_edge_node_offset, int* h_n_octree_nodes)
{
    return guard([&] {
        need(ctx && tree, NBH_INVALID_ARGUMENT, "context and tree must not be null");
        need(tree->n_leaves > 0, NBH_INVALID_ARGUMENT, "The radix tree has no leaves");
        need(d_unique_keys != nullptr, NBH_LOGIC_ERROR,
             "Morton unique keys are not allocated");
        need(tree->parent != nullptr, NBH_LOGIC_ERROR,
             "Radix-tree parent array is not allocated");
        need(tree->n_leaves == 1 || tree->prefix_length != nullptr,
             NBH_LOGIC_ERROR, "Radix-tree prefix metadata is not allocated");
        need(d_radix_level && d_edge_node_count && d_edge_node_offset,
             NBH_LOGIC_ERROR, "RadixToOctreePlan memory is not allocated");
        need(bits_per_axis == 10 || bits_per_axis == 21, NBH_INVALID_ARGUMENT,
             "bits_per_axis must be 10 or 21");
        const int nl = tree->n_leaves;
        const size_t words = scan_status_words(2 * static_cast<size_t>(nl));
        char* misc = static_cast<char*>(ctx->ws.get(SLOT_SCAN_STATUS, words * 8 + 256));
        auto* status = reinterpret_cast<unsigned long long*>(misc);
        int* d_ints = reinterpret_cast<int*>(misc + words * 8);
        launch_fill_u64(ctx->stream, status, 0ull, words);
        launch_fill_i32(ctx->stream, d_ints, nl, 1);
        launch_fill_i32(ctx->stream, d_ints + 1, 0, 1);
        clear_errors(ctx);
        int* err = error_words(ctx) + ERR_PLAN;
        if (bits_per_axis == 10)
            launch_plan_scan<10>(ctx->stream, d_ints, nl, tree->parent,
                                 tree->prefix_length,
                                 static_cast<const uint32_t*>(d_unique_keys),
                                 d_radix_level, d_edge_node_count,
                                 d_edge_node_offset, d_ints + 1, err, status);
        else
            launch_plan_scan<21>(ctx->stream, d_ints, nl, tree->parent,
                                 tree->prefix_length,
                                 static_cast<const uint64_t*>(d_unique_keys),
                                 d_radix_level, d_edge_node_count,
                                 d_edge_node_offset, d_ints + 1, err, status);
        const int code = read_error(ctx, ERR_PLAN);
        if (code != 0)
            fail(NBH_RUNTIME_ERROR, plan_message(code), code);
        int m = 0;
        sync_read(ctx, d_ints + 1, &m, 1);
        if (m <= 0)
            fail(NBH_OVERFLOW_ERROR,
                 "The computed octree node count does not fit in int");
        if (h_n_octree_nodes)
            *h_n_octree_nodes = m;
    });
}

int nbh_octree_build_topology(nbh_context* ctx, const nbh_octree_view* octree,
                              const nbh_tree_view* tree,
                              const void* d_unique_keys,
                              const int* d_group_first,
                              const int* d_group_count, int n_groups,
                              int n_particles, int bits_per_axis,
                              const int* d_radix_level,
                              const int* d_edge_node_count,
                              const int* d_edge_node_offset)
{
    return guard([&] {
        need(ctx && octree && tree, NBH_INVALID_ARGUMENT,
             "context, octree and tree must not be null");
        need(n_particles > 0 && n_groups > 0, NBH_INVALID_ARGUMENT,
             "Morton groups are empty");
        need(n_groups == tree->n_leaves, NBH_INVALID_ARGUMENT,
             "Morton groups and radix tree have different leaf counts");
        need(d_unique_keys && d_group_first && d_group_count, NBH_LOGIC_ERROR,
             "Morton group arrays are not allocated");
        need(octree->n_nodes > 0, NBH_INVALID_ARGUMENT,
             "Radix-to-octree plan has no octree nodes");
        need(octree->children && octree->parent && octree->level &&
                 octree->prefix && octree->first_particle &&
                 octree->particle_count,
             NBH_LOGIC_ERROR, "Octree topology memory is not allocated");
        need(tree->parent != nullptr, NBH_LOGIC_ERROR,
             "Radix parent array is not allocated");
        need(tree->n_leaves == 1 || (tree->range_first && tree->prefix_length),
             NBH_LOGIC_ERROR, "Radix metadata is incomplete");
        need(d_radix_level && d_edge_node_count && d_edge_node_offset,
             NBH_LOGIC_ERROR, "Radix-to-octree plan arrays are not allocated");
        need(bits_per_axis == 10 || bits_per_axis == 21, NBH_INVALID_ARGUMENT,
             "bits_per_axis must be 10 or 21");
        clear_errors(ctx);
        int* err = error_words(ctx) + ERR_OCTREE;
        if (bits_per_axis == 10)
            launch_octree_build<10>(ctx->stream, *octree, n_groups, tree->parent,
                                    tree->range_first,
                                    static_cast<const uint32_t*>(d_unique_keys),
                                    d_group_first, d_group_count, d_radix_level,
                                    d_edge_node_count, d_edge_node_offset, err,
                                    false);
        else
            launch_octree_build<21>(ctx->stream, *octree, n_groups, tree->parent,
                                    tree->range_first,
                                    static_cast<const uint64_t*>(d_unique_keys),
                                    d_group_first, d_group_count, d_radix_level,
                                    d_edge_node_count, d_edge_node_offset, err,
                                    false);
        const int code = read_error(ctx, ERR_OCTREE);
        if (code != 0)
            fail(NBH_RUNTIME_ERROR, octree_message(code), code);
    });
}

int nbh_octree_geometry(nbh_context* ctx, const nbh_octree_view* octree,
                        int bits_per_axis, const double* h_box)
{
    return guard([&] {
        need(ctx && octree, NBH_INVALID_ARGUMENT, "context and octree must not be null");
        need(octree->n_nodes > 0 && octree->root == 0, NBH_LOGIC_ERROR,
             "Octree topology is not valid");
        need(octree->level && octree->prefix, NBH_LOGIC_ERROR,
             "Octree topology arrays required for geometry are not allocated");
        need(octree->center_x && octree->center_y && octree->center_z &&
                 octree->half_size,
             NBH_LOGIC_ERROR, "Octree geometry arrays are not allocated");
        need(h_box != nullptr, NBH_INVALID_ARGUMENT, "box must not be null");
        need(std::isfinite(h_box[0]) && std::isfinite(h_box[1]) &&
                 std::isfinite(h_box[2]) && std::isfinite(h_box[3]) &&
                 h_box[3] > 0.0,
             NBH_INVALID_ARGUMENT, "Bounding box geometry is not valid");
        need(bits_per_axis == 10 || bits_per_axis == 21, NBH_INVALID_ARGUMENT,
             "bits_per_axis must be 10 or 21");
        clear_errors(ctx);
        int* err = error_words(ctx) + ERR_GEOMETRY;
        if (bits_per_axis == 10)
            launch_geometry<10>(ctx->stream, *octree, h_box[0], h_box[1],
                                h_box[2], h_box[3], err);
        else
            launch_geometry<21>(ctx->stream, *octree, h_box[0], h_box[1],
                                h_box[2], h_box[3], err);
        const int code = read_error(ctx, ERR_GEOMETRY);
        if (code != 0)
            fail(NBH_RUNTIME_ERROR, geometry_message(code), code);
    });
}

int nbh_octree_mass_com(nbh_context* ctx, const nbh_octree_view* octree,
                        const uint32_t* d_sorted_indices, const double* d_mass,
                        const double* d_x, const double* d_y,
                        const double* d_z)
{
    return guard([&] {
        need(ctx && octree, NBH_INVALID_ARGUMENT, "context and octree must not be null");
        need(octree->n_particles > 0 && octree->n_leaves > 0 &&
                 octree->n_nodes > 0 && octree->root == 0,
             NBH_LOGIC_ERROR, "Octree topology is not valid");
        need(octree->children && octree->parent && octree->first_particle &&
                 octree->particle_count,
             NBH_LOGIC_ERROR, "Octree topology arrays are not allocated");
        need(octree->mass && octree->com_x && octree->com_y && octree->com_z &&
                 octree->pending_children && octree->child_count,
             NBH_LOGIC_ERROR, "Octree physical arrays are not allocated");
        need(d_sorted_indices != nullptr, NBH_INVALID_ARGUMENT,
             "d_sortedIndices must not be null");
        need(d_mass && d_x && d_y && d_z, NBH_INVALID_ARGUMENT,
             "Particle arrays must not be null");
        const size_t m = static_cast<size_t>(octree->n_nodes);
        hipStream_t s = ctx->stream;
        launch_fill_f64(s, octree->mass, 0.0, m);
        launch_fill_f64(s, octree->com_x, 0.0, m);
        launch_fill_f64(s, octree->com_y, 0.0, m);
        launch_fill_f64(s, octree->com_z, 0.0, m);
        launch_fill_i32(s, octree->pending_children, 0, m);
        launch_fill_i32(s, octree->child_count, 0, m);
        clear_errors(ctx);
        int* err = error_words(ctx) + ERR_PHYSICS;
        launch_count_children(s, *octree, err);
        int code = read_error(ctx, ERR_PHYSICS);
        if (code != 0)
            fail(NBH_RUNTIME_ERROR, physics_message(code), code);
        launch_leaf_init(s, *octree, d_sorted_indices, d_mass, d_x, d_y, d_z, err);
        code = read_error(ctx, ERR_PHYSICS);
        if (code != 0)
            fail(NBH_RUNTIME_ERROR, physics_message(code), code);
        launch_propagate(s, *octree, err);
        code = read_error(ctx, ERR_PHYSICS);
        if (code != 0)
            fail(NBH_RUNTIME_ERROR, physics_message(code), code);
    });
}

int nbh_bh_acceleration(nbh_context* ctx, const nbh_octree_view* octree,
                        const uint32_t* d_sorted_indices, const double* d_mass,
                        const double* d_x, const double* d_y,
                        const double* d_z, double* d_ax, double* d_ay,
                        double* d_az, size_t particle_count, double theta,
                        double softening, double gravitational_constant,
                        int particle_order)
{
    return guard([&] {
        need(ctx && octree, NBH_INVALID_ARGUMENT, "context and octree must not be null");
        need(octree->n_particles > 0 && octree->n_nodes > 0 && octree->root == 0,
             NBH_LOGIC_ERROR, "Octree is not valid for Barnes-Hut traversal");
        need(particle_count == static_cast<size_t>(octree->n_particles),
             NBH_INVALID_ARGUMENT, "Particle count does not match the octree");
        need(octree->children && octree->parent && octree->first_particle &&
                 octree->particle_count,
             NBH_LOGIC_ERROR,
             "Octree topology arrays required for traversal are not allocated");
        need(octree->mass && octree->com_x && octree->com_y && octree->com_z,
             NBH_LOGIC_ERROR, "Octree physical data is not allocated");
        need(octree->center_x && octree->center_y && octree->center_z &&
                 octree->half_size,
             NBH_LOGIC_ERROR, "Octree geometry is not allocated");
        need(d_sorted_indices && d_mass && d_x && d_y && d_z,
             NBH_INVALID_ARGUMENT, "Barnes-Hut particle arrays must not be null");
        need(d_ax && d_ay && d_az, NBH_INVALID_ARGUMENT,
             "Barnes-Hut acceleration arrays must not be null");
        need(std::isfinite(theta) && theta >= 0.0, NBH_INVALID_ARGUMENT,
             "Barnes-Hut theta must be finite and non-negative");
        need(std::isfinite(softening) && softening >= 0.0, NBH_INVALID_ARGUMENT,
             "Barnes-Hut softening must be finite and non-negative");
        need(std::isfinite(gravitational_constant) && gravitational_constant > 0.0,
             NBH_INVALID_ARGUMENT,
             "Gravitational constant must be finite and positive");
        need(particle_order == NBH_ORDER_ORIGINAL ||
                 particle_order == NBH_ORDER_MORTON,
             NBH_INVALID_ARGUMENT, "Unknown Barnes-Hut particle order");
        const int n = octree->n_particles;
        const size_t m = static_cast<size_t>(octree->n_nodes);
        hipStream_t s = ctx->stream;
        auto* rec = static_cast<double4*>(ctx->ws.get(SLOT_TMP0, n * sizeof(double4)));
        int* ints = static_cast<int*>(ctx->ws.get(SLOT_TMP1, 4 * m * sizeof(int) + 256));
        int* size = ints;
        int* local = ints + m;
        int* cc = ints + 2 * m;
        int* pend = ints + 3 * m;
        int* wcount = ints + 4 * m;
        // records of positions 0 .. m (m: the terminal record)
        auto* w64 = static_cast<Walk64Rec*>(ctx->ws.get(SLOT_TMP2, (m + 1) * sizeof(Walk64Rec)));
        auto* queue = static_cast<int*>(ctx->ws.get(SLOT_TMP3, kWalkQueueInts * sizeof(int)));
        const bool contains = theta >= kContainsTheta;
        auto* wcenter = contains ? static_cast<double4*>(ctx->ws.get(SLOT_TMP5, m * sizeof(double4)))
                                 : nullptr;
        nbh_octree_view o = *octree;
        o.child_count = cc;
        o.pending_children = pend;
        clear_errors(ctx);
        int* err = error_words(ctx) + ERR_TRAVERSAL;
        launch_fill_i32(s, ints, 0, 4 * m + 1);
        launch_pack_records(s, d_sorted_indices, d_mass, d_x, d_y, d_z, n,
                            particle_order == NBH_ORDER_MORTON, rec, err);
        launch_generic_sizes(s, o, size, local, err);
        launch_walk_build(s, o, size, local, theta, rec, contains, softening,
                          gravitational_constant, w64, wcenter, wcount);
        launch_walk_fp64(s, w64, m + 1, wcenter, wcount, rec, d_sorted_indices, nullptr,
                         n, softening * softening, gravitational_constant, contains, 0.0,
                         d_ax, d_ay, d_az, queue, err, nullptr);
        const int code = read_error(ctx, ERR_TRAVERSAL);
        if (code != 0)
            fail(NBH_RUNTIME_ERROR, traversal_message(code), code);
    });
}

int nbh_leapfrog_kick_drift(nbh_context* ctx, double* d_x, double* d_y,
                            double* d_z, double* d_vx, double* d_vy,
                            double* d_vz, const double* d_ax,
                            const double* d_ay, const double* d_az,
                            size_t count, double dt)
{
    return guard([&] {
        need(ctx != nullptr, NBH_INVALID_ARGUMENT, "context must not be null");
        need(std::isfinite(dt) && dt > 0.0, NBH_INVALID_ARGUMENT,
             "Leapfrog timestep must be finite and positive");
        if (count == 0)
            return;
        need(d_x && d_y && d_z, NBH_INVALID_ARGUMENT,
             "Leapfrog position arrays must not be null");
        need(d_vx && d_vy && d_vz, NBH_INVALID_ARGUMENT,
             "Leapfrog velocity arrays must not be null");
        need(d_ax && d_ay && d_az, NBH_INVALID_ARGUMENT,
             "Leapfrog acceleration arrays must not be null");
        launch_kick_drift(ctx->stream, d_x, d_y, d_z, d_vx, d_vy, d_vz, d_ax,
                          d_ay, d_az, count, dt);
        NBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int nbh_leapfrog_kick(nbh_context* ctx, double* d_vx, double* d_vy,
                      double* d_vz, const double* d_ax, const double* d_ay,
                      const double* d_az, size_t count, double dt)
{
    return guard([&] {
        need(ctx != nullptr, NBH_INVALID_ARGUMENT, "context must not be null");
        need(std::isfinite(dt) && dt > 0.0, NBH_INVALID_ARGUMENT,
             "Leapfrog timestep must be finite and positive");
        if (count == 0)
            return;
        need(d_vx && d_vy && d_vz, NBH_INVALID_ARGUMENT,
             "Leapfrog velocity arrays must not be null");
        need(d_ax && d_ay && d_az, NBH_INVALID_ARGUMENT,
             "Leapfrog acceleration arrays must not be null");
        launch_kick(ctx->stream, d_vx, d_vy, d_vz, d_ax, d_ay, d_az, count, dt);
        NBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int nbh_gather_morton(nbh_context* ctx, const uint32_t* d_sorted_indices,
                      const double* d_mass, const double* d_x,
                      const double* d_y, const double* d_z, double* d_out_mass,
                      double* d_out_x, double* d_out_y, double* d_out_z,
                      size_t count)
{
    return guard([&] {
        need(ctx != nullptr, NBH_INVALID_ARGUMENT, "context must not be null");
        need(d_sorted_indices && d_mass && d_x && d_y && d_z,
             NBH_INVALID_ARGUMENT,
             "Morton particle gather input arrays must not be null");
        need(d_out_mass && d_out_x && d_out_y && d_out_z, NBH_INVALID_ARGUMENT,
             "Morton particle gather output arrays must not be null");
        if (count == 0)
            return;
        launch_gather4(ctx->stream, d_sorted_indices, d_mass, d_x, d_y, d_z,
                       d_out_mass, d_out_x, d_out_y, d_out_z, count);
        NBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

int nbh_direct_sum(nbh_context* ctx, const double* d_mass, const double* d_x,
                   const double* d_y, const double* d_z, size_t count,
                   const uint32_t* d_targets, size_t n_targets,
                   double softening, double gravitational_constant,
                   double* d_ax, double* d_ay, double* d_az)
{
    return guard([&] {
        need(ctx != nullptr, NBH_INVALID_ARGUMENT, "context must not be null");
        need(d_mass && d_x && d_y && d_z && d_targets && d_ax && d_ay && d_az,
             NBH_INVALID_ARGUMENT, "direct-sum pointers must not be null");
        launch_direct(ctx->stream, d_mass, d_x, d_y, d_z, count, d_targets,
                      n_targets, softening * softening, gravitational_constant,
                      d_ax, d_ay, d_az);
        NBH_HIP(hipStreamSynchronize(ctx->stream));
    });
}

} // extern "C"
