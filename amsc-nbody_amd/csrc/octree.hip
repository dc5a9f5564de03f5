// octree.hip -- sparse octree from the radix tree, its geometry, and the
// reference-shaped mass / centre-of-mass phases.
//
// Node numbering is the reference's (octree_builder.cu:153-174): node 0 is
// the root; radix node r owns nodes 1 + offset[r] ... 1 + offset[r] +
// count[r] - 1, a chain from its parent's level down to its own.  Every
// derived, traversal-ordered array is built from this numbering, never in
// place of it.
#include "device_utils.hpp"
#include "nbh_internal.hpp"
#include "octree_device.hpp"

namespace nbh {

namespace {

// Initial values of octree_builder.cu:289-321 (children/parent/level/
// firstParticle -1, prefix/particleCount 0, root level 0 prefix 0), plus
// zeroed child and arrival counters for the pipeline.
template <int BITS>
__global__ __launch_bounds__(256) void octree_init_kernel(
    nbh_octree_view o, int* child_count, int* pending)
{
    using Key = typename Layout<BITS>::Key;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    const size_t m = static_cast<size_t>(o.n_nodes);
    int4* ch = reinterpret_cast<int4*>(o.children);
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < m; i += stride) {
        ch[2 * i] = make_int4(-1, -1, -1, -1);
        ch[2 * i + 1] = make_int4(-1, -1, -1, -1);
        o.parent[i] = -1;
        o.level[i] = (i == 0) ? 0 : -1;
        static_cast<Key*>(o.prefix)[i] = 0;
        o.first_particle[i] = -1;
        o.particle_count[i] = 0;
        if (child_count) {
            child_count[i] = 0;
            pending[i] = 0;
        }
    }
}

// materializeSparseOctreeKernel (octree_builder.cu:86-223), error codes 1-7.
template <int BITS>
__global__ __launch_bounds__(256) void materialize_kernel(
    nbh_octree_view o, int n_leaves, const int* __restrict__ tparent,
    const int* __restrict__ range_first,
    const typename Layout<BITS>::Key* __restrict__ unique,
    const int* __restrict__ gfirst, const int* __restrict__ gcount,
    const int* __restrict__ plevel, const int* __restrict__ pcount,
    const int* __restrict__ poffset, int* err, int* child_count)
{
    using L = Layout<BITS>;
    using Key = typename L::Key;
    const int R = 2 * n_leaves - 1;
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= R)
        return;
    const int count = pcount[node];
    const int lv = plevel[node];
    if (count < 0 || lv < 0 || lv > L::bits) {
        raise_error(err, 1);
        return;
    }
    if (count == 0) {
        if (node < n_leaves)
            raise_error(err, 5);
        return;
    }
    const int parent_level = lv - count;
    if (parent_level < 0) {
        raise_error(err, 1);
        return;
    }
    Key key;
    if (node < n_leaves) {
        key = unique[node];
    } else {
        const int f = range_first[node - n_leaves];
        if (f < 0 || f >= n_leaves) {
            raise_error(err, 7);
            return;
        }
        key = unique[f];
    }
    int parent_node = 0;
    int anc = tparent[node];
    int iterations = 0;
    while (anc != -1) {
        if (anc < 0 || anc >= R) {
            raise_error(err, 2);
            return;
        }
        const int ac = pcount[anc];
        if (ac > 0) {
            if (plevel[anc] != parent_level) {
                raise_error(err, 3);
                return;
            }
            parent_node = 1 + poffset[anc] + ac - 1;
            if (parent_node <= 0 || parent_node >= o.n_nodes) {
                raise_error(err, 6);
                return;
            }
            break;
        }
        anc = tparent[anc];
        if (++iterations > R) {
            raise_error(err, 2);
            return;
        }
    }
    if (anc == -1 && parent_level != 0) {
        raise_error(err, 3);
        return;
    }
    const int first_node = 1 + poffset[node];
    Key* prefix = static_cast<Key*>(o.prefix);
    for (int local = 0; local < count; ++local) {
        const int nd = first_node + local;
        if (nd <= 0 || nd >= o.n_nodes) {
            raise_error(err, 6);
            return;
        }
        const int l = parent_level + local + 1;
        const int p = (local == 0) ? parent_node : nd - 1;
        o.level[nd] = l;
        prefix[nd] = prefix_at<BITS>(key, l);
        o.parent[nd] = p;
        const int oct = octant_at<BITS>(key, l);
        const int prev = atomicCAS(&o.children[static_cast<size_t>(p) * 8 + oct], -1, nd);
        if (prev != -1 && prev != nd) {
            raise_error(err, 4);
            return;
        }
        if (child_count)
            atomicAdd(&child_count[p], 1);
    }
    if (node < n_leaves) {
        const int leaf = first_node + count - 1;
        const int leaf_level = o.level[leaf];
        if (leaf_level <= 0 || leaf_level > L::bits) {
            raise_error(err, 5);
            return;
        }
        o.first_particle[leaf] = gfirst[node];
        o.particle_count[leaf] = gcount[node];
    }
}

// computeOctreeGeometryKernel (octree_geometry.cu:56-102).
template <int BITS>
__global__ __launch_bounds__(256) void geometry_kernel(nbh_octree_view o,
                                                       double rcx, double rcy,
                                                       double rcz, double side,
                                                       int* err)
{
    using Key = typename Layout<BITS>::Key;
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= o.n_nodes)
        return;
    double cx, cy, cz, h;
    int code = 0;
    if (!node_geometry<BITS>(static_cast<const Key*>(o.prefix)[node],
                             o.level[node], rcx, rcy, rcz, side, cx, cy, cz, h,
                             code)) {
        raise_error(err, code);
        return;
    }
    o.center_x[node] = cx;
    o.center_y[node] = cy;
    o.center_z[node] = cz;
    o.half_size[node] = h;
}

// countOctreeChildrenKernel (octree_physics.cu:63-92).
__global__ __launch_bounds__(256) void count_children_kernel(nbh_octree_view o,
                                                             int* err)
{
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= o.n_nodes)
        return;
    int c = 0;
    for (int oc = 0; oc < 8; ++oc) {
        const int ch = o.children[static_cast<size_t>(node) * 8 + oc];
        if (ch == -1)
            continue;
        if (ch <= 0 || ch >= o.n_nodes) {
            raise_error(err, 1);
            return;
        }
        ++c;
    }
    o.child_count[node] = c;
    if (o.particle_count[node] > 0 && c != 0)
        raise_error(err, 4);
}

// initializeOctreeLeavesKernel (octree_physics.cu:94-153).
__global__ __launch_bounds__(256) void leaf_init_kernel(
    nbh_octree_view o, const uint32_t* __restrict__ idx,
    const double* __restrict__ m, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, int* err)
{
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= o.n_nodes)
        return;
    const int count = o.particle_count[node];
    if (count <= 0)
        return;
    const int f = o.first_particle[node];
    if (f < 0 || f + count > o.n_particles) {
        raise_error(err, 2);
        return;
    }
    double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
    for (int k = 0; k < count; ++k) {
        const uint32_t p = idx[f + k];
        if (p >= static_cast<uint32_t>(o.n_particles)) {
            raise_error(err, 3);
            return;
        }
        const double pm = m[p];
        tm += pm;
        wx = fma(pm, x[p], wx);
        wy = fma(pm, y[p], wy);
        wz = fma(pm, z[p], wz);
    }
    o.mass[node] = tm;
    if (tm > 0.0) {
        const double inv = 1.0 / tm;
        o.com_x[node] = wx * inv;
        o.com_y[node] = wy * inv;
        o.com_z[node] = wz * inv;
    } else {
        o.com_x[node] = 0.0;
        o.com_y[node] = 0.0;
        o.com_z[node] = 0.0;
    }
}

// propagateOctreeMassKernel (octree_physics.cu:155-244).  The arrival
// counter is an agent-scope acquire/release RMW: the releasing side makes
// this child's mass/COM visible, the acquiring last arriver sees all
// siblings' results even when they were written on another XCD.
__global__ __launch_bounds__(256) void propagate_kernel(nbh_octree_view o,
                                                        int* err)
{
    const int leaf = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (leaf >= o.n_nodes)
        return;
    if (o.particle_count[leaf] <= 0)
        return;
    if (o.child_count[leaf] != 0) {
        raise_error(err, 4);
        return;
    }
    int p = o.parent[leaf];
    while (p != -1) {
        if (p < 0 || p >= o.n_nodes) {
            raise_error(err, 5);
            return;
        }
        const int required = o.child_count[p];
        if (required <= 0 || required > 8) {
            raise_error(err, 6);
            return;
        }
        const int done = fetch_add_acq_rel(&o.pending_children[p], 1) + 1;
        if (done < required)
            return;
        if (done > required) {
            raise_error(err, 7);
            return;
        }
        double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
        for (int oc = 0; oc < 8; ++oc) {
            const int ch = o.children[static_cast<size_t>(p) * 8 + oc];
            if (ch == -1)
                continue;
            const double cm = o.mass[ch];
            tm += cm;
            wx = fma(cm, o.com_x[ch], wx);
            wy = fma(cm, o.com_y[ch], wy);
            wz = fma(cm, o.com_z[ch], wz);
        }
        o.mass[p] = tm;
        if (tm > 0.0) {
            const double inv = 1.0 / tm;
            o.com_x[p] = wx * inv;
            o.com_y[p] = wy * inv;
            o.com_z[p] = wz * inv;
        } else {
            o.com_x[p] = 0.0;
            o.com_y[p] = 0.0;
            o.com_z[p] = 0.0;
        }
        p = o.parent[p];
    }
}

} // namespace

template <int BITS>
void launch_octree_build(hipStream_t s, const nbh_octree_view& o,
                         int n_leaves, const int* tparent,
                         const int* range_first,
                         const typename Layout<BITS>::Key* unique,
                         const int* gfirst, const int* gcount,
                         const int* plevel, const int* pcount,
                         const int* poffset, int* err, bool count_children)
{
    unsigned blocks = blocks_for(o.n_nodes, 256);
    if (blocks > 8192)
        blocks = 8192;
    octree_init_kernel<BITS><<<blocks, 256, 0, s>>>(
        o, count_children ? o.child_count : nullptr,
        count_children ? o.pending_children : nullptr);
    NBH_HIP(hipGetLastError());
    const int R = 2 * n_leaves - 1;
    materialize_kernel<BITS><<<blocks_for(R, 256), 256, 0, s>>>(
        o, n_leaves, tparent, range_first, unique, gfirst, gcount, plevel,
        pcount, poffset, err, count_children ? o.child_count : nullptr);
    NBH_HIP(hipGetLastError());
}

template <int BITS>
void launch_geometry(hipStream_t s, const nbh_octree_view& o, double cx,
                     double cy, double cz, double side, int* err)
{
    geometry_kernel<BITS><<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(
        o, cx, cy, cz, side, err);
    NBH_HIP(hipGetLastError());
}

void launch_count_children(hipStream_t s, const nbh_octree_view& o, int* err)
{
    count_children_kernel<<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(o, err);
    NBH_HIP(hipGetLastError());
}

void launch_leaf_init(hipStream_t s, const nbh_octree_view& o,
                      const uint32_t* idx, const double* m, const double* x,
                      const double* y, const double* z, int* err)
{
    leaf_init_kernel<<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(o, idx, m, x,
                                                                y, z, err);
    NBH_HIP(hipGetLastError());
}

void launch_propagate(hipStream_t s, const nbh_octree_view& o, int* err)
{
    propagate_kernel<<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(o, err);
    NBH_HIP(hipGetLastError());
}

#define NBH_OCTREE_INST(B)                                                          \
    template void launch_octree_build<B>(hipStream_t, const nbh_octree_view&, int,   \
                                         const int*, const int*,                     \
                                         const typename Layout<B>::Key*, const int*, \
                                         const int*, const int*, const int*,         \
                                         const int*, int*, bool);                    \
    template void launch_geometry<B>(hipStream_t, const nbh_octree_view&, double,    \
                                     double, double, double, int*);
NBH_OCTREE_INST(10)
NBH_OCTREE_INST(21)
#undef NBH_OCTREE_INST

} // namespace nbh
