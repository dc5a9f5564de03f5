// nbh_internal.hpp -- shared host/device definitions of the MI355X pipeline.
//
// gfx950 only: wave64, 160 KB LDS per CU, 8 XCDs.  Every kernel in this
// library is hand-written for that target; there is no other code path.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

#include "nbody_amd.h"

namespace nbh {

// ------------------------------------------------------------- errors
// Host-side failure carrying the ABI status, the device error code (0 when
// the failure was detected on the host) and the message the reference's
// exception would carry.
struct Error {
    int status;
    int code;
    std::string message;
};

[[noreturn]] void fail(int status, const std::string& message, int code = 0);
void hip_check(hipError_t e, const char* what);
#define NBH_HIP(call) ::nbh::hip_check((call), #call)
void set_last_error(int code, const std::string& message);

// Runs f, mapping every failure to an ABI status and the thread's last
// error message / device code.
template <typename F>
int guard(F&& f)
{
    set_last_error(0, "");
    try {
        f();
        return NBH_OK;
    } catch (const Error& e) {
        set_last_error(e.code, e.message);
        return e.status;
    } catch (const std::bad_alloc&) {
        set_last_error(0, "host allocation failed");
        return NBH_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        set_last_error(0, e.what());
        return NBH_RUNTIME_ERROR;
    }
}

// --------------------------------------------------------- key layouts
// Reference layout: 10 bits per axis, keys in uint32_t with 2 padding bits
// (src/radix_to_octree.hpp:10-19).  Extended layout: 21 bits per axis, keys
// in uint64_t with 1 padding bit.
template <int BITS>
struct Layout;

template <>
struct Layout<10> {
    using Key = uint32_t;
    static constexpr int bits = 10;
    static constexpr int spatial_bits = 30;
    static constexpr int storage_bits = 32;
    static constexpr int padding_bits = 2;
    static constexpr Key spatial_mask = 0x3FFFFFFFu;
    static constexpr int sort_passes = 4; // 8-bit digits over 30 bits
};

template <>
struct Layout<21> {
    using Key = uint64_t;
    static constexpr int bits = 21;
    static constexpr int spatial_bits = 63;
    static constexpr int storage_bits = 64;
    static constexpr int padding_bits = 1;
    static constexpr Key spatial_mask = 0x7FFFFFFFFFFFFFFFull;
    static constexpr int sort_passes = 8; // 8-bit digits over 63 bits
};

// ------------------------------------------------------------- workspace
// Grow-only device scratch owned by a context.  Each slot is reused by the
// stage that names it; growth reallocates (rare after the first step).
enum Slot : int {
    SLOT_SORT_KEYS_ALT = 0,
    SLOT_SORT_VALS_ALT,
    SLOT_SORT_STATUS,
    SLOT_SORT_HIST,
    SLOT_SCAN_STATUS,
    SLOT_ERRORS,
    SLOT_MISC,
    SLOT_TMP0,
    SLOT_TMP1,
    SLOT_TMP2,
    SLOT_TMP3,
    SLOT_TMP4,
    SLOT_TMP5,
    SLOT_TMP6,
    SLOT_TMP7,
    SLOT_COUNT
};

struct Workspace {
    void* ptr[SLOT_COUNT] = {};
    size_t bytes[SLOT_COUNT] = {};
    void* get(int slot, size_t need);
    void release();
};

// Error words: one int per stage, cleared by the stage that owns it.
enum ErrorWord : int {
    ERR_BBOX = 0,   // 1 = non-finite coordinate
    ERR_PLAN,       // radix_to_octree.cu:32-51 codes 1-4
    ERR_OCTREE,     // octree_builder.cu:18-46 codes 1-7
    ERR_GEOMETRY,   // octree_geometry.cu:21-34 codes 1-2
    ERR_PHYSICS,    // octree_physics.cu:18-61 codes 1-7
    ERR_TRAVERSAL,  // barnes_hut.cu:21-52 codes 1-8
    ERR_GROUPS,     // group count sanity
    ERR_WORD_COUNT
};

} // namespace nbh

struct nbh_context {
    hipStream_t stream = nullptr;
    nbh::Workspace ws;
    int* host_pinned = nullptr; // 64 ints of pinned readback space
};

namespace nbh {

// Status words in device memory for each context (SLOT_ERRORS).
int* error_words(nbh_context* ctx);
// Copy the error words (and up to `extra` ints from `d_extra`) to the host
// after a stream synchronisation.
void sync_read(nbh_context* ctx, const int* d_src, int* h_dst, int count);

// ------------------------------------------------------- launch helpers
constexpr int kWave = 64;

inline unsigned blocks_for(size_t n, unsigned threads)
{
    const size_t b = (n + threads - 1) / threads;
    if (b > 0x7FFFFFFFull)
        fail(NBH_OVERFLOW_ERROR, "kernel grid exceeds the launch limit");
    return static_cast<unsigned>(b == 0 ? 1 : b);
}

// ---------------------------------------------------- stage entry points
// Implemented in the .hip files and shared by the C ABI and the simulation.

// common.hip
void launch_fill_u32(hipStream_t s, uint32_t* p, uint32_t v, size_t n);
void launch_fill_i32(hipStream_t s, int* p, int v, size_t n);
void launch_fill_f64(hipStream_t s, double* p, double v, size_t n);
void launch_fill_u64(hipStream_t s, unsigned long long* p,
                     unsigned long long v, size_t n);

// bbox_morton.hip
size_t bbox_partial_doubles();
// d_box receives {centre x, y, z, side}.  With m and pos4 given, both also
// write the packed {x, y, z, m} copy of every particle (original order) that
// the Morton gather reads, and d_box[4] = the largest |m| (1 when every mass
// is zero), the mass unit of the mixed traversal's records.
void launch_bbox(nbh_context* ctx, const double* x, const double* y,
                 const double* z, size_t n, double* d_box, double* partials,
                 int* err, const double* m = nullptr, double4* pos4 = nullptr);
void launch_kick_drift_bbox(hipStream_t s, double* x, double* y, double* z,
                            double* vx, double* vy, double* vz,
                            const double* ax, const double* ay,
                            const double* az, size_t n, double dt,
                            double* d_box, double* partials, int* err,
                            const double* m = nullptr, double4* pos4 = nullptr);
template <int BITS>
void launch_encode(hipStream_t s, const double* x, const double* y,
                   const double* z, size_t n, const double* d_box,
                   typename Layout<BITS>::Key* keys, uint32_t* idx,
                   uint32_t* hist /* passes*256, zeroed */);
// pos4 (optional): packed {x, y, z, m} in original order, read instead of
// the four planar arrays.
void launch_gather(hipStream_t s, const uint32_t* idx, const double* m,
                   const double* x, const double* y, const double* z,
                   size_t n, double4* rec64, int4* rec32, const double* box,
                   const double4* pos4 = nullptr);

// radix_sort.hip
size_t sort_status_words(size_t n, int passes);
template <typename K>
void launch_histogram(hipStream_t s, const K* keys, size_t n, int passes,
                      uint32_t* hist);
template <typename K>
void launch_radix_passes(hipStream_t s, K* keys, uint32_t* vals, K* keys_alt,
                         uint32_t* vals_alt, size_t n, int passes,
                         uint32_t* hist, uint32_t* status);

// Octree levels the pipeline's level passes handle (63-bit keys give 22),
// and the spacing of the per-level internal-node counters (one 128-byte
// line each, so that tiles adding to different levels do not share a line).
constexpr int kMaxLevels = 32;
constexpr int kLevelCountStride = 32;

// scan.hip
size_t scan_status_words(size_t n);
template <typename K>
void launch_leaf_groups(hipStream_t s, const K* keys, int n, K* unique,
                        int* first, int* count, int* d_ngroups,
                        unsigned long long* status);
template <int BITS>
void launch_plan_scan(hipStream_t s, const int* d_nleaves, int max_leaves,
                      const int* tparent, const int* prefix_len,
                      const typename Layout<BITS>::Key* unique, int* level,
                      int* count, int* offset, int* d_noctree, int* err,
                      unsigned long long* status, int* level_count = nullptr,
                      const int* parent_prefix_len = nullptr);

// karras.hip
template <typename K>
void launch_karras(hipStream_t s, const K* keys, const int* d_nleaves,
                   int max_leaves, int* left, int* right, int* parent,
                   int* range_first, int* range_last, int* prefix_len,
                   int* parent_prefix_len = nullptr);
void launch_tree_com(hipStream_t s, const nbh_tree_view& t, const int* first,
                     const int* count, const uint32_t* idx, const double* m,
                     const double* x, const double* y, const double* z);

// octree.hip
template <int BITS>
void launch_octree_build(hipStream_t s, const nbh_octree_view& o,
                         int n_leaves, const int* tparent,
                         const int* range_first,
                         const typename Layout<BITS>::Key* unique,
                         const int* gfirst, const int* gcount,
                         const int* plevel, const int* pcount,
                         const int* poffset, int* err, bool count_children);
template <int BITS>
void launch_geometry(hipStream_t s, const nbh_octree_view& o, double cx,
                     double cy, double cz, double side, int* err);
void launch_count_children(hipStream_t s, const nbh_octree_view& o, int* err);
void launch_leaf_init(hipStream_t s, const nbh_octree_view& o,
                      const uint32_t* idx, const double* m, const double* x,
                      const double* y, const double* z, int* err);
void launch_propagate(hipStream_t s, const nbh_octree_view& o, int* err);

// One position of the FP64 walk, 64 bytes, fetched by one scalar load:
// the node's centre of mass and mass (a single particle's position and mass
// for a one-particle leaf), the opening threshold on d^2, the link and G m.
//   thr: the traversal takes the record at d^2 > thr.  Internal and
//        childless nodes hold mac_threshold(crit) (exactly the reference's
//        sqrt(d^2) > crit); one-particle leaves and zero-mass nodes hold
//        -inf (always taken as a monopole: the leaf's exact particle term,
//        skipping the target's own particle by index, or nothing for a zero
//        mass); leaves of several particles and error records hold +inf.
//   kind: KIND_INTERNAL for internal nodes and the two always-taken kinds,
//        else the record's kind (a childless node that a lane does not take
//        is error 5 for that lane).
//   first, count: a leaf's particle range (count 1 for a one-particle leaf;
//        -1, 0 for nodes).
//   gm: G m, rounded once, as the reference's G * m * r^-3 rounds it.
struct alignas(64) Walk64Rec {
    double x, y, z, m;
    double thr;
    int skip, kind, first, count;
    double gm;
};
static_assert(sizeof(Walk64Rec) == 64, "FP64 walk records are 64 bytes");

// Per-node state of the pipeline's level passes (one entry per octree
// node in each array; sum4, aux and crit for internal nodes only: a leaf's
// values follow from its particles):
//   sum4  {com x, y, z, mass}        written bottom-up by the node
//   aux   {size, hcount, 0, c}       bottom-up by the node: walk positions
//                                    of the subtree, its particle count, and
//                                    c = 1 when the mixed layout collapsed
//                                    it into its single child
//   crit  opening radius             bottom-up by the node
//   link  {local, hoffset}           bottom-up by the parent: position and
//                                    Hilbert-curve offsets inside the
//                                    parent's subtree
//   pos   {pre, hoffset}             top-down by the node: absolute
//                                    pre-order position and curve offset
// A parent reads each child's sum4 and aux; the top-down passes read the
// node's own words and the parent's pos.
struct LevelArrays {
    double4* sum4;
    int4* aux;
    double* crit;
    int2* link;
    int2* pos;
};

// walk_layout.hip, walk_fp64.hip, walk_mixed.hip
void launch_generic_sizes(hipStream_t s, const nbh_octree_view& o, int* size,
                          int* local, int* err);
// The caller-octree walk build (FP64 only): w64 holds n_nodes + 1 records.
void launch_walk_build(hipStream_t s, const nbh_octree_view& o, const int* size,
                       const int* local, double theta, const double4* rec64,
                       bool need_center, double eps, double G, Walk64Rec* w64,
                       double4* w64center, int* wcount);
int walk_blocks_padded(int n);
// queue: kWalkQueueInts ints of work-queue counters (cleared by the launch).
constexpr int kWalkQueueInts = 256;
// The FP64 traversal: w64 records (wcap of them), rec the Morton-order
// {x, y, z, m} particles, order the sorted target indices in packet order
// (null: Morton order).  Accelerations go to the targets' original indices.
// side: the side of a cube holding every particle (0 when unknown), which
// selects the walk without range tests when eps2 and side allow it.
void launch_walk_fp64(hipStream_t s, const Walk64Rec* w64, size_t wcap,
                      const double4* wcenter, const int* d_wcount,
                      const double4* rec, const uint32_t* idx,
                      const uint32_t* order, int n, double eps2, double G,
                      bool contains, double side, double* ax, double* ay, double* az,
                      int* queue, int* err, unsigned long long* counters);
// wpairs: interleaved {node, link} int4 pairs, 32 bytes per position.
// Split traversal (few packets): at most kWalkSplitMax subtrees per packet,
// for walks of at most kWalkSplitPackets packets.  split: 1 + 4 *
// kWalkSplitMax ints; part: kWalkSplitMax * 64 * 3 doubles per packet.
constexpr int kWalkSplitMax = 8;
constexpr int kWalkSplitPackets = 16384; // N <= 1,048,576
constexpr int kWalkSplitInts = 1 + 4 * kWalkSplitMax;
// Whether the mixed traversal of n targets is split (a function of n only,
// so that results do not depend on the device).
bool walk_mixed_splits(int n);
// Marks the subtree starts of a split mixed walk below its first record and
// fills the split table (kWalkSplitInts ints); a layout that does not split
// gets a one-unit table covering the whole walk.
void launch_walk_split(hipStream_t s, int4* wpairs, const int* d_wcount, int* split);
// wcap: records the wpairs buffer holds.
// order: sorted target indices in the packets' order (null: Morton order).
// wpairs: the {node, link} record pairs; a split walk marks its subtrees'
// first records in place (see walk_split_kernel).
// perm, pcost (one entry per unit each: packets, or packets times
// kWalkSplitMax for a split walk): the walk hands out its units in the
// order built from pcost (packet_schedule_kernel) and, unless it counts
// interactions, records their new visit counts in pcost.  Zero costs give
// the curve order.
void launch_walk_mixed(hipStream_t s, int4* wpairs, size_t wcap,
                       const int4* wcenter, const int* d_wcount,
                       const int4* rec, const uint32_t* idx,
                       const uint32_t* order, int n,
                       float eps2, double out_scale, bool contains, bool soft,
                       double* ax, double* ay, double* az, int* queue,
                       int* split, double* part, uint32_t* perm, uint32_t* pcost,
                       int* err, unsigned long long* counters);
void launch_direct(hipStream_t s, const double* m, const double* x,
                   const double* y, const double* z, size_t n,
                   const uint32_t* targets, size_t n_targets, double eps2,
                   double G, double* ax, double* ay, double* az);

// integrate.hip
void launch_kick_drift(hipStream_t s, double* x, double* y, double* z,
                       double* vx, double* vy, double* vz, const double* ax,
                       const double* ay, const double* az, size_t n, double dt);
void launch_kick(hipStream_t s, double* vx, double* vy, double* vz,
                 const double* ax, const double* ay, const double* az,
                 size_t n, double dt);
void launch_gather4(hipStream_t s, const uint32_t* idx, const double* m,
                    const double* x, const double* y, const double* z,
                    double* om, double* ox, double* oy, double* oz, size_t n);
void launch_pack_records(hipStream_t s, const uint32_t* idx, const double* m,
                         const double* x, const double* y, const double* z,
                         size_t n, bool morton, double4* rec, int* err);

// The opening test needs cell geometry only from this theta on.
constexpr double kContainsTheta = 1.15;

// eps^2 of the mixed traversal in grid units (scale = 2^31 / side), with a
// floor of 1e-20 under positive softening.  Separations of distinct grid
// points are at least 1, so the floor changes no interaction's FP32 r^2,
// while a coincident pair (the target's own record included) keeps a finite
// inv^3 and so adds exactly zero.
__host__ __device__ inline double mixed_eps2(double eps, double scale)
{
    const double eq = eps * scale;
    return eps > 0.0 ? fmax(eq * eq, 1e-20) : 0.0;
}

} // namespace nbh
