// walk_layout.hip -- the derived depth-first layout the traversals walk.
//
// Layout.  The reference walks the octree per target with parent climbs and
// 8-slot sibling scans, visiting octants 7 -> 0 (barnes_hut.cu:90-162).  Here
// the same visiting order is laid out once per build: nodes are stored at
// their depth-first pre-order position (children 7 -> 0), each record
// carrying `skip` = pre + subtree size, the position that follows its
// subtree.  "Open" is then pre + 1 and "done with this subtree" is skip: no
// stack, no parent reads.
//
// Packets.  One wavefront owns 64 targets (Morton-consecutive in the FP64
// walk, consecutive along a Hilbert curve in the mixed one).  The wave
// walks the layout in increasing position.  A lane is active at position i
// when its own next position is <= i; active lanes evaluate the node exactly
// as the per-target walk would (leaf -> exact particles, zero mass -> skip,
// else MAC: accept -> monopole, reject -> open).  If any lane opens, the
// wave steps to i + 1, else to skip[i].  Every lane therefore sees precisely
// the node sequence of its own per-target walk, in the same order, while node
// records are fetched once per wave with scalar loads.
//
// Opening test.  The reference accepts a node when the target is outside the
// cell, d^2 > 0, theta > 0 and d > s/theta + delta (barnes_hut.cu:273-294).
// For theta < 2/sqrt(3) the containment test never decides: inside the cell
// d <= delta + sqrt(3) h < 2h/theta + delta.  The kernels therefore load
// cell geometry only when theta >= 1.15.
//
// Precision.  FP64 reproduces the reference arithmetic (and is bit-exact
// with the oracle).  Mixed evaluates the opening test and the interaction in
// FP32 on exact integer differences of 31-bit fixed-point coordinates over
// the bounding box -- coordinates relative to the target, so the rounding of
// each separation is relative to the separation itself -- and sums each
// target's acceleration with Kahan compensation.
#include "octree_device.hpp"
#include "walk_common.hpp"
#include "hilbert.hpp"

namespace nbh {

namespace {

// ------------------------------------------------- generic tree sizes
// For octrees handed in through the C ABI: child ids must be valid (else
// code 1, barnes_hut.cu:101-105) and every child must name its parent (else
// code 4, barnes_hut.cu:150-154).  Writes the child counts to the scratch
// array o.child_count.
__global__ __launch_bounds__(256) void check_parents_kernel(nbh_octree_view o,
                                                            int* err)
{
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= o.n_nodes)
        return;
    int count = 0;
    for (int oc = 0; oc < 8; ++oc) {
        const int c = o.children[static_cast<size_t>(node) * 8 + oc];
        if (c == -1)
            continue;
        if (c <= 0 || c >= o.n_nodes) {
            raise_error(err, 1);
            return;
        }
        if (o.parent[c] != node)
            raise_error(err, 4);
        ++count;
    }
    o.child_count[node] = count;
}

// Subtree sizes and pre-order offsets, started from every childless node.
__global__ __launch_bounds__(256) void generic_sizes_kernel(
    nbh_octree_view o, int* __restrict__ size, int* __restrict__ local,
    int* err)
{
    const int node = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (node >= o.n_nodes)
        return;
    if (o.child_count[node] != 0)
        return;
    size[node] = 1;
    int v = node;
    int guard = 0;
    while (v != 0) {
        const int p = o.parent[v];
        if (p < 0 || p >= o.n_nodes || ++guard > o.n_nodes) {
            // not connected to the root: never visited by the traversal
            return;
        }
        const int required = o.child_count[p];
        const int done = fetch_add_acq_rel(&o.pending_children[p], 1) + 1;
        if (done < required)
            return;
        if (done > required) {
            raise_error(err, 4);
            return;
        }
        int off = 1;
        for (int oc = 7; oc >= 0; --oc) {
            const int c = o.children[static_cast<size_t>(p) * 8 + oc];
            if (c <= 0 || c >= o.n_nodes)
                continue;
            local[c] = off;
            off += size[c];
        }
        size[p] = off;
        v = p;
    }
    local[0] = 0;
}

// Pre-order position of a node: the sum of the sibling offsets along its
// path to the root.  Returns -1 when the path does not reach node 0.
__device__ __forceinline__ int preorder(const nbh_octree_view& o,
                                        const int* __restrict__ local, int v)
{
    int pre = 0;
    int guard = 0;
    while (v != 0) {
        pre += local[v];
        v = o.parent[v];
        if (v < 0 || ++guard > 64)
            return -1;
    }
    return pre;
}

// --------------------------------------------------------- walk records
// Outputs of the walk build (records indexed by pre-order position).
// The mixed traversal's target order along the Hilbert curve (see
// hilbert.cpp), built alongside the level passes; all pointers are null
// when the order is not wanted.  A node's curve offset is the rank of its
// subtree's first particle along the curve: the bottom-up pass writes each
// child's offset relative to its parent, the top-down pass adds the
// parent's absolute offset (as for pre-order positions).
struct HilbertOut {
    int* count;          // subtree particle count per node (bottom-up)
    int* offset;         // curve offset, relative to the parent until made absolute
    uint32_t* order;     // curve rank -> sorted particle index
    const uint8_t* tab;  // HilbertTables: rank[24][8] then next[24][8]
};

constexpr int kHilbertTabBytes = 2 * 8 * kHilbertStates;

// Curve orientation state of the level-L cell with (left-aligned) Morton
// prefix p: the root's state 0 followed down the prefix's octant digits.
template <int BITS>
__device__ __forceinline__ int curve_state(typename Layout<BITS>::Key p, int L,
                                           const uint8_t* next)
{
    int st = 0;
    for (int d = 1; d <= L; ++d) {
        const int oct = static_cast<int>((p >> (Layout<BITS>::spatial_bits - 3 * d)) & 0x7u);
        st = next[8 * st + oct];
    }
    return st;
}

struct WalkOut {
    Walk64Rec* w64;
    double4* w64center;
    int4* w32node;
    int4* w32link;
    int4* w32center;
    double G; // FP64 records hold G m
};

// The record of node v at pre-order position `pre`.  GENERIC: a
// caller-supplied octree, with the reference's per-node failure conditions
// encoded as error kinds and the opening radius computed here.  Otherwise:
// the pipeline's own octree with its precomputed opening radius, packed
// {com, mass} in sum4[v] and the cell geometry {centre, half} in `geom`.
template <bool GENERIC, bool MIXED>
__device__ __forceinline__ void write_record(
    const nbh_octree_view& o, int v, int pre, const int* __restrict__ size,
    const double* __restrict__ crit_in, double theta,
    const double4* __restrict__ rec64, const int4* __restrict__ rec32,
    const double* __restrict__ box, bool need_center, double eps,
    const WalkOut& w, const double4* __restrict__ sum4, double4 geom)
{
    double4* w64center = w.w64center;
    int4* w32node = w.w32node;
    int4* w32link = w.w32link;
    int4* w32center = w.w32center;
    const int skip = pre + size[v];
    const int pc = o.particle_count[v];
    const double h = GENERIC ? o.half_size[v] : geom.w;
    double4 mc;
    if (GENERIC)
        mc = make_double4(o.com_x[v], o.com_y[v], o.com_z[v], o.mass[v]);
    else
        mc = sum4[v];
    int kind;
    int first = 0, count = 0;
    if (GENERIC) {
        const int nch = o.child_count[v];
        if (!(h > 0.0)) {
            kind = KIND_ERROR + 8;
        } else if (pc > 0) {
            first = o.first_particle[v];
            count = pc;
            if (nch != 0)
                kind = KIND_ERROR + 7;
            else if (first < 0 || pc > o.n_particles ||
                     first > o.n_particles - pc)
                kind = KIND_ERROR + 2;
            else
                kind = KIND_LEAF;
        } else if (o.mass[v] == 0.0) {
            kind = KIND_SKIP;
        } else if (nch == 0) {
            kind = KIND_CHILDLESS;
        } else {
            kind = KIND_INTERNAL;
        }
    } else {
        if (pc > 0) {
            kind = KIND_LEAF;
            first = o.first_particle[v];
            count = pc;
        } else if (mc.w == 0.0) {
            kind = KIND_SKIP;
        } else {
            kind = KIND_INTERNAL;
        }
    }
    const double m = mc.w;
    const double cx = mc.x, cy = mc.y, cz = mc.z;
    double crit = 0.0;
    if (kind == KIND_INTERNAL || kind == KIND_CHILDLESS) {
        if (GENERIC) {
            const double ox = cx - o.center_x[v];
            const double oy = cy - o.center_y[v];
            const double oz = cz - o.center_z[v];
            const double o2 = fma(oz, oz, fma(ox, ox, oy * oy));
            crit = (2.0 * h) / theta + sqrt(o2);
        } else {
            crit = crit_in[v];
        }
    }
    if (!MIXED) {
        // One-particle leaves and zero-mass nodes are always-taken
        // monopoles (thr = -inf): a leaf's exact particle term (the walk
        // skips the target's own particle by index) or nothing for a zero
        // mass; the cell test of large theta never applies to them (half -1).
        Walk64Rec r;
        const bool fast = (kind == KIND_LEAF && count == 1) || kind == KIND_SKIP;
        if (kind == KIND_LEAF && count == 1) {
            const double4 p = rec64[first];
            r.x = p.x;
            r.y = p.y;
            r.z = p.z;
            r.m = p.w;
        } else {
            r.x = cx;
            r.y = cy;
            r.z = cz;
            r.m = m;
        }
        r.skip = skip;
        r.first = (kind == KIND_LEAF) ? first : -1;
        r.count = (kind == KIND_LEAF) ? count : 0;
        r.gm = w.G * r.m;
        if (fast) {
            r.kind = KIND_INTERNAL;
            r.thr = -INFINITY;
        } else if (kind == KIND_INTERNAL || kind == KIND_CHILDLESS) {
            r.kind = kind;
            r.thr = mac_threshold(crit);
        } else {
            r.kind = kind;
            r.thr = INFINITY;
        }
        w.w64[pre] = r;
        if (need_center) {
            double4 c = GENERIC ? make_double4(o.center_x[v], o.center_y[v], o.center_z[v], h)
                                : geom;
            if (fast)
                c.w = -1.0;
            w64center[pre] = c;
        }
    } else {
        const double side = box[3];
        const double scale = 2147483648.0 / side;
        const double ox = box[0] - side / 2.0;
        const double oy = box[1] - side / 2.0;
        const double oz = box[2] - side / 2.0;
        const double lim = 2147483647.0;
        if (kind == KIND_LEAF && count == 1) {
            w32node[2 * static_cast<size_t>(pre)] = rec32[first];
        } else {
            const int qx = static_cast<int>(fmin(fmax(rint((cx - ox) * scale), 0.0), lim));
            const int qy = static_cast<int>(fmin(fmax(rint((cy - oy) * scale), 0.0), lim));
            const int qz = static_cast<int>(fmin(fmax(rint((cz - oz) * scale), 0.0), lim));
            const float gm = static_cast<float>(m / box[4]);
            w32node[2 * static_cast<size_t>(pre)] = make_int4(qx, qy, qz, __float_as_int(gm));
        }
        // The traversal compares the softened squared distance d^2 + eps^2
        // (its interaction's r^2) with crit^2 + eps^2, both in fixed-point
        // units.
        const double cq = crit * scale;
        const double eq2 = mixed_eps2(eps, scale);
        float crit2 = static_cast<float>(fma(cq, cq, eq2));
        // With positive softening a single-particle leaf and a zero-mass
        // node are "always accept" monopoles (crit^2 = -1): the target's own
        // particle and a zero mass both add exactly zero.  Records of every
        // other non-internal kind carry crit^2 = +inf, so that the
        // traversal's MAC never accepts them.
        const bool soft = eps > 0.0;
        const bool fast = soft && ((kind == KIND_LEAF && count == 1) || kind == KIND_SKIP);
        int kind_out = fast ? KIND_INTERNAL : kind;
        // The link word holds the leaf's particle count above its kind byte,
        // 23 bits: a larger leaf (that many particles on one grid key)
        // becomes an error record, which every walk reaching it reports.
        if (kind == KIND_LEAF && count >= kMixedLeafCountLimit) {
            kind_out = KIND_ERROR + kErrMixedLeafTooLarge;
            count = 0;
        }
        if (fast)
            crit2 = -1.0f;
        else if (kind_out != KIND_INTERNAL)
            crit2 = __int_as_float(0x7F800000);
        w32link[2 * static_cast<size_t>(pre)] = make_int4(__float_as_int(crit2), skip,
                                 kind_out | (count << 8), first);
        if (need_center) {
            const double gx = GENERIC ? o.center_x[v] : geom.x;
            const double gy = GENERIC ? o.center_y[v] : geom.y;
            const double gz = GENERIC ? o.center_z[v] : geom.z;
            const int qx = static_cast<int>(rint((gx - ox) * scale));
            const int qy = static_cast<int>(rint((gy - oy) * scale));
            const int qz = static_cast<int>(rint((gz - oz) * scale));
            const int qh = fast ? -1 : static_cast<int>(fmin(rint(h * scale), lim));
            w32center[pre] = make_int4(qx, qy, qz, qh);
        }
    }
}

template <bool GENERIC, bool MIXED>
__global__ __launch_bounds__(256) void walk_build_kernel(
    nbh_octree_view o, const int* __restrict__ size,
    const int* __restrict__ local, const double* __restrict__ crit_in,
    double theta, const double4* __restrict__ rec64,
    const int4* __restrict__ rec32, const double* __restrict__ box,
    bool need_center, double eps, WalkOut w, int* __restrict__ wcount)
{
    const int v = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= o.n_nodes)
        return;
    const int pre = preorder(o, local, v);
    if (pre < 0 || pre >= o.n_nodes)
        return;
    if (v == 0)
        *wcount = size[0] < o.n_nodes ? size[0] : o.n_nodes;
    write_record<GENERIC, MIXED>(o, v, pre, size, crit_in, theta, rec64, rec32,
                                 box, need_center, eps, w, nullptr,
                                 make_double4(0.0, 0.0, 0.0, 0.0));
}

// ------------------------------------------ level-synchronous passes
// Bottom-up over one octree level: every child of a level-L node sits at
// level L+1 and was completed by the previous launch, so no arrival
// counters (and no cache write-back/invalidate per node) are needed.  A
// node's values are the reference's: leaf mass/COM summed over its
// particles in sorted order (octree_physics.cu:94-153), internal nodes
// combined in octant order 0..7 (octree_physics.cu:211-243), cell geometry
// from the prefix path (octree_geometry.cu:56-102), the opening radius
// side/theta + |com - centre| (barnes_hut.cu:277-291); plus the subtree size
// and the pre-order offsets of the children (octants 7 -> 0).
template <int BITS>
__global__ __launch_bounds__(256) void level_up_kernel(
    nbh_octree_view o, int L, const double4* __restrict__ rec, double theta,
    const double* __restrict__ box, int* __restrict__ size,
    int* __restrict__ local, double* __restrict__ crit, bool collapse,
    const int* __restrict__ order, const int* __restrict__ start,
    double4* __restrict__ sum4, HilbertOut h)
{
    using Key = typename Layout<BITS>::Key;
    __shared__ uint8_t s_tab[kHilbertTabBytes];
    if (h.count)
        for (int t = static_cast<int>(threadIdx.x); t < kHilbertTabBytes; t += blockDim.x)
            s_tab[t] = h.tab[t];
    __syncthreads();
    const double rcx = box[0], rcy = box[1], rcz = box[2], side = box[3];
    const Key* prefix = static_cast<const Key*>(o.prefix);
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    const int end = start[L + 1];
    for (int k = start[L] + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
         k < end; k += stride) {
        const int v = order[k];
        const Key pv = prefix[v];
        double gx, gy, gz, gh;
        int code = 0;
        node_geometry<BITS>(pv, L, rcx, rcy, rcz, side, gx, gy, gz, gh, code);
        double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
        const int pc = o.particle_count[v];
        int nch = 0, only = 0;
        if (pc > 0) {
            const int f = o.first_particle[v];
            for (int k = 0; k < pc; ++k) {
                const double4 p = rec[f + k];
                tm += p.w;
                wx = fma(p.w, p.x, wx);
                wy = fma(p.w, p.y, wy);
                wz = fma(p.w, p.z, wz);
            }
            size[v] = 1;
            if (h.count)
                h.count[v] = pc;
        } else {
            const int4 c0 = reinterpret_cast<const int4*>(o.children)[2 * static_cast<size_t>(v)];
            const int4 c1 = reinterpret_cast<const int4*>(o.children)[2 * static_cast<size_t>(v) + 1];
            const int ch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
#pragma unroll
            for (int oc = 0; oc < 8; ++oc) {
                if (ch[oc] == -1)
                    continue;
                const double4 c = sum4[ch[oc]];
                tm += c.w;
                wx = fma(c.w, c.x, wx);
                wy = fma(c.w, c.y, wy);
                wz = fma(c.w, c.z, wz);
                only = ch[oc];
                ++nch;
            }
            // A collapsed single-child node takes no walk position of its
            // own: its child starts at the node's position and overwrites
            // the node's record on the next top-down level.  Only internal
            // children and single-particle leaves qualify (a leaf holding
            // several particles is summed exactly, never as a monopole).
            // Curve offsets of the children: in v's orientation state they
            // take consecutive particle ranges in the order of their ranks.
            if (h.count) {
                int cnt[8];
                int hc = 0;
#pragma unroll
                for (int oc = 0; oc < 8; ++oc) {
                    cnt[oc] = ch[oc] != -1 ? h.count[ch[oc]] : 0;
                    hc += cnt[oc];
                }
                h.count[v] = hc;
                const uint8_t* rank =
                    s_tab + 8 * curve_state<BITS>(pv, L, s_tab + 8 * kHilbertStates);
#pragma unroll
                for (int oc = 0; oc < 8; ++oc) {
                    if (ch[oc] == -1)
                        continue;
                    int before = 0;
#pragma unroll
                    for (int b = 0; b < 8; ++b)
                        before += (rank[b] < rank[oc]) ? cnt[b] : 0;
                    h.offset[ch[oc]] = before;
                }
            }
            int off = (collapse && nch == 1 && o.particle_count[only] <= 1) ? 0 : 1;
#pragma unroll
            for (int oc = 7; oc >= 0; --oc) {
                if (ch[oc] == -1)
                    continue;
                local[ch[oc]] = off;
                off += size[ch[oc]];
            }
            size[v] = off;
        }
        double cx = 0.0, cy = 0.0, cz = 0.0;
        if (tm > 0.0) {
            const double inv = 1.0 / tm;
            cx = wx * inv;
            cy = wy * inv;
            cz = wz * inv;
        }
        sum4[v] = make_double4(cx, cy, cz, tm);
        if (pc <= 0) {
            const double ox = cx - gx;
            const double oy = cy - gy;
            const double oz = cz - gz;
            const double o2 = fma(oz, oz, fma(ox, ox, oy * oy));
            crit[v] = (2.0 * gh) / theta + sqrt(o2);
        }
        if (v == 0)
            local[0] = 0;
    }
}

// Top-down over one level: pre-order position = parent's + own offset, then
// the node's walk record.
template <int BITS, bool MIXED>
__global__ __launch_bounds__(256) void level_down_kernel(
    nbh_octree_view o, int L, const int* __restrict__ size,
    const int* __restrict__ local, int* __restrict__ pre_of,
    const double* __restrict__ crit, double theta,
    const double4* __restrict__ rec64, const int4* __restrict__ rec32,
    const double* __restrict__ box, bool need_center, double eps,
    WalkOut w, int* __restrict__ wcount, const int* __restrict__ order,
    const int* __restrict__ start, const double4* __restrict__ sum4,
    HilbertOut h)
{
    using Key = typename Layout<BITS>::Key;
    const Key* prefix = static_cast<const Key*>(o.prefix);
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    const int end = start[L + 1];
    for (int k = start[L] + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
         k < end; k += stride) {
        const int v = order[k];
        int pre = 0;
        if (L == 0) {
            if (h.count)
                h.offset[v] = 0;
            const int wc = size[0] < o.n_nodes ? size[0] : o.n_nodes;
            *wcount = wc;
            // Terminal record at the walk length ("error 4, skip to the
            // end"): a walk reaches position wc only by finishing, and the
            // traversal's internal-node loop stops at any record of another
            // kind, so it needs no separate bound test.  The buffers hold
            // more than n_nodes records, and position wc may still hold a
            // record of an earlier, longer walk.
            if (MIXED)
                w.w32link[2 * static_cast<size_t>(wc)] =
                    make_int4(0x7F800000, wc, KIND_ERROR + 4, 0);
            else
                w.w64[wc] = error_record64(wc);
        } else {
            const int parent = o.parent[v];
            pre = pre_of[parent] + local[v];
            if (h.count)
                h.offset[v] += h.offset[parent];
        }
        pre_of[v] = pre;
        if (pre < 0 || pre >= o.n_nodes)
            continue;
        double4 geom = make_double4(0.0, 0.0, 0.0, 0.0);
        if (need_center) {
            int code = 0;
            node_geometry<BITS>(prefix[v], L, box[0], box[1], box[2], box[3],
                                geom.x, geom.y, geom.z, geom.w, code);
        }
        write_record<false, MIXED>(o, v, pre, size, crit, theta, rec64, rec32,
                                   box, need_center, eps, w, sum4, geom);
    }
}

// ------------------------------------------------- nodes grouped by level
// The octree numbers nodes by radix-tree node, not by level.  A counting
// sort groups them by level once per build, so that each level pass reads
// only its own nodes.  Order within a level is irrelevant: every node of a
// level is processed independently.  Three launches with the same grid:
// per-block level counts, one scan over (level, block), and a scatter
// through per-block cursors in LDS (no global atomics).
// info layout: [0, kMaxLevels + 1) level starts (start[L + 1] ends level L),
// then the kMaxLevels x blocks per-block counts / bases, level-major.
constexpr int kMaxLevels = 32; // 63-bit keys give at most 22 levels
constexpr int kLevelBlocks = 256;
constexpr int kLevelCounts = kMaxLevels * kLevelBlocks;
constexpr int kLevelInfoInts = kMaxLevels + 1 + kLevelCounts;

// Internal nodes only: leaves are handled by the per-group leaf passes.
__device__ __forceinline__ int node_level(const nbh_octree_view& o, int v)
{
    if (o.particle_count[v] > 0)
        return -1;
    const int L = o.level[v];
    return (L >= 0 && L < kMaxLevels) ? L : -1;
}

__global__ __launch_bounds__(256) void level_count_kernel(nbh_octree_view o,
                                                          int* __restrict__ info)
{
    __shared__ int cnt[kMaxLevels];
    if (threadIdx.x < kMaxLevels)
        cnt[threadIdx.x] = 0;
    __syncthreads();
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    for (int v = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
         v < o.n_nodes; v += stride) {
        const int L = node_level(o, v);
        if (L >= 0)
            atomicAdd(&cnt[L], 1);
    }
    __syncthreads();
    if (threadIdx.x < kMaxLevels)
        info[kMaxLevels + 1 + threadIdx.x * gridDim.x + blockIdx.x] = cnt[threadIdx.x];
}

// One block of 1024 threads: exclusive scan of the level-major counts (zero
// beyond `blocks`) staged in LDS.
__global__ __launch_bounds__(1024) void level_scan_kernel(int* __restrict__ info,
                                                          int blocks)
{
    constexpr int kPer = kLevelCounts / 1024;
    __shared__ int buf[kLevelCounts];
    __shared__ int part[1024];
    int* bc = info + kMaxLevels + 1;
    const int total = kMaxLevels * blocks;
    const int t = static_cast<int>(threadIdx.x);
    for (int k = t; k < kLevelCounts; k += 1024)
        buf[k] = k < total ? bc[k] : 0;
    __syncthreads();
    int sum = 0;
#pragma unroll
    for (int j = 0; j < kPer; ++j)
        sum += buf[t * kPer + j];
    part[t] = sum;
    __syncthreads();
    for (int off = 1; off < 1024; off <<= 1) {
        const int add = t >= off ? part[t - off] : 0;
        __syncthreads();
        part[t] += add;
        __syncthreads();
    }
    int run = part[t] - sum;
#pragma unroll
    for (int j = 0; j < kPer; ++j) {
        const int c = buf[t * kPer + j];
        buf[t * kPer + j] = run;
        run += c;
    }
    __syncthreads();
    for (int k = t; k < total; k += 1024)
        bc[k] = buf[k];
    if (t < kMaxLevels)
        info[t] = buf[t * blocks];
    if (t == 0)
        info[kMaxLevels] = part[1023];
}

__global__ __launch_bounds__(256) void level_scatter_kernel(nbh_octree_view o,
                                                            const int* __restrict__ info,
                                                            int* __restrict__ order)
{
    __shared__ int cursor[kMaxLevels];
    if (threadIdx.x < kMaxLevels)
        cursor[threadIdx.x] =
            info[kMaxLevels + 1 + threadIdx.x * gridDim.x + blockIdx.x];
    __syncthreads();
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    for (int v = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
         v < o.n_nodes; v += stride) {
        const int L = node_level(o, v);
        if (L >= 0)
            order[atomicAdd(&cursor[L], 1)] = v;
    }
}

template <int BITS>
__global__ __launch_bounds__(256) void octree_views_kernel(
    nbh_octree_view o, const double4* __restrict__ sum4,
    const double* __restrict__ box)
{
    using Key = typename Layout<BITS>::Key;
    const int v = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (v >= o.n_nodes)
        return;
    const Key* prefix = static_cast<const Key*>(o.prefix);
    double gx = 0.0, gy = 0.0, gz = 0.0, gh = 0.0;
    int code = 0;
    node_geometry<BITS>(prefix[v], o.level[v], box[0], box[1], box[2], box[3],
                        gx, gy, gz, gh, code);
    o.center_x[v] = gx;
    o.center_y[v] = gy;
    o.center_z[v] = gz;
    o.half_size[v] = gh;
    const double4 mc = sum4[v];
    o.mass[v] = mc.w;
    o.com_x[v] = mc.x;
    o.com_y[v] = mc.y;
    o.com_z[v] = mc.z;
    int nch = 0;
    if (o.particle_count[v] <= 0)
        for (int oc = 0; oc < 8; ++oc)
            nch += o.children[8 * static_cast<size_t>(v) + oc] != -1;
    o.child_count[v] = nch;
    o.pending_children[v] = nch;
}

// ------------------------------------------------------------ leaf passes
// Leaves depend on their own particles only, so they need no level order:
// one thread per Morton group g, whose leaf is the last chain node of radix
// leaf g, node edgeNodeOffset[g] + edgeNodeCount[g] (octree numbering of
// octree_builder.cu:49-224).  Groups, their particles and their leaves are
// all in Morton order, so these passes read and write nearly sequentially.
__device__ __forceinline__ int leaf_of_group(const int* __restrict__ poffset,
                                             const int* __restrict__ pcount,
                                             const nbh_octree_view& o, int g)
{
    const int v = poffset[g] + pcount[g];
    return (v > 0 && v < o.n_nodes && o.particle_count[v] > 0) ? v : -1;
}

// Leaf sums (octree_physics.cu:94-153, particles in sorted order) and
// subtree size 1, ahead of the internal levels.
__global__ __launch_bounds__(256) void leaf_up_kernel(
    nbh_octree_view o, const int* __restrict__ poffset,
    const int* __restrict__ pcount, int n_groups,
    const double4* __restrict__ rec, double4* __restrict__ sum4,
    int* __restrict__ size, HilbertOut h)
{
    const int g = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= n_groups)
        return;
    const int v = leaf_of_group(poffset, pcount, o, g);
    if (v < 0)
        return;
    const int f = o.first_particle[v];
    const int pc = o.particle_count[v];
    double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
    for (int k = 0; k < pc; ++k) {
        const double4 p = rec[f + k];
        tm += p.w;
        wx = fma(p.w, p.x, wx);
        wy = fma(p.w, p.y, wy);
        wz = fma(p.w, p.z, wz);
    }
    double cx = 0.0, cy = 0.0, cz = 0.0;
    if (tm > 0.0) {
        const double inv = 1.0 / tm;
        cx = wx * inv;
        cy = wy * inv;
        cz = wz * inv;
    }
    sum4[v] = make_double4(cx, cy, cz, tm);
    size[v] = 1;
    if (h.count)
        h.count[v] = pc;
}

// Leaf positions and records, after every internal level has its position.
template <int BITS, bool MIXED>
__global__ __launch_bounds__(256) void leaf_down_kernel(
    nbh_octree_view o, const int* __restrict__ poffset,
    const int* __restrict__ pcount, int n_groups, const int* __restrict__ size,
    const int* __restrict__ local, int* __restrict__ pre_of,
    const double* __restrict__ crit, double theta,
    const double4* __restrict__ rec64, const int4* __restrict__ rec32,
    const double* __restrict__ box, bool need_center, double eps,
    WalkOut w, const double4* __restrict__ sum4, HilbertOut h)
{
    using Key = typename Layout<BITS>::Key;
    const int g = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= n_groups)
        return;
    const int v = leaf_of_group(poffset, pcount, o, g);
    if (v < 0)
        return;
    const int parent = o.parent[v];
    const bool has_parent = parent >= 0 && parent < o.n_nodes;
    if (h.order) {
        // the leaf's particles, in sorted order, at its curve offset (its
        // parent's offset is absolute once every internal level is done)
        const int off = (v == 0) ? 0 : has_parent ? h.offset[v] + h.offset[parent] : -1;
        const int f = o.first_particle[v];
        const int pc = o.particle_count[v];
        if (off >= 0 && pc > 0 && off <= o.n_particles - pc)
            for (int k = 0; k < pc; ++k)
                h.order[off + k] = static_cast<uint32_t>(f + k);
    }
    if (!has_parent)
        return;
    const int pre = pre_of[parent] + local[v];
    pre_of[v] = pre;
    if (pre < 0 || pre >= o.n_nodes)
        return;
    double4 geom = make_double4(0.0, 0.0, 0.0, 0.0);
    if (need_center) {
        int code = 0;
        const Key* prefix = static_cast<const Key*>(o.prefix);
        node_geometry<BITS>(prefix[v], o.level[v], box[0], box[1], box[2], box[3],
                            geom.x, geom.y, geom.z, geom.w, code);
    }
    write_record<false, MIXED>(o, v, pre, size, crit, theta, rec64, rec32, box,
                               need_center, eps, w, sum4, geom);
}

// Every link starts as "error 4, skip to the end", so a position that the
// build leaves unwritten (a node off the root's tree) stops the walk with a
// topology error instead of reading an undefined record.
__global__ __launch_bounds__(256) void link_sentinel_kernel(int4* __restrict__ link,
                                                            int4 v, int n,
                                                            int stride)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        link[static_cast<size_t>(i) * stride] = v;
}

// The same for FP64 records: n records "error 4, skip to `skip`".
__global__ __launch_bounds__(256) void record_sentinel_kernel(Walk64Rec* __restrict__ w,
                                                              int n, int skip)
{
    const int i = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (i < n)
        w[i] = error_record64(skip);
}

} // namespace

void launch_generic_sizes(hipStream_t s, const nbh_octree_view& o, int* size,
                          int* local, int* err)
{
    check_parents_kernel<<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(o, err);
    NBH_HIP(hipGetLastError());
    generic_sizes_kernel<<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(o, size,
                                                                    local, err);
    NBH_HIP(hipGetLastError());
}

// Positions [0, M] start as "error 4, skip to M", position M stays the
// terminal record, and every node on the root's tree writes its own.
void launch_walk_build(hipStream_t s, const nbh_octree_view& o, const int* size,
                       const int* local, double theta, const double4* rec64,
                       bool need_center, double eps, double G, Walk64Rec* w64,
                       double4* w64center, int* wcount)
{
    const int M = o.n_nodes;
    record_sentinel_kernel<<<blocks_for(static_cast<size_t>(M) + 1, 256), 256, 0, s>>>(
        w64, M + 1, M);
    NBH_HIP(hipGetLastError());
    const WalkOut w{w64, w64center, nullptr, nullptr, nullptr, G};
    walk_build_kernel<true, false><<<blocks_for(M, 256), 256, 0, s>>>(
        o, size, local, nullptr, theta, rec64, nullptr, nullptr, need_center, eps, w,
        wcount);
    NBH_HIP(hipGetLastError());
}

// The pipeline's mass/COM + walk layout: levels max_level..0 bottom-up, then
// 0..max_level top-down.  Each launch strides over the node list of its
// level (`order`, grouped by a counting sort over the node levels).
template <int BITS>
void launch_level_passes(hipStream_t s, const nbh_octree_view& o, int max_level,
                         bool mixed, const double4* rec64, const int4* rec32,
                         double theta, const double* box,
                         bool need_center, double eps, int* size, int* local, double* crit,
                         int* pre, int* order, int* level_info, double4* sum4,
                         const int* poffset, const int* pcount, int n_groups,
                         double G, Walk64Rec* w64, double4* w64center, int4* w32node,
                         int4* w32link, int4* w32center, int* wcount,
                         hipEvent_t between, int* hcount, int* hoffset,
                         uint32_t* horder, const uint8_t* htab)
{
    // Mixed records are interleaved {node, link} pairs: w32node is the array
    // base and w32link = w32node + 1, both indexed at 2 * position.
    (void)w32link;
    const HilbertOut h{hcount, hoffset, horder, htab};
    int4* w32pair_link = w32node ? w32node + 1 : nullptr;
    // Single-child nodes are collapsed into their child in the mixed walk:
    // they carry the child's mass and centre of mass, and for theta below
    // 2/sqrt(3) accepting the parent implies accepting the child, so every
    // lane's interaction list is unchanged.
    const bool collapse = mixed && theta < kContainsTheta;
    if (max_level >= kMaxLevels)
        fail(NBH_LOGIC_ERROR, "octree depth exceeds the level-pass limit");
    unsigned blocks = blocks_for(o.n_nodes, 256);
    if (blocks > 2048)
        blocks = 2048;
    unsigned lblocks = blocks_for(o.n_nodes, 256);
    if (lblocks > kLevelBlocks)
        lblocks = kLevelBlocks;
    level_count_kernel<<<lblocks, 256, 0, s>>>(o, level_info);
    NBH_HIP(hipGetLastError());
    level_scan_kernel<<<1, 1024, 0, s>>>(level_info, static_cast<int>(lblocks));
    NBH_HIP(hipGetLastError());
    level_scatter_kernel<<<lblocks, 256, 0, s>>>(o, level_info, order);
    NBH_HIP(hipGetLastError());
    const int* start = level_info;
    const unsigned gblocks = blocks_for(n_groups > 0 ? n_groups : 1, 256);
    leaf_up_kernel<<<gblocks, 256, 0, s>>>(o, poffset, pcount, n_groups, rec64, sum4,
                                           size, h);
    NBH_HIP(hipGetLastError());
    for (int L = max_level; L >= 0; --L) {
        level_up_kernel<BITS><<<blocks, 256, 0, s>>>(o, L, rec64, theta, box, size,
                                                     local, crit, collapse, order,
                                                     start, sum4, h);
        NBH_HIP(hipGetLastError());
    }
    if (between)
        NBH_HIP(hipEventRecord(between, s));
    // The pipeline's octree comes from its own builder: every node lies on
    // the root's tree and every position below the walk length is written
    // here on each rebuild, as is the terminal record at the walk length.
    // The error sentinel of the other positions is written once, when the
    // record buffers are allocated (launch_walk_sentinel); the buffers hold
    // at least n_nodes + 1024 records.
    const WalkOut w{w64, w64center, w32node, w32pair_link,
                    w32center, G};
    for (int L = 0; L <= max_level; ++L) {
        if (mixed)
            level_down_kernel<BITS, true><<<blocks, 256, 0, s>>>(
                o, L, size, local, pre, crit, theta, rec64, rec32, box,
                need_center, eps, w, wcount, order, start, sum4, h);
        else
            level_down_kernel<BITS, false><<<blocks, 256, 0, s>>>(
                o, L, size, local, pre, crit, theta, rec64, rec32, box,
                need_center, eps, w, wcount, order, start, sum4, h);
        NBH_HIP(hipGetLastError());
    }
    // Leaves last: a leaf that collapsed its single-child parent overwrites
    // the parent's record at the shared position.
    if (mixed)
        leaf_down_kernel<BITS, true><<<gblocks, 256, 0, s>>>(
            o, poffset, pcount, n_groups, size, local, pre, crit, theta, rec64, rec32,
            box, need_center, eps, w, sum4, h);
    else
        leaf_down_kernel<BITS, false><<<gblocks, 256, 0, s>>>(
            o, poffset, pcount, n_groups, size, local, pre, crit, theta, rec64, rec32,
            box, need_center, eps, w, sum4, h);
    NBH_HIP(hipGetLastError());
}

size_t level_info_ints() { return kLevelInfoInts; }

// Writes the error sentinel "error 4, skip to `cap`" into every link of
// freshly allocated walk records.  Records are later overwritten only by
// the layout passes, whose skip links always point forward, so a walk over
// these buffers always ends.
void launch_walk_sentinel(hipStream_t s, int4* w32node, Walk64Rec* w64, int cap)
{
    if (w32node)
        link_sentinel_kernel<<<blocks_for(cap, 256), 256, 0, s>>>(
            w32node + 1, make_int4(0x7F800000, cap, KIND_ERROR + 4, 0), cap, 2);
    // FP64: skip to the last position, at or beyond which every walk ends
    if (w64)
        record_sentinel_kernel<<<blocks_for(cap, 256), 256, 0, s>>>(w64, cap, cap - 1);
    NBH_HIP(hipGetLastError());
}

// The reference-layout per-node arrays of the pipeline's octree (mass and
// centre of mass, cell centre and half size, child counts), filled from the
// packed sums on request: the level passes themselves only keep sum4.
template <int BITS>
void launch_octree_views(hipStream_t s, const nbh_octree_view& o,
                         const double4* sum4, const double* box)
{
    octree_views_kernel<BITS><<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(o, sum4, box);
    NBH_HIP(hipGetLastError());
}
template void launch_octree_views<10>(hipStream_t, const nbh_octree_view&,
                                      const double4*, const double*);
template void launch_octree_views<21>(hipStream_t, const nbh_octree_view&,
                                      const double4*, const double*);

#define NBH_LEVEL_INST(B)                                                            \
    template void launch_level_passes<B>(hipStream_t, const nbh_octree_view&, int, bool, \
                                         const double4*, const int4*, double,           \
                                         const double*, bool, double, int*, int*,       \
                                         double*, int*, int*, int*, double4*,           \
                                         const int*, const int*, int, double, Walk64Rec*, \
                                         double4*, int4*, int4*, int4*, int*, hipEvent_t, \
                                         int*, int*, uint32_t*, const uint8_t*);
NBH_LEVEL_INST(10)
NBH_LEVEL_INST(21)
#undef NBH_LEVEL_INST


} // namespace nbh
