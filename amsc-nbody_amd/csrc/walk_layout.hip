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
#include "sim_internal.hpp"
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
// The traversals' target order along the Hilbert curve (see hilbert.cpp) is
// built alongside the level passes.  A node's curve offset is the rank of
// its subtree's first particle along the curve: the bottom-up pass writes
// each child's offset relative to its parent, the top-down pass adds the
// parent's absolute offset (as for pre-order positions).
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

// Outputs of the walk build (records indexed by pre-order position).
struct WalkOut {
    Walk64Rec* w64;
    double4* w64center;
    int4* w32node;
    int4* w32link;
    int4* w32center;
    double G; // FP64 records hold G m
};

// What one node's walk record states: packed {com, mass}, cell geometry
// {centre, half}, the opening radius (internal and childless nodes), the
// kind, a leaf's particle range and the position after the subtree.
struct NodeVals {
    double4 mc;
    double4 geom;
    double crit;
    int kind, first, count, skip;
};

// The reference's per-node conditions for a caller-supplied octree
// (barnes_hut.cu:101-162, 273-294), encoded as record kinds, with the
// opening radius computed here.
__device__ __forceinline__ NodeVals generic_vals(const nbh_octree_view& o, int v,
                                                 int pre, const int* __restrict__ size,
                                                 double theta)
{
    NodeVals n;
    n.skip = pre + size[v];
    n.mc = make_double4(o.com_x[v], o.com_y[v], o.com_z[v], o.mass[v]);
    n.geom = make_double4(o.center_x[v], o.center_y[v], o.center_z[v], o.half_size[v]);
    n.first = 0;
    n.count = 0;
    n.crit = 0.0;
    const int pc = o.particle_count[v];
    const int nch = o.child_count[v];
    const double h = n.geom.w;
    if (!(h > 0.0)) {
        n.kind = KIND_ERROR + 8;
    } else if (pc > 0) {
        n.first = o.first_particle[v];
        n.count = pc;
        if (nch != 0)
            n.kind = KIND_ERROR + 7;
        else if (n.first < 0 || pc > o.n_particles || n.first > o.n_particles - pc)
            n.kind = KIND_ERROR + 2;
        else
            n.kind = KIND_LEAF;
    } else if (o.mass[v] == 0.0) {
        n.kind = KIND_SKIP;
    } else if (nch == 0) {
        n.kind = KIND_CHILDLESS;
    } else {
        n.kind = KIND_INTERNAL;
    }
    if (n.kind == KIND_INTERNAL || n.kind == KIND_CHILDLESS) {
        const double ox = n.mc.x - n.geom.x;
        const double oy = n.mc.y - n.geom.y;
        const double oz = n.mc.z - n.geom.z;
        const double o2 = fma(oz, oz, fma(ox, ox, oy * oy));
        n.crit = (2.0 * h) / theta + sqrt(o2);
    }
    return n;
}

// The record of a node at pre-order position `pre`.
template <bool MIXED>
__device__ __forceinline__ void emit_record(const NodeVals& nv, int pre,
                                            const double4* __restrict__ rec64,
                                            const int4* __restrict__ rec32,
                                            const double* __restrict__ box,
                                            bool need_center, double eps,
                                            const WalkOut& w)
{
    const int kind = nv.kind;
    const int first = nv.first;
    int count = nv.count;
    const double4 mc = nv.mc;
    const double crit = nv.crit;
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
            r.x = mc.x;
            r.y = mc.y;
            r.z = mc.z;
            r.m = mc.w;
        }
        r.skip = nv.skip;
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
            double4 c = nv.geom;
            if (fast)
                c.w = -1.0;
            w.w64center[pre] = c;
        }
    } else {
        const double side = box[3];
        const double scale = 2147483648.0 / side;
        const double ox = box[0] - side / 2.0;
        const double oy = box[1] - side / 2.0;
        const double oz = box[2] - side / 2.0;
        const double lim = 2147483647.0;
        if (kind == KIND_LEAF && count == 1) {
            w.w32node[2 * static_cast<size_t>(pre)] = rec32[first];
        } else {
            const int qx = static_cast<int>(fmin(fmax(rint((mc.x - ox) * scale), 0.0), lim));
            const int qy = static_cast<int>(fmin(fmax(rint((mc.y - oy) * scale), 0.0), lim));
            const int qz = static_cast<int>(fmin(fmax(rint((mc.z - oz) * scale), 0.0), lim));
            const float gm = static_cast<float>(mc.w / box[4]);
            w.w32node[2 * static_cast<size_t>(pre)] = make_int4(qx, qy, qz, __float_as_int(gm));
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
        w.w32link[2 * static_cast<size_t>(pre)] =
            make_int4(__float_as_int(crit2), nv.skip, kind_out | (count << 8), first);
        if (need_center) {
            const int qx = static_cast<int>(rint((nv.geom.x - ox) * scale));
            const int qy = static_cast<int>(rint((nv.geom.y - oy) * scale));
            const int qz = static_cast<int>(rint((nv.geom.z - oz) * scale));
            const int qh = fast ? -1 : static_cast<int>(fmin(rint(nv.geom.w * scale), lim));
            w.w32center[pre] = make_int4(qx, qy, qz, qh);
        }
    }
}

__global__ __launch_bounds__(256) void walk_build_kernel(
    nbh_octree_view o, const int* __restrict__ size,
    const int* __restrict__ local, double theta, const double4* __restrict__ rec64,
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
    emit_record<false>(generic_vals(o, v, pre, size, theta), pre, rec64, nullptr, nullptr,
                       need_center, eps, w);
}


// ------------------------------------------ level-synchronous passes
// Bottom-up over octree levels: every child of a level-L node sits at level
// L+1 and was completed by the previous launch, so no arrival counters (and
// no cache write-back/invalidate per node) are needed.  A node's values are
// the reference's: leaf mass/COM summed over its particles in sorted order
// (octree_physics.cu:94-153, done by the octree materialisation), internal
// nodes combined in octant order 0..7 (octree_physics.cu:211-243), cell
// geometry from the prefix path (octree_geometry.cu:56-102), the opening
// radius side/theta + |com - centre| (barnes_hut.cu:277-291); plus the
// subtree size, the pre-order offsets of the children (octants 7 -> 0) and
// the children's Hilbert-curve offsets.
// A leaf's mass / centre of mass over its pc particles in sorted order
// (octree_physics.cu:94-153), given its first particle p0 = rec[f].
__device__ __forceinline__ double4 leaf_sum(double4 p0, const double4* __restrict__ rec,
                                            int f, int pc)
{
    double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
    if (pc > 0) {
        tm += p0.w;
        wx = fma(p0.w, p0.x, wx);
        wy = fma(p0.w, p0.y, wy);
        wz = fma(p0.w, p0.z, wz);
    }
    for (int k = 1; k < pc; ++k) {
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
    return make_double4(cx, cy, cz, tm);
}

// One internal node v of level L (s_tab: the curve tables in LDS).
template <int BITS>
__device__ __forceinline__ void level_up_node(const nbh_octree_view& o, int L, int v,
                                              const LevelArrays& la,
                                              const double4* __restrict__ rec64,
                                              const double* __restrict__ box, double theta,
                                              bool collapse, const uint8_t* s_tab)
{
    using Key = typename Layout<BITS>::Key;
    const double rcx = box[0], rcy = box[1], rcz = box[2], side = box[3];
    {
        const Key pv = static_cast<const Key*>(o.prefix)[v];
        const int4 c0 = reinterpret_cast<const int4*>(o.children)[2 * static_cast<size_t>(v)];
        const int4 c1 =
            reinterpret_cast<const int4*>(o.children)[2 * static_cast<size_t>(v) + 1];
        const int ch[8] = {c0.x, c0.y, c0.z, c0.w, c1.x, c1.y, c1.z, c1.w};
        // Every present child's words are requested before any is used.
        // Leaves are octree nodes 1 .. n_leaves (radix leaf g owns node
        // g + 1): a leaf child's sums follow from its particles, an internal
        // child's come from its own level pass.
        double4 cs[8];
        int4 ca[8]; // size, hcount, particle count
        int lf[8];  // a leaf child's first particle
#pragma unroll
        for (int oc = 0; oc < 8; ++oc) {
            ca[oc] = make_int4(0, 0, 0, 0);
            lf[oc] = 0;
            const int c = ch[oc];
            if (c != -1 && c <= o.n_leaves) {
                lf[oc] = o.first_particle[c];
                const int pc = o.particle_count[c];
                ca[oc] = make_int4(1, pc, pc, 0);
            }
        }
#pragma unroll
        for (int oc = 0; oc < 8; ++oc) {
            cs[oc] = make_double4(0.0, 0.0, 0.0, 0.0);
            const int c = ch[oc];
            if (c != -1) {
                if (c <= o.n_leaves) {
                    if (ca[oc].z > 0)
                        cs[oc] = rec64[lf[oc]]; // the leaf's first particle
                } else {
                    cs[oc] = la.sum4[c];
                    ca[oc] = la.aux[c];
                }
            }
        }
#pragma unroll
        for (int oc = 0; oc < 8; ++oc)
            if (ch[oc] != -1 && ch[oc] <= o.n_leaves)
                cs[oc] = leaf_sum(cs[oc], rec64, lf[oc], ca[oc].z);
        double tm = 0.0, wx = 0.0, wy = 0.0, wz = 0.0;
        int nch = 0, only_pc = 0; // child count, the last child's particle count
#pragma unroll
        for (int oc = 0; oc < 8; ++oc) {
            if (ch[oc] == -1)
                continue;
            const double4 c = cs[oc];
            tm += c.w;
            wx = fma(c.w, c.x, wx);
            wy = fma(c.w, c.y, wy);
            wz = fma(c.w, c.z, wz);
            only_pc = ca[oc].z;
            ++nch;
        }
        // Curve offsets of the children: in v's orientation state they take
        // consecutive particle ranges in the order of their ranks.
        int hc = 0;
#pragma unroll
        for (int oc = 0; oc < 8; ++oc)
            hc += ca[oc].y;
        const uint8_t* rank =
            s_tab + 8 * curve_state<BITS>(pv, L, s_tab + 8 * kHilbertStates);
        // A collapsed single-child node takes no walk position of its own:
        // its child starts at the node's position and overwrites the node's
        // record on the next top-down level.  Only internal children and
        // single-particle leaves qualify (a leaf holding several particles is
        // summed exactly, never as a monopole).
        int off = (collapse && nch == 1 && only_pc <= 1) ? 0 : 1;
#pragma unroll
        for (int oc = 7; oc >= 0; --oc) {
            if (ch[oc] == -1)
                continue;
            int before = 0;
#pragma unroll
            for (int b = 0; b < 8; ++b)
                before += (rank[b] < rank[oc]) ? ca[b].y : 0;
            la.link[ch[oc]] = make_int2(off, before);
            off += ca[oc].x;
        }
        double cx = 0.0, cy = 0.0, cz = 0.0;
        if (tm > 0.0) {
            const double inv = 1.0 / tm;
            cx = wx * inv;
            cy = wy * inv;
            cz = wz * inv;
        }
        double gx, gy, gz, gh;
        int code = 0;
        node_geometry<BITS>(pv, L, rcx, rcy, rcz, side, gx, gy, gz, gh, code);
        const double ox = cx - gx;
        const double oy = cy - gy;
        const double oz = cz - gz;
        const double o2 = fma(oz, oz, fma(ox, ox, oy * oy));
        la.sum4[v] = make_double4(cx, cy, cz, tm);
        // aux.w: collapsed into its child (no walk position of its own)
        la.aux[v] = make_int4(off, hc, 0, (collapse && nch == 1 && only_pc <= 1) ? 1 : 0);
        la.crit[v] = (2.0 * gh) / theta + sqrt(o2);
    }
}

template <int BITS>
__global__ __launch_bounds__(256) void level_up_kernel(
    nbh_octree_view o, int L, int begin, int end, const int* __restrict__ order,
    LevelArrays la, const double4* __restrict__ rec64, const double* __restrict__ box,
    double theta, bool collapse, const uint8_t* __restrict__ tab)
{
    __shared__ uint8_t s_tab[kHilbertTabBytes];
    for (int t = static_cast<int>(threadIdx.x); t < kHilbertTabBytes; t += blockDim.x)
        s_tab[t] = tab[t];
    __syncthreads();
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    for (int k = begin + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); k < end;
         k += stride) {
        const int v = order[k];
        if (v < 0 || v >= o.n_nodes)
            continue;
        level_up_node<BITS>(o, L, v, la, rec64, box, theta, collapse, s_tab);
    }
}

// Consecutive levels hi .. lo (hi >= lo) of at most kLevelRunNodes nodes
// each, bottom-up in one block: a small level is bound by the latency of its
// dependent loads, which a launch per level pays from HBM and this block
// pays from L2 (the previous level's writes).  The block's barrier orders
// the levels (all of a workgroup's wavefronts share one CU and its L1).
constexpr int kLevelRunThreads = 512;
constexpr int kLevelRunNodes = 512;

template <int BITS>
__global__ __launch_bounds__(kLevelRunThreads) void level_up_run_kernel(
    nbh_octree_view o, int hi, int lo, LevelStarts st, const int* __restrict__ order,
    LevelArrays la, const double4* __restrict__ rec64, const double* __restrict__ box,
    double theta, bool collapse, const uint8_t* __restrict__ tab)
{
    __shared__ uint8_t s_tab[kHilbertTabBytes];
    for (int t = static_cast<int>(threadIdx.x); t < kHilbertTabBytes; t += blockDim.x)
        s_tab[t] = tab[t];
    __syncthreads();
    for (int L = hi; L >= lo; --L) {
        for (int k = st.s[L] + static_cast<int>(threadIdx.x); k < st.s[L + 1];
             k += kLevelRunThreads) {
            const int v = order[k];
            if (v >= 0 && v < o.n_nodes)
                level_up_node<BITS>(o, L, v, la, rec64, box, theta, collapse, s_tab);
        }
        __syncthreads();
    }
}

// Top-down over one level: pre-order position = parent's + own offset, the
// curve offset made absolute, then the node's walk record.  The root
// (v = 0, level 0) also writes the walk length and the terminal record at
// it ("error 4, skip to the end"): a walk reaches that position only by
// finishing, and the traversal's internal-node loop stops at any record of
// another kind, so it needs no separate bound test.  The record buffers
// hold more than n_nodes records, and the position may still hold a record
// of an earlier, longer walk.
template <int BITS, bool MIXED>
__device__ __forceinline__ void level_down_node(const nbh_octree_view& o, int L, int v,
                                                const LevelArrays& la,
                                                const double4* __restrict__ rec64,
                                                const int4* __restrict__ rec32,
                                                const double* __restrict__ box,
                                                bool need_center, double eps, const WalkOut& w,
                                                int* __restrict__ wcount)
{
    using Key = typename Layout<BITS>::Key;
    {
        const int4 a = la.aux[v];
        int pre = 0, hoff = 0;
        if (v == 0) {
            const int wc = a.x < o.n_nodes ? a.x : o.n_nodes;
            *wcount = wc;
            if (MIXED)
                w.w32link[2 * static_cast<size_t>(wc)] =
                    make_int4(0x7F800000, wc, KIND_ERROR + 4, 0);
            else
                w.w64[wc] = error_record64(wc);
        } else {
            const int parent = o.parent[v];
            if (parent < 0 || parent >= o.n_nodes)
                return;
            const int2 lk = la.link[v];
            const int2 p = la.pos[parent];
            pre = p.x + lk.x;
            hoff = p.y + lk.y;
        }
        la.pos[v] = make_int2(pre, hoff);
        if (pre < 0 || pre >= o.n_nodes)
            return;
        NodeVals nv;
        nv.mc = la.sum4[v];
        nv.geom = make_double4(0.0, 0.0, 0.0, 0.0);
        nv.crit = la.crit[v];
        nv.kind = (nv.mc.w == 0.0) ? KIND_SKIP : KIND_INTERNAL;
        nv.first = 0;
        nv.count = 0;
        nv.skip = pre + a.x;
        if (need_center) {
            int code = 0;
            node_geometry<BITS>(static_cast<const Key*>(o.prefix)[v], L, box[0], box[1],
                                box[2], box[3], nv.geom.x, nv.geom.y, nv.geom.z,
                                nv.geom.w, code);
        }
        emit_record<MIXED>(nv, pre, rec64, rec32, box, need_center, eps, w);
    }
}

template <int BITS, bool MIXED>
__global__ __launch_bounds__(256) void level_down_kernel(
    nbh_octree_view o, int L, int begin, int end, const int* __restrict__ order,
    LevelArrays la, const double4* __restrict__ rec64, const int4* __restrict__ rec32,
    const double* __restrict__ box, bool need_center, double eps, WalkOut w,
    int* __restrict__ wcount)
{
    const int stride = static_cast<int>(gridDim.x * blockDim.x);
    for (int k = begin + static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x); k < end;
         k += stride) {
        const int v = order[k];
        if (v < 0 || v >= o.n_nodes)
            continue;
        level_down_node<BITS, MIXED>(o, L, v, la, rec64, rec32, box, need_center, eps, w,
                                     wcount);
    }
}

// Consecutive small levels lo .. hi top-down in one block (as
// level_up_run_kernel).
template <int BITS, bool MIXED>
__global__ __launch_bounds__(kLevelRunThreads) void level_down_run_kernel(
    nbh_octree_view o, int lo, int hi, LevelStarts st, const int* __restrict__ order,
    LevelArrays la, const double4* __restrict__ rec64, const int4* __restrict__ rec32,
    const double* __restrict__ box, bool need_center, double eps, WalkOut w,
    int* __restrict__ wcount)
{
    for (int L = lo; L <= hi; ++L) {
        for (int k = st.s[L] + static_cast<int>(threadIdx.x); k < st.s[L + 1];
             k += kLevelRunThreads) {
            const int v = order[k];
            if (v >= 0 && v < o.n_nodes)
                level_down_node<BITS, MIXED>(o, L, v, la, rec64, rec32, box, need_center, eps,
                                             w, wcount);
        }
        __syncthreads();
    }
}

// ------------------------------------------------- nodes grouped by level
// The octree numbers nodes by radix-tree node, not by level.  The radix ->
// octree plan already fixes every octree node's level: internal radix node
// r owns the chain of octree nodes 1 + offset[r] + k at levels
// level[r] - count[r] + 1 + k, k < count[r] (octree_builder.cu:153-174), a
// radix leaf owns one octree leaf, and the root is node 0 at level 0.  The
// plan scan counts the internal nodes per level (the counts join the
// mid-step readback and size the level launches); this kernel groups the
// nodes by level.  A block counts its nodes per level in LDS, reserves
// ranges of each level with one global add per level, and hands out slots
// in LDS.  Order within a level is irrelevant: every node of a level is
// processed independently.
constexpr int kScatterItems = 4;
constexpr int kScatterThreads = 256;

__global__ __launch_bounds__(kScatterThreads) void level_scatter_kernel(
    const int* __restrict__ d_nleaves, const int* __restrict__ plevel,
    const int* __restrict__ pcount, const int* __restrict__ poffset,
    const int* __restrict__ level_count, int* __restrict__ cursor,
    int* __restrict__ order)
{
    __shared__ int s_cnt[kMaxLevels];
    __shared__ int s_slot[kMaxLevels];
    const int t = static_cast<int>(threadIdx.x);
    if (t < kMaxLevels)
        s_cnt[t] = 0;
    __syncthreads();
    const int nl = *d_nleaves;
    const int ni = nl > 1 ? nl - 1 : 0;
    const int base = static_cast<int>(blockIdx.x) * (kScatterThreads * kScatterItems);
    int c[kScatterItems], fl[kScatterItems], fn[kScatterItems];
#pragma unroll
    for (int j = 0; j < kScatterItems; ++j) {
        const int i = base + j * kScatterThreads + t;
        const bool in = i < ni;
        const int r = nl + i;
        c[j] = in ? pcount[r] : 0;
        if (c[j] < 0 || c[j] > kMaxLevels)
            c[j] = 0; // a failed plan: reported by its own error word
        fl[j] = in ? plevel[r] - c[j] + 1 : 0;
        fn[j] = in ? 1 + poffset[r] : 0;
    }
    const bool root = blockIdx.x == 0 && t == 0;
    if (t < 64)
        (void)wave_claim(s_cnt, 0, root);
#pragma unroll
    for (int j = 0; j < kScatterItems; ++j)
        for (int k = 0; __ballot(k < c[j]) != 0ull; ++k) {
            const int L = fl[j] + k;
            (void)wave_claim(s_cnt, L, k < c[j] && L >= 0 && L < kMaxLevels);
        }
    __syncthreads();
    if (t < 64) {
        // level starts: exclusive sum of the per-level counts (plus the root
        // at level 0), lanes 0..kMaxLevels-1 of the first wave
        const int n_l = t < kMaxLevels ? level_count[t * kLevelCountStride] + (t == 0) : 0;
        const int start = wave_inclusive_sum(n_l) - n_l;
        if (t < kMaxLevels) {
            const int mine = s_cnt[t];
            s_slot[t] = start + (mine != 0 ? atomicAdd(&cursor[t * kLevelCountStride], mine) : 0);
        }
    }
    __syncthreads();
    if (t < 64) {
        const int slot = wave_claim(s_slot, 0, root);
        if (root)
            order[slot] = 0;
    }
#pragma unroll
    for (int j = 0; j < kScatterItems; ++j)
        for (int k = 0; __ballot(k < c[j]) != 0ull; ++k) {
            const int L = fl[j] + k;
            const bool act = k < c[j] && L >= 0 && L < kMaxLevels;
            const int slot = wave_claim(s_slot, L, act);
            if (act)
                order[slot] = fn[j] + k;
        }
}

template <int BITS>
__global__ __launch_bounds__(256) void octree_views_kernel(
    nbh_octree_view o, LevelArrays la, const double4* __restrict__ rec64,
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
    const int pc = o.particle_count[v];
    const double4 mc = pc > 0 ? leaf_sum(rec64[o.first_particle[v]], rec64,
                                         o.first_particle[v], pc)
                              : la.sum4[v];
    o.mass[v] = mc.w;
    o.com_x[v] = mc.x;
    o.com_y[v] = mc.y;
    o.com_z[v] = mc.z;
    int nch = 0;
    if (pc <= 0)
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
// A leaf's sums are formed from its particles where they are needed.
__device__ __forceinline__ int leaf_of_group(const int* __restrict__ poffset,
                                             const int* __restrict__ pcount,
                                             const nbh_octree_view& o, int g)
{
    const int v = poffset[g] + pcount[g];
    return (v > 0 && v < o.n_nodes && o.particle_count[v] > 0) ? v : -1;
}

// Leaf positions, curve ranks and records, after every internal level has
// its position.
template <int BITS, bool MIXED>
__global__ __launch_bounds__(256) void leaf_down_kernel(
    nbh_octree_view o, const int* __restrict__ poffset,
    const int* __restrict__ pcount, int n_groups, LevelArrays la,
    const double4* __restrict__ rec64, const int4* __restrict__ rec32,
    const double* __restrict__ box, bool need_center, double eps, WalkOut w,
    uint32_t* __restrict__ horder)
{
    using Key = typename Layout<BITS>::Key;
    const int g = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= n_groups)
        return;
    const int v = leaf_of_group(poffset, pcount, o, g);
    if (v < 0)
        return;
    const int parent = o.parent[v];
    if (parent < 0 || parent >= o.n_nodes)
        return;
    const int2 lk = la.link[v];
    const int2 p = la.pos[parent];
    const int f = o.first_particle[v];
    const int pc = o.particle_count[v];
    const int pre = p.x + lk.x;
    const int off = p.y + lk.y;
    if (horder) {
        // the leaf's particles, in sorted order, at its curve offset (its
        // parent's offset is absolute once every internal level is done)
        if (off >= 0 && pc > 0 && off <= o.n_particles - pc)
            for (int k = 0; k < pc; ++k)
                horder[off + k] = static_cast<uint32_t>(f + k);
    }
    if (pre < 0 || pre >= o.n_nodes)
        return;
    NodeVals nv;
    // a one-particle leaf's record is its particle's: no sums needed
    nv.mc = pc > 1 ? leaf_sum(rec64[f], rec64, f, pc) : make_double4(0.0, 0.0, 0.0, 0.0);
    nv.geom = make_double4(0.0, 0.0, 0.0, 0.0);
    nv.crit = 0.0;
    nv.kind = KIND_LEAF;
    nv.first = f;
    nv.count = pc;
    nv.skip = pre + 1;
    if (need_center) {
        int code = 0;
        node_geometry<BITS>(static_cast<const Key*>(o.prefix)[v], o.level[v], box[0],
                            box[1], box[2], box[3], nv.geom.x, nv.geom.y, nv.geom.z,
                            nv.geom.w, code);
    }
    emit_record<MIXED>(nv, pre, rec64, rec32, box, need_center, eps, w);
}

// Leaf positions {pre, curve offset} for the callers that read them per
// node (the force VJP's push-down); the layout passes keep them for the
// internal nodes only.
__global__ __launch_bounds__(256) void leaf_positions_kernel(
    nbh_octree_view o, const int* __restrict__ poffset, const int* __restrict__ pcount,
    int n_groups, LevelArrays la)
{
    const int g = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (g >= n_groups)
        return;
    const int v = leaf_of_group(poffset, pcount, o, g);
    if (v < 0)
        return;
    const int parent = o.parent[v];
    if (parent < 0 || parent >= o.n_nodes)
        return;
    const int2 lk = la.link[v];
    const int2 p = la.pos[parent];
    la.pos[v] = make_int2(p.x + lk.x, p.y + lk.y);
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
    walk_build_kernel<<<blocks_for(M, 256), 256, 0, s>>>(o, size, local, theta, rec64,
                                                         need_center, eps, w, wcount);
    NBH_HIP(hipGetLastError());
}

LevelStarts level_starts(const int* level_count)
{
    LevelStarts st{};
    long long run = 0;
    for (int L = 0; L < kMaxLevels; ++L) {
        st.s[L] = static_cast<int>(run);
        const int c = level_count[L * kLevelCountStride] + (L == 0 ? 1 : 0);
        if (c < 0)
            fail(NBH_LOGIC_ERROR, "inconsistent octree level counts");
        run += c;
        if (run > 0x7FFFFFFFll)
            fail(NBH_OVERFLOW_ERROR, "octree level counts exceed the int range");
    }
    st.s[kMaxLevels] = static_cast<int>(run);
    return st;
}

void launch_level_scatter(hipStream_t s, const int* d_nleaves, int n_leaves,
                          const int* plevel, const int* pcount, const int* poffset,
                          const int* level_count, int* cursor, int* order)
{
    const size_t ni = n_leaves > 1 ? static_cast<size_t>(n_leaves) - 1 : 0;
    const unsigned blocks = blocks_for(ni, kScatterThreads * kScatterItems);
    level_scatter_kernel<<<blocks, kScatterThreads, 0, s>>>(d_nleaves, plevel, pcount,
                                                            poffset, level_count, cursor,
                                                            order);
    NBH_HIP(hipGetLastError());
}

void launch_leaf_positions(hipStream_t s, const nbh_octree_view& o, const int* poffset,
                           const int* pcount, int n_groups, const LevelArrays& la)
{
    leaf_positions_kernel<<<blocks_for(n_groups > 0 ? n_groups : 1, 256), 256, 0, s>>>(
        o, poffset, pcount, n_groups, la);
    NBH_HIP(hipGetLastError());
}

// The pipeline's mass/COM + walk layout: the internal levels max_level..0
// bottom-up (each leaf's sums are formed by its parent from its particles), then
// 0..max_level top-down, then the leaves' records; one launch per level,
// sized to its node count, or one block per run of small levels.
template <int BITS>
void launch_level_passes(hipStream_t s, const nbh_octree_view& o, const LevelStarts& st,
                         bool mixed, const double4* rec64, const int4* rec32, double theta,
                         const double* box, bool need_center, double eps,
                         const LevelArrays& la, const int* order, const int* poffset,
                         const int* pcount, int n_groups, double G, Walk64Rec* w64,
                         double4* w64center, int4* w32node, int4* w32center, int* wcount,
                         hipEvent_t between, uint32_t* horder, const uint8_t* htab)
{
    // Single-child nodes are collapsed into their child in the mixed walk:
    // they carry the child's mass and centre of mass, and for theta below
    // 2/sqrt(3) accepting the parent implies accepting the child, so every
    // lane's interaction list is unchanged.
    const bool collapse = mixed && theta < kContainsTheta;
    if (st.s[0] != 0 || st.s[kMaxLevels] > o.n_nodes || st.s[1] != 1)
        fail(NBH_LOGIC_ERROR, "inconsistent octree level counts");
    int max_level = 0;
    for (int L = 0; L < kMaxLevels; ++L)
        if (st.s[L + 1] > st.s[L])
            max_level = L;
    const auto grid = [&](int L) {
        unsigned b = blocks_for(static_cast<size_t>(st.s[L + 1] - st.s[L]), 256);
        return b > 8192u ? 8192u : b;
    };
    // Runs of consecutive levels of at most kLevelRunNodes nodes each (the
    // tree's top levels, and its deepest ones) take one block each; the
    // other levels a launch each.
    const auto small = [&](int L) { return st.s[L + 1] - st.s[L] <= kLevelRunNodes; };
    for (int L = max_level; L >= 0;) {
        if (small(L)) {
            int lo = L;
            while (lo > 0 && small(lo - 1))
                --lo;
            level_up_run_kernel<BITS><<<1, kLevelRunThreads, 0, s>>>(
                o, L, lo, st, order, la, rec64, box, theta, collapse, htab);
            NBH_HIP(hipGetLastError());
            L = lo - 1;
            continue;
        }
        level_up_kernel<BITS><<<grid(L), 256, 0, s>>>(o, L, st.s[L], st.s[L + 1], order, la,
                                                      rec64, box, theta, collapse, htab);
        NBH_HIP(hipGetLastError());
        --L;
    }
    if (between)
        NBH_HIP(hipEventRecord(between, s));
    // The pipeline's octree comes from its own builder: every node lies on
    // the root's tree and every position below the walk length is written
    // here on each rebuild, as is the terminal record at the walk length.
    // The error sentinel of the other positions is written once, when the
    // record buffers are allocated (launch_walk_sentinel); the buffers hold
    // at least n_nodes + 1024 records.  Mixed records are interleaved
    // {node, link} pairs: w32node is the array base and the links start one
    // int4 later, both indexed at 2 * position.
    const WalkOut w{w64, w64center, w32node, w32node ? w32node + 1 : nullptr, w32center, G};
    for (int L = 0; L <= max_level;) {
        if (small(L)) {
            int hi = L;
            while (hi < max_level && small(hi + 1))
                ++hi;
            if (mixed)
                level_down_run_kernel<BITS, true><<<1, kLevelRunThreads, 0, s>>>(
                    o, L, hi, st, order, la, rec64, rec32, box, need_center, eps, w, wcount);
            else
                level_down_run_kernel<BITS, false><<<1, kLevelRunThreads, 0, s>>>(
                    o, L, hi, st, order, la, rec64, rec32, box, need_center, eps, w, wcount);
            NBH_HIP(hipGetLastError());
            L = hi + 1;
            continue;
        }
        if (mixed)
            level_down_kernel<BITS, true><<<grid(L), 256, 0, s>>>(
                o, L, st.s[L], st.s[L + 1], order, la, rec64, rec32, box, need_center, eps,
                w, wcount);
        else
            level_down_kernel<BITS, false><<<grid(L), 256, 0, s>>>(
                o, L, st.s[L], st.s[L + 1], order, la, rec64, rec32, box, need_center, eps,
                w, wcount);
        NBH_HIP(hipGetLastError());
        ++L;
    }
    // Leaves last: a leaf that collapsed its single-child parent overwrites
    // the parent's record at the shared position.
    const unsigned gblocks = blocks_for(n_groups > 0 ? n_groups : 1, 256);
    if (mixed)
        leaf_down_kernel<BITS, true><<<gblocks, 256, 0, s>>>(
            o, poffset, pcount, n_groups, la, rec64, rec32, box, need_center, eps, w, horder);
    else
        leaf_down_kernel<BITS, false><<<gblocks, 256, 0, s>>>(
            o, poffset, pcount, n_groups, la, rec64, rec32, box, need_center, eps, w, horder);
    NBH_HIP(hipGetLastError());
}

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
// level passes' packed sums on request: the passes themselves keep only those.
template <int BITS>
void launch_octree_views(hipStream_t s, const nbh_octree_view& o, const LevelArrays& la,
                         const double4* rec64, const double* box)
{
    octree_views_kernel<BITS><<<blocks_for(o.n_nodes, 256), 256, 0, s>>>(o, la, rec64, box);
    NBH_HIP(hipGetLastError());
}
template void launch_octree_views<10>(hipStream_t, const nbh_octree_view&,
                                      const LevelArrays&, const double4*, const double*);
template void launch_octree_views<21>(hipStream_t, const nbh_octree_view&,
                                      const LevelArrays&, const double4*, const double*);

#define NBH_LEVEL_INST(B)                                                               \
    template void launch_level_passes<B>(                                               \
        hipStream_t, const nbh_octree_view&, const LevelStarts&, bool, const double4*,  \
        const int4*, double, const double*, bool, double, const LevelArrays&,           \
        const int*, const int*, const int*, int, double, Walk64Rec*, double4*, int4*,   \
        int4*, int*, hipEvent_t, uint32_t*, const uint8_t*);
NBH_LEVEL_INST(10)
NBH_LEVEL_INST(21)
#undef NBH_LEVEL_INST

} // namespace nbh
