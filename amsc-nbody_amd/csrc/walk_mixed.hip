// walk_mixed.hip -- mixed-precision packet traversal.
//
// The opening test and every interaction run in FP32 on exact integer
// differences of 31-bit fixed-point coordinates over the bounding box, i.e.
// on coordinates relative to the target: the only rounding of a separation
// is the conversion of the exact integer difference, relative to the
// separation itself.  Node records hold m / m_max (m_max the largest |m|),
// and a target's sums times G m_max / unit^2 are its acceleration in
// physical units.
//
// Units of work.  A unit is a packet of 64 targets and a range of walk
// positions.  With many packets a unit is a whole walk.  With few packets
// (small N) the walk is split below its first record, which every target
// opens: each subtree of that record is a unit of its own, the units write
// FP64 partial sums, and a reduction adds a packet's partials in subtree
// order.  Both forms give every target the same interaction list.  The
// subtrees' first records are marked as boundaries (kind flag, crit^2 =
// +inf, their own link kept aside), so the straight-line loop leaves at
// them as at any rare record: a unit ends at the next boundary, and a walk
// passing one visits it with the kept link.
//
// Hot path.  Internal nodes, single-particle leaves and zero-mass nodes
// share one straight-line path: "accept if d^2 > crit^2", with crit^2 = -1
// for leaves and zero-mass nodes (always accepted; a zero mass adds exactly
// zero, and with positive softening so does the target itself).  A node's
// record is one 32-byte {node, link} pair fetched with a single scalar load
// at the top of the iteration.  Only multi-particle leaves, zero softening
// and error records take the slow path.  Single-child chains are collapsed
// into their lowest member when the layout is built, which removes visits
// without changing any lane's interaction list.
//
// Accumulation.  Each lane adds its interactions into FP32 partial sums and
// folds them into a Kahan-compensated sum after every kFold visits of its
// packet's walk, so the compensated total carries the rounding of at most
// kFold terms per fold, and a target's result does not depend on the
// schedule.
#include "walk_common.hpp"

#include <algorithm>

namespace nbh {

namespace {

constexpr int kFold = 64;
constexpr int kBoundary = kMixedBoundary;

// One position of the mixed walk: the node (fixed-point COM, m / m_max) and
// its link (crit^2 + eps^2, skip, kind | count << 8, first particle).
struct alignas(32) WalkRec {
    int4 node;
    int4 link;
};

// Kahan-compensated FP32 accumulator of one component.
struct Kahan {
    float s = 0.0f;
    float c = 0.0f;
    __device__ __forceinline__ void add(float v)
    {
        const float y = v - c;
        const float t = s + y;
        c = (t - s) - y;
        s = t;
    }
    __device__ __forceinline__ double value() const
    {
        return static_cast<double>(s) - static_cast<double>(c);
    }
};

struct Acc {
    float px = 0.0f, py = 0.0f, pz = 0.0f;
    Kahan kx, ky, kz;
    __device__ __forceinline__ void fold()
    {
        kx.add(px);
        ky.add(py);
        kz.add(pz);
        px = py = pz = 0.0f;
    }
};

__device__ __forceinline__ void separation(const int4& s, const int4& t,
                                           float& dx, float& dy, float& dz,
                                           float& d2)
{
    dx = static_cast<float>(s.x - t.x);
    dy = static_cast<float>(s.y - t.y);
    dz = static_cast<float>(s.z - t.z);
    d2 = fmaf(dz, dz, fmaf(dx, dx, dy * dy));
}

// Separation and softened squared distance d^2 + eps^2, the r^2 of the
// interaction, which the MAC compares with the record's crit^2 + eps^2.
__device__ __forceinline__ void separation_soft(const int4& s, const int4& t,
                                                float eps2, float& dx, float& dy,
                                                float& dz, float& r2)
{
    dx = static_cast<float>(s.x - t.x);
    dy = static_cast<float>(s.y - t.y);
    dz = static_cast<float>(s.z - t.z);
    r2 = fmaf(dz, dz, fmaf(dx, dx, fmaf(dy, dy, eps2)));
}

// Softened FP32 monopole added to (x, y, z); `f` is forced to zero where
// `use` is false.
__device__ __forceinline__ void interact(float dx, float dy, float dz, float d2,
                                         float gm, float eps2, bool use,
                                         float& x, float& y, float& z)
{
    const float r2 = d2 + eps2;
    const float inv = __builtin_amdgcn_rsqf(r2);
    float f = gm * (inv * inv * inv);
    f = use ? f : 0.0f;
    x = fmaf(f, dx, x);
    y = fmaf(f, dy, y);
    z = fmaf(f, dz, z);
}

constexpr int kIdle = 0x7FFFFFFF;

// One packet's walk: 64 targets consecutive in the target order (the
// Hilbert curve through the octree), one per lane.  A lane's
// `next` is the first position it still has to see.  At position i a lane
// is active when next <= i.  Accepting (or finishing a leaf) sets
// next = skip; opening leaves next <= i, so the lane is active again at
// i + 1.  Lanes waiting for a later position always wait for skip[i] or
// beyond (subtrees nest), hence the packet moves to i + 1 when some lane
// opened and to skip[i] otherwise.
struct Walker {
    int4 tq;       // target (fixed point)
    int t;         // sorted index of the target
    bool valid;    // t < n
    int next;
    int code;      // traversal error of this lane, 0 if none
    int i;         // current position (wave-uniform)
    int end;       // end of the unit's positions (the walk length if whole)
    int unit;      // the unit walked (split walks: its partial-sum slot)
    int pk;        // the packet walked (whole walks: its visit count is kept)
    // Visits of the packet's walk (wave-uniform): those since the last fold
    // in the bits below kFold, the folds from bit 2 * kFold up.
    unsigned cnt;
    Acc a;
    unsigned long long npp, npn, visits;

    // Starts the walk of `packet`: targets order[packet * 64 + lane] (sorted
    // indices along the Hilbert curve), or the Morton-consecutive ones when
    // there is no order.  The interaction and visit counts go on
    // accumulating over the packets a slot walks.  The partial sums are
    // folded after every kFold-th visit of the packet's own walk, so the
    // rounding of a target's sum does not depend on which wavefront or slot
    // walked it, or with which other packet.
    __device__ __forceinline__ void start(const int4* __restrict__ rec,
                                          const uint32_t* __restrict__ order,
                                          int packet, int lane, int n)
    {
        const int k = packet * 64 + lane;
        valid = k < n;
        t = k;
        code = 0;
        if (valid && order) {
            const uint32_t o = order[k];
            if (o < static_cast<uint32_t>(n)) {
                t = static_cast<int>(o);
            } else {
                valid = false;
                code = 3;
            }
        }
        tq = valid ? rec[t] : make_int4(0, 0, 0, 0);
        next = valid ? 0 : kIdle;
        i = 0;
        cnt = 0;
        a = Acc();
    }
};

// Lane masks of the two comparisons of a visit, written by the compare
// instructions straight into a scalar register pair (exec is full here).
__device__ __forceinline__ unsigned long long mask_le(int next, int i)
{
    unsigned long long m;
    asm("v_cmp_ge_i32_e64 %0, %1, %2" : "=s"(m) : "s"(i), "v"(next));
    return m;
}

__device__ __forceinline__ unsigned long long mask_gt(float d2, float crit2)
{
    unsigned long long m;
    asm("v_cmp_lt_f32_e64 %0, %1, %2" : "=s"(m) : "s"(crit2), "v"(d2));
    return m;
}

// Per lane: m's bit of the lane ? b : a.
__device__ __forceinline__ float select_f(unsigned long long m, float a, float b)
{
    float r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(m));
    return r;
}

__device__ __forceinline__ int select_i(unsigned long long m, int a, int b)
{
    int r;
    asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(a), "v"(b), "s"(m));
    return r;
}

// Adds the monopole f = gm * inv3 (per lane) times (dx, dy, dz) to
// (x, y, z) in the lanes of m only: the statement narrows exec to m for the
// multiply and the three fused adds and restores it, so the other lanes
// need no zeroing select.
__device__ __forceinline__ void add_in_lanes(unsigned long long m, float gm,
                                             float inv3, float dx, float dy,
                                             float dz, float& x, float& y,
                                             float& z)
{
    unsigned long long saved;
    float f;
    asm volatile("s_and_saveexec_b64 %[sv], %[m]\n\t"
                 "v_mul_f32_e32 %[f], %[gm], %[i3]\n\t"
                 "v_fmac_f32_e32 %[x], %[f], %[dx]\n\t"
                 "v_fmac_f32_e32 %[y], %[f], %[dy]\n\t"
                 "v_fmac_f32_e32 %[z], %[f], %[dz]\n\t"
                 "s_mov_b64 exec, %[sv]"
                 : [x] "+v"(x), [y] "+v"(y), [z] "+v"(z), [sv] "=&s"(saved),
                   [f] "=&v"(f)
                 : [m] "s"(m), [gm] "s"(gm), [i3] "v"(inv3), [dx] "v"(dx),
                   [dy] "v"(dy), [dz] "v"(dz)
                 : "scc"); // s_and_saveexec sets SCC
}

__device__ __forceinline__ bool lane_bit(unsigned long long m)
{
    return ((m >> (threadIdx.x & 63)) & 1ull) != 0ull;
}

// Ends a visit: folds the partial sums after the packet's kFold-th visit
// since the last fold (`cnt` was incremented for this visit); adding kFold
// then carries the kFold bit into the fold count.
__device__ __forceinline__ void end_visit(Walker& w)
{
    if ((w.cnt & kFold) != 0u) {
        w.a.fold();
        w.cnt += kFold;
    }
}

// The visits of a packet's walk from its visit counter.
__device__ __forceinline__ unsigned packet_visits(unsigned cnt)
{
    return (cnt / (2 * kFold)) * kFold + (cnt & (kFold - 1));
}

// The straight-line part of a visit, for records of every kind: the MAC
// and, when some lane accepts, the monopole and the lanes' next position.
// Records of kinds other than internal carry crit^2 = +inf and no lane
// accepts them here.  Returns the active and accepting lane masks; they are
// ballots of the two comparisons themselves, so they stay the comparisons'
// lane masks instead of a bool that is materialised per lane and compared
// again.
// The record's skip link read through the vector memory path: the value
// arrives in a vector register, where the lanes' next-position select can
// take it directly.  (The select can read only one scalar operand, its lane
// mask.)  A wavefront-scope relaxed load is a plain vector load.
__device__ __forceinline__ int skip_vgpr(const WalkRec* r)
{
    return __hip_atomic_load(&r->link.y, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_WAVEFRONT);
}

template <bool CONTAINS, bool COUNT, bool BRANCH = true>
__device__ __forceinline__ void visit_mac(Walker& w, const WalkRec& r, int skipv,
                                          const int4* __restrict__ wcenter,
                                          float eps2, unsigned long long& actm,
                                          unsigned long long& accm)
{
    const int4 lk = r.link;
    const int4 nd = r.node;
    float dx, dy, dz, r2;
    separation_soft(nd, w.tq, eps2, dx, dy, dz, r2);
    actm = mask_le(w.next, w.i);
    accm = actm & mask_gt(r2, __int_as_float(lk.x));
    if (CONTAINS && accm != 0ull) {
        bool acc = lane_bit(accm);
        if (acc) {
            const int4 c = wcenter[w.i];
            acc = !(abs(w.tq.x - c.x) <= c.w && abs(w.tq.y - c.y) <= c.w &&
                    abs(w.tq.z - c.z) <= c.w);
        }
        accm = __ballot(acc);
    }
    if (!BRANCH || accm != 0ull) {
        const float inv = __builtin_amdgcn_rsqf(r2);
        add_in_lanes(accm, __int_as_float(nd.w), inv * inv * inv, dx, dy, dz,
                     w.a.px, w.a.py, w.a.pz);
        if (COUNT && lane_bit(accm) && __int_as_float(nd.w) != 0.0f) {
            if ((lk.z >> 8) == 1) {
                if (lk.w != w.t)
                    ++w.npp;
            } else {
                ++w.npn;
            }
        }
        w.next = select_i(accm, w.next, skipv);
    }
}

typedef int int8v __attribute__((ext_vector_type(8)));

__device__ __forceinline__ WalkRec as_record(const int8v& v)
{
    WalkRec r;
    r.node = make_int4(v[0], v[1], v[2], v[3]);
    r.link = make_int4(v[4], v[5], v[6], v[7]);
    return r;
}

// Loads the records of both packets with one wait for the two: the second
// load issues before the first has returned.  The wait is inside the same
// statement, so no output register is written after the statement ends.
__device__ __forceinline__ void load_pair(const WalkRec* pa, const WalkRec* pb,
                                          WalkRec& ra, WalkRec& rb)
{
    int8v a, b;
    asm volatile("s_load_dwordx8 %0, %2, 0x0\n\t"
                 "s_load_dwordx8 %1, %3, 0x0\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&s"(a), "=&s"(b)
                 : "s"(pa), "s"(pb)
                 : "memory");
    ra = as_record(a);
    rb = as_record(b);
}

// The same with 32-bit byte offsets from the record array's base, used
// while the records span less than 2 GiB: each address is then one shift.
__device__ __forceinline__ void load_pair_off(const WalkRec* base, unsigned oa,
                                              unsigned ob, WalkRec& ra, WalkRec& rb)
{
    int8v a, b;
    asm volatile("s_load_dwordx8 %0, %2, %3\n\t"
                 "s_load_dwordx8 %1, %2, %4\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&s"(a), "=&s"(b)
                 : "s"(base), "s"(oa), "s"(ob)
                 : "memory");
    ra = as_record(a);
    rb = as_record(b);
}

// Buffer resource over the record array (raw, byte-addressed; word 3 is
// the gfx9 default format).
__device__ __forceinline__ __amdgpu_buffer_rsrc_t records_rsrc(const WalkRec* base,
                                                              unsigned bytes)
{
    return __builtin_amdgcn_make_buffer_rsrc(const_cast<WalkRec*>(base), 0,
                                             static_cast<int>(bytes), 0x00020000);
}

// The skip link of the record at byte offset `o`, through the vector memory
// path into a vector register (see skip_vgpr).
__device__ __forceinline__ int skip_vgpr_off(__amdgpu_buffer_rsrc_t rsrc, unsigned o)
{
    return __builtin_amdgcn_raw_buffer_load_b32(rsrc, 20, static_cast<int>(o), 0);
}

// Next position after an internal-node visit: the active lanes that opened
// the node are still at or before i (skip > i, so an accepting lane has
// left).
__device__ __forceinline__ void advance_internal(Walker& w, const WalkRec& r,
                                                 unsigned long long actm,
                                                 unsigned long long accm)
{
    w.i = (accm != actm) ? w.i + 1 : r.link.y;
}

// A record of another kind below the walk length, after visit_mac: a
// multi-particle leaf (or a single particle without softening), a
// zero-mass node to skip, or an error record.
template <bool SOFT, bool COUNT>
__device__ __forceinline__ void visit_rare(Walker& w, const WalkRec& r,
                                        const int4* __restrict__ rec,
                                        float eps2)
{
    const int4 lk = r.link;
    const int skip = lk.y;
    const int kind = lk.z & 0xFF;
    const bool active = w.next <= w.i;
    if (kind == KIND_LEAF) {
        const int count = lk.z >> 8;
        const int first = lk.w;
        for (int k = first; k < first + count; ++k) {
            const int4 s = rec[k];
            float sx, sy, sz, s2;
            separation(s, w.tq, sx, sy, sz, s2);
            const float gm = __int_as_float(s.w);
            bool use = active && k != w.t && gm != 0.0f;
            if (!SOFT && use && s2 == 0.0f) {
                w.code = 6;
                use = false;
            }
            interact(sx, sy, sz, s2, gm, eps2, use, w.a.px, w.a.py, w.a.pz);
            if (COUNT && use)
                ++w.npp;
        }
        w.next = active ? skip : w.next;
    } else if (kind == KIND_SKIP) {
        w.next = active ? skip : w.next;
    } else if (active) {
        w.code = (kind == KIND_CHILDLESS) ? 5 : kind - KIND_ERROR;
    }
    if (w.code != 0)
        w.next = kIdle;
    w.i = (__ballot(w.next <= w.i) != 0ull) ? w.i + 1 : skip;
}

// The visit of a boundary record at w.i that does not end the unit: the
// record with its own link (kept in the split table) takes the usual path.
// The straight-line loop has already evaluated the marked record, whose
// crit^2 = +inf accepts nothing and changes no lane.
template <bool CONTAINS, bool SOFT, bool COUNT>
__device__ __forceinline__ void visit_boundary(Walker& w, const WalkRec& r,
                                               const int* __restrict__ split,
                                               const int4* __restrict__ wcenter,
                                               const int4* __restrict__ rec,
                                               float eps2)
{
    const int subtrees = split[0];
    int k = 0;
    while (k < subtrees && split[1 + 4 * k] != w.i)
        ++k;
    if (k == subtrees) { // not a subtree's first record: inconsistent walk
        w.code = 4;
        w.next = kIdle;
        w.i = r.link.y;
        return;
    }
    WalkRec own = r;
    own.link.x = split[3 + 4 * k];
    own.link.z = split[4 + 4 * k];
    if ((own.link.z & 0xFF) == KIND_INTERNAL) {
        unsigned long long actm, accm;
        visit_mac<CONTAINS, COUNT>(w, own, own.link.y, wcenter, eps2, actm, accm);
        advance_internal(w, own, actm, accm);
    } else {
        visit_rare<SOFT, COUNT>(w, own, rec, eps2);
    }
}

// Walks one packet from its current position to the end.  Position wcount
// holds a record of a non-internal kind, which ends the internal-node loop.
template <bool CONTAINS, bool SOFT, bool COUNT, bool SPLIT>
__device__ __forceinline__ void walk_rest(Walker& w, const WalkRec* __restrict__ wrec,
                                          const int4* __restrict__ wcenter,
                                          const int4* __restrict__ rec,
                                          const int* __restrict__ split,
                                          int wcount, float eps2)
{
    for (;;) {
        WalkRec r;
        unsigned long long actm, accm;
        for (;;) {
            r = wrec[w.i];
            visit_mac<CONTAINS, COUNT>(w, r, r.link.y, wcenter, eps2, actm, accm);
            if ((r.link.z & 0xFF) != KIND_INTERNAL)
                break;
            if (COUNT)
                ++w.visits;
            ++w.cnt;
            end_visit(w);
            advance_internal(w, r, actm, accm);
        }
        if (SPLIT && (r.link.z & kBoundary) != 0) {
            if (w.i == w.end)
                break;
            if (COUNT)
                ++w.visits;
            visit_boundary<CONTAINS, SOFT, COUNT>(w, r, split, wcenter, rec, eps2);
            ++w.cnt;
            end_visit(w);
            continue;
        }
        if (w.i >= wcount)
            break;
        if (COUNT)
            ++w.visits;
        visit_rare<SOFT, COUNT>(w, r, rec, eps2);
        ++w.cnt;
        end_visit(w);
    }
}

// A whole walk writes the target's acceleration; a split unit (part given)
// writes its FP64 partial sums at part[(unit * 64 + lane) * 3].
// Each unit's visit count goes to pcost for the next schedule (a counting
// walk leaves the costs).
template <bool COUNT, bool SPLIT>
__device__ __forceinline__ void finish(Walker& w, const uint32_t* __restrict__ idx,
                                       int n, double out_scale,
                                       double* __restrict__ ax_out,
                                       double* __restrict__ ay_out,
                                       double* __restrict__ az_out,
                                       double* __restrict__ part,
                                       uint32_t* __restrict__ pcost, int* err)
{
    if (!COUNT && (threadIdx.x & 63) == 0)
        pcost[SPLIT ? w.unit : w.pk] = packet_visits(w.cnt);
    w.a.fold();
    if (w.code != 0) {
        raise_error(err, w.code);
        return;
    }
    if (COUNT || !w.valid)
        return;
    if (SPLIT) {
        double* q = part + (static_cast<size_t>(w.unit) * 64 + (threadIdx.x & 63)) * 3;
        q[0] = w.a.kx.value();
        q[1] = w.a.ky.value();
        q[2] = w.a.kz.value();
        return;
    }
    const uint32_t o = idx[w.t];
    if (o >= static_cast<uint32_t>(n)) {
        raise_error(err, 3);
        return;
    }
    ax_out[o] = out_scale * w.a.kx.value();
    ay_out[o] = out_scale * w.a.ky.value();
    az_out[o] = out_scale * w.a.kz.value();
}

// Persistent wavefronts, each with two unit slots walked in lock-step: one
// iteration loads both slots' records, then does both visits, so the two
// record loads (mostly served by L2: the walk's working set exceeds the
// scalar cache) are in flight together and each visit's VALU work overlaps
// the other slot's load.  A slot whose unit ends writes its results and
// takes the next unit from the queue, so the slots stay paired until the
// queue is empty; the remaining slot then finishes alone.  Every wavefront
// leaves once the queue is empty and its slots are done.
// SPLIT (split and part given): unit u walks subtree u % s of packet u / s,
// with split = {s, then per subtree {first, end, own crit^2, own kind
// word}}; only these instantiations handle boundary records.  Queue
// position q holds unit perm[q]; each unit's visit count goes to pcost for
// the next schedule (counting walks leave pcost).
template <bool CONTAINS, bool SOFT, bool COUNT, bool WIDE, bool SPLIT>
__global__ __launch_bounds__(kWalkThreads) void walk_mixed_kernel(
    const WalkRec* __restrict__ wrec, unsigned wbytes,
    const int4* __restrict__ wcenter, const int* __restrict__ d_wcount,
    const int4* __restrict__ rec, const uint32_t* __restrict__ idx,
    const uint32_t* __restrict__ order, int n,
    int* __restrict__ queue, float eps2, double out_scale,
    double* __restrict__ ax_out,
    double* __restrict__ ay_out, double* __restrict__ az_out,
    const int* __restrict__ split, double* __restrict__ part,
    const uint32_t* __restrict__ perm, uint32_t* __restrict__ pcost, int* err,
    unsigned long long* __restrict__ counters)
{
    const int lane = threadIdx.x & 63;
    const int packets = (n + 63) / 64;
    const int subtrees = SPLIT ? split[0] : 1;
    const int units = packets * subtrees;
    Queue q{queue, units, (units + 7) / 8, xcd_id(), 0};
    const int ua = q.take();
    if (ua < 0)
        return;
    const int wcount = *d_wcount;
    const __amdgpu_buffer_rsrc_t rsrc = records_rsrc(wrec, WIDE ? 0u : wbytes);
    // The schedule's pointers and packet numbers live in vector registers:
    // they are used once per packet, and the scalar registers of the walk
    // are spent.
    const uint32_t* permv = perm;
    uint32_t* pcostv = pcost;
    asm volatile("" : "+v"(permv), "+v"(pcostv));
    auto begin = [&](Walker& w, int q) {
        if (SPLIT) {
            // perm is a permutation by construction; the bound keeps
            // every access inside the units whatever it holds
            const int u = __builtin_amdgcn_readfirstlane(
                static_cast<int>(min(permv[q], static_cast<uint32_t>(units - 1))));
            const int p = u / subtrees;
            const int k = u - p * subtrees;
            w.start(rec, order, p, lane, n);
            const int first = split[1 + 4 * k];
            w.i = first;
            w.next = w.valid ? first : kIdle;
            w.end = split[2 + 4 * k];
            w.unit = u;
        } else {
            // perm is a permutation by construction; the bound keeps
            // every access inside the targets whatever it holds
            int p = static_cast<int>(min(permv[q], static_cast<uint32_t>(packets)));
            asm volatile("" : "+v"(p));
            w.start(rec, order, p, lane, n);
            w.pk = p;
        }
    };
    Walker A, B;
    A.npp = A.npn = A.visits = 0;
    B.npp = B.npn = B.visits = 0;
    begin(A, ua);
    const int ub = q.take();
    bool run_a = true, run_b = ub >= 0;
    if (run_b)
        begin(B, ub);
    while (run_a && run_b) {
        // Lock-step iterations until a slot reaches a rare record kind, its
        // end, or its fold point; the loop body has no branch besides its
        // exit, so both record loads issue together.
        WalkRec ra, rb;
        unsigned long long act_a, acc_a, act_b, acc_b;
        static_assert((kFold & (kFold - 1)) == 0, "kFold must be a power of two");
        for (;;) {
            int sa, sb;
            if (WIDE) {
                sa = skip_vgpr(wrec + A.i);
                sb = skip_vgpr(wrec + B.i);
                load_pair(wrec + A.i, wrec + B.i, ra, rb);
            } else {
                const unsigned oa = static_cast<unsigned>(A.i) << 5;
                const unsigned ob = static_cast<unsigned>(B.i) << 5;
                sa = skip_vgpr_off(rsrc, oa);
                sb = skip_vgpr_off(rsrc, ob);
                load_pair_off(wrec, oa, ob, ra, rb);
            }
            visit_mac<CONTAINS, COUNT, false>(A, ra, sa, wcenter, eps2, act_a, acc_a);
            visit_mac<CONTAINS, COUNT, false>(B, rb, sb, wcenter, eps2, act_b, acc_b);
            ++A.cnt;
            ++B.cnt;
            // One exit test in integer arithmetic (cnt < kFold, a power of
            // two, until a fold point).
            if ((((ra.link.z | rb.link.z) & 0xFF) | ((A.cnt | B.cnt) & kFold)) != 0)
                break;
            if (COUNT) {
                ++A.visits;
                ++B.visits;
            }
            advance_internal(A, ra, act_a, acc_a);
            advance_internal(B, rb, act_b, acc_b);
        }
        // Finish both visits of the last iteration; a slot at the end of
        // its walk moves on to the next unit.
        if ((ra.link.z & 0xFF) == KIND_INTERNAL) {
            if (COUNT)
                ++A.visits;
            end_visit(A);
            advance_internal(A, ra, act_a, acc_a);
        } else if (SPLIT && A.i != A.end && (ra.link.z & kBoundary) != 0) {
            if (COUNT)
                ++A.visits;
            visit_boundary<CONTAINS, SOFT, COUNT>(A, ra, split, wcenter, rec, eps2);
            end_visit(A);
        } else if (A.i < (SPLIT ? A.end : wcount)) {
            if (COUNT)
                ++A.visits;
            visit_rare<SOFT, COUNT>(A, ra, rec, eps2);
            end_visit(A);
        } else {
            finish<COUNT, SPLIT>(A, idx, n, out_scale, ax_out, ay_out, az_out, part, pcostv, err);
            const int u = q.take();
            if (u >= 0)
                begin(A, u);
            else
                run_a = false;
        }
        if ((rb.link.z & 0xFF) == KIND_INTERNAL) {
            if (COUNT)
                ++B.visits;
            end_visit(B);
            advance_internal(B, rb, act_b, acc_b);
        } else if (SPLIT && B.i != B.end && (rb.link.z & kBoundary) != 0) {
            if (COUNT)
                ++B.visits;
            visit_boundary<CONTAINS, SOFT, COUNT>(B, rb, split, wcenter, rec, eps2);
            end_visit(B);
        } else if (B.i < (SPLIT ? B.end : wcount)) {
            if (COUNT)
                ++B.visits;
            visit_rare<SOFT, COUNT>(B, rb, rec, eps2);
            end_visit(B);
        } else {
            finish<COUNT, SPLIT>(B, idx, n, out_scale, ax_out, ay_out, az_out, part, pcostv, err);
            const int u = q.take();
            if (u >= 0)
                begin(B, u);
            else
                run_b = false;
        }
    }
    if (run_a) {
        walk_rest<CONTAINS, SOFT, COUNT, SPLIT>(A, wrec, wcenter, rec, split, wcount, eps2);
        finish<COUNT, SPLIT>(A, idx, n, out_scale, ax_out, ay_out, az_out, part, pcostv, err);
    }
    if (run_b) {
        walk_rest<CONTAINS, SOFT, COUNT, SPLIT>(B, wrec, wcenter, rec, split, wcount, eps2);
        finish<COUNT, SPLIT>(B, idx, n, out_scale, ax_out, ay_out, az_out, part, pcostv, err);
    }
    if (COUNT) {
        unsigned long long npp = A.npp + B.npp, npn = A.npn + B.npn;
        for (int o = 32; o > 0; o >>= 1) {
            npp += __shfl_xor(npp, o);
            npn += __shfl_xor(npn, o);
        }
        if (lane == 0) {
            atomicAdd(&counters[0], npp);
            atomicAdd(&counters[1], npn);
            atomicAdd(&counters[2], A.visits + B.visits);
        }
    }
}

// The split of the walk below its first record: the subtrees that follow
// position 0, read off the skip links.  Every target opens position 0 (a
// cell holding the target is never accepted: for theta < 2/sqrt(3) its
// opening radius exceeds every distance inside it, and above that the
// containment test rejects it), so each subtree is walked by every lane
// from its first position.  The subtrees' first records are marked as
// boundaries (kBoundary in the kind byte, crit^2 = +inf) and their own
// crit^2 and kind word kept in the table; a record that is already marked
// was marked by an earlier split of the same layout, whose table entry
// holds its own values.  Without such a split (a one-record walk, or links
// that do not tile the walk) the single unit is the whole walk and no
// record is marked.
// split = {s, then per subtree {first, end, own crit^2, own kind word}}.
__global__ void walk_split_kernel(WalkRec* __restrict__ wrec,
                                  const int* __restrict__ d_wcount,
                                  int* __restrict__ split)
{
    if (threadIdx.x != 0 || blockIdx.x != 0)
        return;
    const int wc = *d_wcount;
    bool ok = wc > 1 && (wrec[0].link.z & 0xFF) == KIND_INTERNAL &&
              __int_as_float(wrec[0].link.x) >= 0.0f;
    int k = 0;
    int p = 1;
    while (ok && p < wc) {
        const int e = wrec[p].link.y;
        if (k == kWalkSplitMax || e <= p || e > wc) {
            ok = false;
            break;
        }
        p = e;
        ++k;
    }
    if (!ok || k == 0) {
        split[0] = 1;
        split[1] = 0;
        split[2] = wc;
        return;
    }
    p = 1;
    for (int j = 0; j < k; ++j) {
        int4& lk = wrec[p].link;
        const int e = lk.y;
        if ((lk.z & kBoundary) == 0) {
            split[3 + 4 * j] = lk.x;
            split[4 + 4 * j] = lk.z;
            lk.x = 0x7F800000;
            lk.z |= kBoundary;
        }
        split[1 + 4 * j] = p;
        split[2 + 4 * j] = e;
        p = e;
    }
    split[0] = k;
}

// Target accelerations of a split walk: each target's unit partials added
// in subtree order, scaled to physical units and written to the target's
// original index.
__global__ __launch_bounds__(256) void walk_split_sum_kernel(
    const double* __restrict__ part, const int* __restrict__ split,
    const uint32_t* __restrict__ order, const uint32_t* __restrict__ idx, int n,
    double out_scale, double* __restrict__ ax_out, double* __restrict__ ay_out,
    double* __restrict__ az_out)
{
    const int k = static_cast<int>(blockIdx.x * blockDim.x + threadIdx.x);
    if (k >= n)
        return;
    const int subtrees = split[0];
    const int packet = k >> 6;
    const int lane = k & 63;
    double sx = 0.0, sy = 0.0, sz = 0.0;
    for (int j = 0; j < subtrees; ++j) {
        const double* q =
            part + ((static_cast<size_t>(packet) * subtrees + j) * 64 + lane) * 3;
        sx += q[0];
        sy += q[1];
        sz += q[2];
    }
    // out-of-range entries were reported by the walk (error 3)
    const uint32_t t = order ? order[k] : static_cast<uint32_t>(k);
    if (t >= static_cast<uint32_t>(n))
        return;
    const uint32_t o = idx[t];
    if (o >= static_cast<uint32_t>(n))
        return;
    ax_out[o] = out_scale * sx;
    ay_out[o] = out_scale * sy;
    az_out[o] = out_scale * sz;
}

} // namespace

bool walk_mixed_splits(int n)
{
    return n > 0 && (n + 63) / 64 <= kWalkSplitPackets;
}

void launch_walk_split(hipStream_t s, int4* wpairs, const int* d_wcount, int* split)
{
    walk_split_kernel<<<1, 64, 0, s>>>(reinterpret_cast<WalkRec*>(wpairs), d_wcount, split);
    NBH_HIP(hipGetLastError());
}

void launch_walk_mixed(hipStream_t s, int4* wpairs, size_t wcap,
                       const int4* wcenter, const int* d_wcount,
                       const int4* rec, const uint32_t* idx,
                       const uint32_t* order, int n,
                       float eps2, double out_scale, bool contains, bool soft,
                       double* ax, double* ay, double* az, int* queue,
                       int* split, double* part, uint32_t* perm, uint32_t* pcost,
                       int* err, unsigned long long* counters)
{
    // 32-bit record offsets while the record array spans less than 2 GiB
    const size_t wbytes = wcap * sizeof(WalkRec);
    const bool wide = wbytes > 0x7FFFFFFFull;
    const bool splits = split != nullptr;
    if (splits && (wide || !part || !walk_mixed_splits(n)))
        fail(NBH_LOGIC_ERROR, "a split traversal needs fewer packets and a partial-sum buffer");
    queue_clear_kernel<<<1, 64, 0, s>>>(queue);
    NBH_HIP(hipGetLastError());
    const long long packets = (static_cast<long long>(n) + 63) / 64;
    auto* wrec = reinterpret_cast<WalkRec*>(wpairs);
    if (splits) {
        walk_split_kernel<<<1, 64, 0, s>>>(wrec, d_wcount, split);
        NBH_HIP(hipGetLastError());
    }
    // Whole walks: one pair of packets per wavefront, 8 resident per SIMD.
    // Split walks: up to 4,096 packets at least four subtree units per
    // wavefront and at most 2 wavefronts per SIMD, whose short units run
    // faster with few wavefronts (SI blob, cost-ordered units, step time:
    // 250K particles 1.42 ms at 2 per SIMD against 1.50 at 3 and 1.58 at 4;
    // 100K 1.21 ms, 30K 0.97 ms); beyond, at least 16 units per wavefront
    // and at most 8 per SIMD (Plummer traversal: 512K 1.41 ms at 4 per SIMD
    // against 1.46 at 8, 1M 2.28 ms at 8 against 2.38 at 6 and 2.97 at 4;
    // split against whole packets: 512K 1.41 against 1.77 ms, 1M 2.28
    // against 2.73 ms; 1M uniform and clumps equal).
    const int nb = splits ? walk_persistent_blocks(packets * kWalkSplitMax,
                                                   packets > 4096 ? 16 : 4,
                                                   packets > 4096 ? 8 : 2)
                          : walk_persistent_blocks(packets, 2, kWalkMixedWavesPerSimd);
    if (nb == 0)
        return;
    // The queue hands out the units in the order built from their costs.
    // Whole walks: each range ends with as many of its cheapest packets as
    // the range has slots (two per wavefront), the last ones started, the
    // others in curve order.  With no more packets than slots every packet
    // starts at once and the curve order stays (reordering then only
    // scatters the packets that the slots of a wavefront and of an XCD walk
    // together: 1M uniform particles, 16,384 packets, took 3.5 % longer).
    // Split walks, whose units differ most in cost (the subtree holding the
    // packet's own targets carries most of its near field) and whose records
    // fit the caches: every unit in decreasing cost (SI blob, 100K: 1.30 ->
    // 1.21 ms per step; 250K: 1.68 -> 1.58 ms).
    if (!perm || !pcost)
        fail(NBH_LOGIC_ERROR, "a traversal needs its unit schedule buffers");
    {
        const long long slots = static_cast<long long>(nb) * 4 * 2;
        const int tail = splits ? static_cast<int>(packets * kWalkSplitMax)
                                : (packets > slots ? static_cast<int>((slots + 7) / 8) : 0);
        packet_schedule_kernel<<<8, kScheduleThreads, 0, s>>>(
            pcost, static_cast<int>(packets), splits ? split : nullptr, tail, perm);
        NBH_HIP(hipGetLastError());
    }
#define NBH_ARGS                                                                 \
    wrec, static_cast<unsigned>(wide ? 0 : wbytes), wcenter, d_wcount, rec, idx, \
        order, n, queue, eps2, out_scale, ax, ay, az, splits ? split : nullptr, \
        splits ? part : nullptr, perm, pcost, err, \
        counters
#define NBH_LAUNCH(C, S, K)                                                      \
    do {                                                                         \
        if (splits)                                                              \
            walk_mixed_kernel<C, S, K, false, true><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS); \
        else if (wide)                                                           \
            walk_mixed_kernel<C, S, K, true, false><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS); \
        else                                                                     \
            walk_mixed_kernel<C, S, K, false, false><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS); \
    } while (0)
    const bool cnt = counters != nullptr;
    if (contains) {
        if (soft) {
            if (cnt) NBH_LAUNCH(true, true, true); else NBH_LAUNCH(true, true, false);
        } else {
            if (cnt) NBH_LAUNCH(true, false, true); else NBH_LAUNCH(true, false, false);
        }
    } else {
        if (soft) {
            if (cnt) NBH_LAUNCH(false, true, true); else NBH_LAUNCH(false, true, false);
        } else {
            if (cnt) NBH_LAUNCH(false, false, true); else NBH_LAUNCH(false, false, false);
        }
    }
#undef NBH_LAUNCH
#undef NBH_ARGS
    NBH_HIP(hipGetLastError());
    if (splits && !cnt) {
        walk_split_sum_kernel<<<blocks_for(static_cast<size_t>(n), 256), 256, 0, s>>>(
            part, split, order, idx, n, out_scale, ax, ay, az);
        NBH_HIP(hipGetLastError());
    }
}

} // namespace nbh
