// walk_fp64.hip -- FP64 packet traversal (the reference arithmetic, bit-exact
// with the oracle) and the FP64 direct sum used as accuracy reference.
#include "fp64_arith.hpp"
#include "walk_common.hpp"

namespace nbh {

namespace {

constexpr int kIdle64 = 0x7FFFFFFF;

// One packet of the FP64 walk: 64 targets consecutive in the target order,
// one per lane.  A lane's `next` is the first position it still has to see;
// at position i it is active when next <= i.  Taking a record (a monopole,
// a leaf or a zero mass) sets next = skip; opening leaves next <= i, so the
// lane is active again at i + 1.  The packet moves to i + 1 when some lane
// opened and to skip[i] otherwise; every lane sees its own per-target list
// in the reference's order (octants 7 -> 0), whatever the packet.
struct Walker64 {
    double tx, ty, tz;
    int t;       // sorted index of the target
    bool valid;  // the lane holds a target
    int next;
    int code;    // traversal error of this lane, 0 if none
    int i;       // current position (wave-uniform)
    double ax, ay, az;
    unsigned long long npp, npn, visits;

    __device__ __forceinline__ void start(const double4* __restrict__ rec,
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
        tx = ty = tz = 0.0;
        if (valid) {
            const double4 p = rec[t];
            tx = p.x;
            ty = p.y;
            tz = p.z;
        }
        next = valid ? 0 : kIdle64;
        i = 0;
        ax = ay = az = 0.0;
    }
};

// barnes_hut.cu:56-80 in FP64 with the oracle's contraction order, for a
// source of G m = gm at separation (dx, dy, dz), d2 = |d|^2 (unsoftened):
// inv = 1 / sqrt(d2 + eps2) with both operations correctly rounded (the
// reduced sequences in the normal range, the general ones outside it).
// FAST: the caller has established that every r2 of the walk lies in the
// reduced sequences' range (eps^2 >= 2^-600 and 3 side^2 + eps^2 <= 2^600
// for the bounding box's side), so r2 > 0 and no range test is needed.
template <bool FAST>
__device__ __forceinline__ void add_fp64(Walker64& w, double dx, double dy, double dz,
                                         double d2, double gm, double eps2)
{
    const double r2 = d2 + eps2;
    double inv;
    if (FAST) {
        inv = rcp_rn_fast(sqrt_rn_fast(r2));
    } else {
        if (r2 <= 0.0) {
            w.code = 6;
            return;
        }
        if (in_fast_range(r2))
            inv = rcp_rn_fast(sqrt_rn_fast(r2));
        else
            inv = 1.0 / sqrt(r2);
    }
    const double inv3 = inv * inv * inv;
    const double f = gm * inv3;
    w.ax = fma(f, dx, w.ax);
    w.ay = fma(f, dy, w.ay);
    w.az = fma(f, dz, w.az);
}

// The straight-line visit of a record of any kind: the opening test
// d^2 > thr (-inf for one-particle leaves and zero masses, +inf for leaves
// of several particles and error records, which no lane takes here) and,
// when some lane takes the record, its monopole for the lanes that do (a
// leaf skips the target's own particle by index, a zero mass adds nothing)
// and their next position.  Returns whether some active lane did not take
// it (opened it, for an internal node).
template <bool CONTAINS, bool COUNT, bool FAST>
__device__ __forceinline__ bool visit_fast(Walker64& w, const Walk64Rec& r, int skipv,
                                           const double4* __restrict__ wcenter,
                                           double eps2)
{
    const bool act = w.next <= w.i;
    const double dx = r.x - w.tx;
    const double dy = r.y - w.ty;
    const double dz = r.z - w.tz;
    const double d2 = fma(dz, dz, fma(dx, dx, dy * dy));
    const bool mac = d2 > r.thr;
    bool acc = act && mac;
    // lane masks, as ballots of the comparisons themselves (a ballot of
    // their conjunction would be materialised per lane): the lanes that
    // opened the record are the active ones that did not take it
    const unsigned long long actm = __ballot(act);
    unsigned long long accm = actm & __ballot(mac);
    if (CONTAINS && accm != 0ull) {
        if (acc) {
            const double4 c = wcenter[w.i];
            acc = !(fabs(w.tx - c.x) <= c.w && fabs(w.ty - c.y) <= c.w &&
                    fabs(w.tz - c.z) <= c.w);
        }
        accm = __ballot(acc);
    }
    if (accm != 0ull) {
        // a zero mass adds nothing (tested on the record's bits, uniform)
        const bool massive = (__double_as_longlong(r.m) << 1) != 0;
        if (massive && acc && r.first != w.t) {
            add_fp64<FAST>(w, dx, dy, dz, d2, r.gm, eps2);
            if (COUNT) {
                if (r.count == 1)
                    ++w.npp;
                else
                    ++w.npn;
            }
        }
        w.next = acc ? skipv : w.next;
    }
    return actm != accm;
}

// A record of a slow kind below the walk length, after visit_fast: a leaf
// of several particles (exact sum in sorted order, the target's own
// particle skipped), a childless node (the lanes still active did not take
// it: error 5) or an error record.
template <bool COUNT>
__device__ __forceinline__ void visit_rare(Walker64& w, const Walk64Rec& r,
                                           const double4* __restrict__ rec, double eps2,
                                           double G)
{
    const bool active = w.next <= w.i;
    if (r.kind == KIND_LEAF) {
        for (int k = r.first; k < r.first + r.count; ++k) {
            const double4 s = rec[k];
            if (active && k != w.t && s.w != 0.0) {
                const double dx = s.x - w.tx;
                const double dy = s.y - w.ty;
                const double dz = s.z - w.tz;
                add_fp64<false>(w, dx, dy, dz, fma(dz, dz, fma(dx, dx, dy * dy)), G * s.w,
                                eps2);
                if (COUNT)
                    ++w.npp;
            }
        }
        w.next = active ? r.skip : w.next;
    } else if (active) {
        w.code = (r.kind == KIND_CHILDLESS) ? 5 : r.kind - KIND_ERROR;
    }
    if (w.code != 0)
        w.next = kIdle64;
    w.i = (__ballot(w.next <= w.i) != 0ull) ? w.i + 1 : r.skip;
}

typedef int int16v __attribute__((ext_vector_type(16)));

// Loads the records of both packets with one wait for the two (the second
// load issues before the first returns).  The wait is inside the statement,
// so no output register is written after it ends.
__device__ __forceinline__ void load_pair64(const Walk64Rec* pa, const Walk64Rec* pb,
                                            Walk64Rec& ra, Walk64Rec& rb)
{
    int16v a, b;
    asm volatile("s_load_dwordx16 %0, %2, 0x0\n\t"
                 "s_load_dwordx16 %1, %3, 0x0\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&s"(a), "=&s"(b)
                 : "s"(pa), "s"(pb)
                 : "memory");
    ra = __builtin_bit_cast(Walk64Rec, a);
    rb = __builtin_bit_cast(Walk64Rec, b);
}

// The same with 32-bit byte offsets from the record array's base, while the
// records span less than 2 GiB.
__device__ __forceinline__ void load_pair64_off(const Walk64Rec* base, unsigned oa,
                                                unsigned ob, Walk64Rec& ra, Walk64Rec& rb)
{
    int16v a, b;
    asm volatile("s_load_dwordx16 %0, %2, %3\n\t"
                 "s_load_dwordx16 %1, %2, %4\n\t"
                 "s_waitcnt lgkmcnt(0)"
                 : "=&s"(a), "=&s"(b)
                 : "s"(base), "s"(oa), "s"(ob)
                 : "memory");
    ra = __builtin_bit_cast(Walk64Rec, a);
    rb = __builtin_bit_cast(Walk64Rec, b);
}

// A record's skip link through the vector memory path, into a vector
// register where the lanes' next-position select takes it directly.
__device__ __forceinline__ int skip64_vgpr(const Walk64Rec* r)
{
    return __hip_atomic_load(&r->skip, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WAVEFRONT);
}

__device__ __forceinline__ int skip64_vgpr_off(__amdgpu_buffer_rsrc_t rsrc, unsigned o)
{
    return __builtin_amdgcn_raw_buffer_load_b32(rsrc, 40, static_cast<int>(o), 0);
}

// Ends a packet: writes each target's acceleration at its original index.
template <bool COUNT>
__device__ __forceinline__ void finish64(Walker64& w, const uint32_t* __restrict__ idx,
                                         int n, double* __restrict__ ax_out,
                                         double* __restrict__ ay_out,
                                         double* __restrict__ az_out, int* err)
{
    if (w.code != 0) {
        raise_error(err, w.code);
        return;
    }
    if (COUNT || !w.valid)
        return;
    const uint32_t o = idx[w.t];
    if (o >= static_cast<uint32_t>(n)) {
        raise_error(err, 3);
        return;
    }
    ax_out[o] = w.ax;
    ay_out[o] = w.ay;
    az_out[o] = w.az;
}

// One slot walking alone once the queue has no packet for the other: its
// walk to the end, then further packets until the queue is empty.
template <bool CONTAINS, bool COUNT, bool WIDE, bool FAST>
__device__ __forceinline__ void walk_alone(Walker64& W, Queue& q, const Walk64Rec* __restrict__ wrec,
                                        __amdgpu_buffer_rsrc_t rsrc,
                                        const double4* __restrict__ wcenter, int wcount,
                                        const double4* __restrict__ rec,
                                        const uint32_t* __restrict__ idx,
                                        const uint32_t* __restrict__ order, int n, int lane,
                                        double eps2, double G, double* __restrict__ ax_out,
                                        double* __restrict__ ay_out,
                                        double* __restrict__ az_out, int* err)
{
    for (;;) {
        const Walk64Rec r = wrec[W.i];
        const int sv = WIDE ? skip64_vgpr(wrec + W.i)
                            : skip64_vgpr_off(rsrc, static_cast<unsigned>(W.i) << 6);
        const bool open = visit_fast<CONTAINS, COUNT, FAST>(W, r, sv, wcenter, eps2);
        if (r.kind == KIND_INTERNAL) {
            if (COUNT)
                ++W.visits;
            W.i = open ? W.i + 1 : r.skip;
        } else if (W.i < wcount) {
            if (COUNT)
                ++W.visits;
            visit_rare<COUNT>(W, r, rec, eps2, G);
        } else {
            finish64<COUNT>(W, idx, n, ax_out, ay_out, az_out, err);
            const int p = q.take();
            if (p < 0)
                break;
            W.start(rec, order, p, lane, n);
        }
    }
}

// Persistent wavefronts with two packet slots walked in lock-step, fed by
// the per-XCD work queue: one iteration loads both slots' 64-byte records
// under one wait and then does both visits, so each record load overlaps
// the other slot's arithmetic.  A slot whose packet ends writes its results
// and takes the next packet; the remaining slot finishes alone once the
// queue is empty.  Position wcount (and every position a walk can reach
// beyond its records) holds a record of a slow kind, which ends the
// straight-line loop, so the loop needs no bound test.
template <bool CONTAINS, bool COUNT, bool WIDE, bool FAST>
__global__ __launch_bounds__(kWalkThreads) void walk_fp64_kernel(
    const Walk64Rec* __restrict__ wrec, unsigned wbytes,
    const double4* __restrict__ wcenter, const int* __restrict__ d_wcount,
    const double4* __restrict__ rec, const uint32_t* __restrict__ idx,
    const uint32_t* __restrict__ order, int n, int* __restrict__ queue, double eps2,
    double G, double* __restrict__ ax_out, double* __restrict__ ay_out,
    double* __restrict__ az_out, int* err, unsigned long long* __restrict__ counters)
{
    const int lane = threadIdx.x & 63;
    const int packets = (n + 63) / 64;
    Queue q{queue, packets, (packets + 7) / 8, xcd_id(), 0};
    const int pa = q.take();
    if (pa < 0)
        return;
    const int wcount = *d_wcount;
    const __amdgpu_buffer_rsrc_t rsrc = __builtin_amdgcn_make_buffer_rsrc(
        const_cast<Walk64Rec*>(wrec), 0, static_cast<int>(WIDE ? 0u : wbytes), 0x00020000);
    Walker64 A, B;
    A.npp = A.npn = A.visits = 0;
    B.npp = B.npn = B.visits = 0;
    A.start(rec, order, pa, lane, n);
    const int pb = q.take();
    bool run_a = true, run_b = pb >= 0;
    if (run_b)
        B.start(rec, order, pb, lane, n);
    while (run_a && run_b) {
        Walk64Rec ra, rb;
        bool open_a, open_b;
        for (;;) {
            int sa, sb;
            if (WIDE) {
                sa = skip64_vgpr(wrec + A.i);
                sb = skip64_vgpr(wrec + B.i);
                load_pair64(wrec + A.i, wrec + B.i, ra, rb);
            } else {
                const unsigned oa = static_cast<unsigned>(A.i) << 6;
                const unsigned ob = static_cast<unsigned>(B.i) << 6;
                sa = skip64_vgpr_off(rsrc, oa);
                sb = skip64_vgpr_off(rsrc, ob);
                load_pair64_off(wrec, oa, ob, ra, rb);
            }
            open_a = visit_fast<CONTAINS, COUNT, FAST>(A, ra, sa, wcenter, eps2);
            open_b = visit_fast<CONTAINS, COUNT, FAST>(B, rb, sb, wcenter, eps2);
            if ((ra.kind | rb.kind) != KIND_INTERNAL)
                break;
            if (COUNT) {
                ++A.visits;
                ++B.visits;
            }
            A.i = open_a ? A.i + 1 : ra.skip;
            B.i = open_b ? B.i + 1 : rb.skip;
        }
        if (ra.kind == KIND_INTERNAL) {
            if (COUNT)
                ++A.visits;
            A.i = open_a ? A.i + 1 : ra.skip;
        } else if (A.i < wcount) {
            if (COUNT)
                ++A.visits;
            visit_rare<COUNT>(A, ra, rec, eps2, G);
        } else {
            finish64<COUNT>(A, idx, n, ax_out, ay_out, az_out, err);
            const int p = q.take();
            if (p >= 0)
                A.start(rec, order, p, lane, n);
            else
                run_a = false;
        }
        if (rb.kind == KIND_INTERNAL) {
            if (COUNT)
                ++B.visits;
            B.i = open_b ? B.i + 1 : rb.skip;
        } else if (B.i < wcount) {
            if (COUNT)
                ++B.visits;
            visit_rare<COUNT>(B, rb, rec, eps2, G);
        } else {
            finish64<COUNT>(B, idx, n, ax_out, ay_out, az_out, err);
            const int p = q.take();
            if (p >= 0)
                B.start(rec, order, p, lane, n);
            else
                run_b = false;
        }
    }
    // the slot still walking finishes alone
    if (run_a)
        walk_alone<CONTAINS, COUNT, WIDE, FAST>(A, q, wrec, rsrc, wcenter, wcount, rec, idx, order,
                                          n, lane, eps2, G, ax_out, ay_out, az_out, err);
    else if (run_b)
        walk_alone<CONTAINS, COUNT, WIDE, FAST>(B, q, wrec, rsrc, wcenter, wcount, rec, idx, order,
                                          n, lane, eps2, G, ax_out, ay_out, az_out, err);
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

// Softened FP64 direct sum on a target list: one block per target, fixed
// reduction order.
__global__ __launch_bounds__(256) void direct_kernel(
    const double* __restrict__ m, const double* __restrict__ x,
    const double* __restrict__ y, const double* __restrict__ z, size_t n,
    const uint32_t* __restrict__ targets, double eps2, double G,
    double* __restrict__ ax, double* __restrict__ ay, double* __restrict__ az)
{
    __shared__ double s_red[3][4];
    const uint32_t t = targets[blockIdx.x];
    if (t >= n)
        return;
    const double tx = x[t], ty = y[t], tz = z[t];
    double a0 = 0.0, a1 = 0.0, a2 = 0.0;
    bool sing = false;
    for (size_t s = threadIdx.x; s < n; s += blockDim.x) {
        if (s == t)
            continue;
        pm_fp64(tx, ty, tz, x[s], y[s], z[s], m[s], G, eps2, a0, a1, a2, sing);
    }
    for (int o = 32; o > 0; o >>= 1) {
        a0 += __shfl_xor(a0, o);
        a1 += __shfl_xor(a1, o);
        a2 += __shfl_xor(a2, o);
    }
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    if (lane == 0) {
        s_red[0][wave] = a0;
        s_red[1][wave] = a1;
        s_red[2][wave] = a2;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        ax[blockIdx.x] = (s_red[0][0] + s_red[0][1]) + (s_red[0][2] + s_red[0][3]);
        ay[blockIdx.x] = (s_red[1][0] + s_red[1][1]) + (s_red[1][2] + s_red[1][3]);
        az[blockIdx.x] = (s_red[2][0] + s_red[2][1]) + (s_red[2][2] + s_red[2][3]);
    }
}

} // namespace

// Blocks of 4 packets, padded to a multiple of 8 for the XCD mapping.
int walk_blocks_padded(int n)
{
    const int packets = (n + 63) / 64;
    const int blocks = (packets + 3) / 4;
    return (blocks + 7) / 8 * 8;
}

void launch_walk_fp64(hipStream_t s, const Walk64Rec* w64, size_t wcap,
                      const double4* wcenter, const int* d_wcount,
                      const double4* rec, const uint32_t* idx,
                      const uint32_t* order, int n, double eps2, double G,
                      bool contains, double side, double* ax, double* ay, double* az,
                      int* queue, int* err, unsigned long long* counters)
{
    // Every separation lies inside the box, so d2 <= 3 side^2: with these
    // bounds every r2 = d2 + eps2 is in the reduced sequences' range.
    const bool fast = side > 0.0 && eps2 >= 0x1p-600 && 3.0 * side * side + eps2 <= 0x1p600;
    // 32-bit record offsets while the records span less than 2 GiB
    const size_t wbytes = wcap * sizeof(Walk64Rec);
    const bool wide = wbytes > 0x7FFFFFFFull;
    queue_clear_kernel<<<1, 64, 0, s>>>(queue);
    NBH_HIP(hipGetLastError());
    const long long packets = (static_cast<long long>(n) + 63) / 64;
    // one pair of packets per wavefront, as many as stay resident: 7 per
    // SIMD for the reduced arithmetic without the containment test, 6 for
    // the other variants (more registers)
    const int nb = walk_persistent_blocks(
        packets, 2, (fast && !contains) ? kWalkFp64WavesPerSimd : kWalkFp64WavesPerSimdOther);
    if (nb == 0)
        return;
#define NBH_ARGS                                                                 \
    w64, static_cast<unsigned>(wide ? 0 : wbytes), wcenter, d_wcount, rec, idx,  \
        order, n, queue, eps2, G, ax, ay, az, err, counters
#define NBH_LAUNCH(C, K)                                                                \
    do {                                                                                \
        if (wide)                                                                       \
            walk_fp64_kernel<C, K, true, false><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS);  \
        else if (fast)                                                                  \
            walk_fp64_kernel<C, K, false, true><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS);  \
        else                                                                            \
            walk_fp64_kernel<C, K, false, false><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS); \
    } while (0)
    const bool cnt = counters != nullptr;
    if (contains) {
        if (cnt) NBH_LAUNCH(true, true); else NBH_LAUNCH(true, false);
    } else {
        if (cnt) NBH_LAUNCH(false, true); else NBH_LAUNCH(false, false);
    }
#undef NBH_LAUNCH
#undef NBH_ARGS
    NBH_HIP(hipGetLastError());
}

void launch_direct(hipStream_t s, const double* m, const double* x,
                   const double* y, const double* z, size_t n,
                   const uint32_t* targets, size_t n_targets, double eps2,
                   double G, double* ax, double* ay, double* az)
{
    if (n_targets == 0)
        return;
    direct_kernel<<<static_cast<unsigned>(n_targets), 256, 0, s>>>(
        m, x, y, z, n, targets, eps2, G, ax, ay, az);
    NBH_HIP(hipGetLastError());
}

} // namespace nbh
