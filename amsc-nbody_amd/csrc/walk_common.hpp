// walk_common.hpp -- definitions shared by the walk layout and the
// traversal kernels.
#pragma once

#include "device_utils.hpp"
#include "mac_threshold.hpp"
#include "nbh_internal.hpp"

namespace nbh {

// Node kinds of the walk records.
enum : int {
    KIND_INTERNAL = 0,   // MAC decides accept / open
    KIND_LEAF = 1,       // exact particle interactions
    KIND_SKIP = 2,       // zero mass: contributes nothing, never opened
    KIND_CHILDLESS = 3,  // internal with mass but no children: open -> error 5
    KIND_ERROR = 100     // kind - 100 is the traversal error code
};

// Mixed walk records keep a leaf's particle count in the 23 bits above the
// kind byte of the link word; a leaf of kMixedLeafCountLimit or more
// particles is written as an error record of this traversal code.
constexpr int kMixedLeafCountLimit = 1 << 23;
constexpr int kErrMixedLeafTooLarge = 9;
// Kind-byte flag of the mixed records that start a split walk's subtrees
// (crit^2 = +inf while marked; their own crit^2 and kind word are kept in
// the split table).
constexpr int kMixedBoundary = 0x80;


// "Error 4 (inconsistent topology), skip to `skip`": the FP64 walk's
// terminal record at the walk length and the initial content of every
// record buffer position.
__host__ __device__ inline Walk64Rec error_record64(int skip)
{
    Walk64Rec r;
    r.x = r.y = r.z = r.m = 0.0;
    r.thr = INFINITY;
    r.skip = skip;
    r.kind = KIND_ERROR + 4;
    r.first = -1;
    r.count = 0;
    r.gm = 0.0;
    return r;
}

constexpr int kWalkThreads = 256;

// ints between the work-queue counters of two ranges (one per 128 bytes)
constexpr int kQueueStride = kWalkQueueInts / 8;

// The work queue of the persistent traversals: packets (units) are dealt
// from eight ranges, one eighth of the target order each, through one
// counter per range.  A wavefront takes packets from the range of its own
// XCD, so each XCD's L2 serves one region of space, and from the other
// ranges once its own is empty.
struct Queue {
    int* counters; // counter of range r at counters[r * kQueueStride]
    int packets;
    int per;       // packets per range
    int home;      // this wavefront's XCD
    int tried;     // ranges found empty so far, in order home, home + 1, ...

    __device__ __forceinline__ int take()
    {
        while (tried < 8) {
            const int r = (home + tried) & 7;
            const int lo = r * per;
            const int hi = min(lo + per, packets);
            int k = 0;
            if ((threadIdx.x & 63) == 0)
                k = atomicAdd(&counters[r * kQueueStride], 1);
            k = __builtin_amdgcn_readfirstlane(k);
            if (lo + k < hi)
                return lo + k;
            ++tried;
        }
        return -1;
    }
};

__device__ __forceinline__ int xcd_id()
{
    unsigned v;
    asm volatile("s_getreg_b32 %0, hwreg(HW_REG_XCC_ID)" : "=s"(v));
    return static_cast<int>(v & 7u);
}

namespace {

__global__ void queue_clear_kernel(int* __restrict__ queue)
{
    if (threadIdx.x < 8)
        queue[threadIdx.x * kQueueStride] = 0;
}

// The order in which the queue hands out the units of a traversal (its
// packets, or with a split table packets times subtrees), from each unit's
// visit count in the previous traversal of the same particles (unit k then
// held nearly the same targets along the curve).  Each of the eight queue
// ranges ends with its `tail` cheapest units (to the cost bin) in
// decreasing cost; the others come first, in curve order.  The last units
// a range hands out are then its cheapest, so the slots run out of work
// close together instead of waiting on units taken just before the queue
// emptied, while the others keep the curve order that the slots of an XCD
// share records by.  Costs of zero (no previous traversal) or tail = 0
// leave the curve order.  perm[q] is the unit handed out at queue position
// q; units within one cost bin of the tail may come in any order.  One
// block of kScheduleThreads per range.
constexpr int kScheduleThreads = 1024;
constexpr int kScheduleBins = 256;

__global__ __launch_bounds__(kScheduleThreads) void packet_schedule_kernel(
    const uint32_t* __restrict__ cost, int packets, const int* __restrict__ split, int tail,
    uint32_t* __restrict__ perm)
{
    __shared__ unsigned s_bin[kScheduleBins]; // histogram, then cursors
    __shared__ unsigned s_wave[kScheduleThreads / 64];
    __shared__ unsigned s_cut[2];             // last tail bin, units in the tail
    const int units = split ? packets * split[0] : packets;
    const int per = (units + 7) / 8;
    const int lo = static_cast<int>(blockIdx.x) * per;
    const int hi = min(lo + per, units);
    if (lo >= hi)
        return;
    const int tid = static_cast<int>(threadIdx.x);
    const int lane = tid & 63;
    const int wave = tid >> 6;
    unsigned mx = 0;
    for (int k = lo + tid; k < hi; k += kScheduleThreads)
        mx = max(mx, cost[k]);
    for (int o = 32; o > 0; o >>= 1)
        mx = max(mx, static_cast<unsigned>(__shfl_xor(static_cast<int>(mx), o)));
    if (lane == 0)
        s_wave[wave] = mx;
    for (int b = tid; b < kScheduleBins; b += kScheduleThreads)
        s_bin[b] = 0;
    __syncthreads();
    mx = 0;
    for (int w = 0; w < kScheduleThreads / 64; ++w)
        mx = max(mx, s_wave[w]);
    if (mx == 0 || tail <= 0) {
        for (int k = lo + tid; k < hi; k += kScheduleThreads)
            perm[k] = static_cast<uint32_t>(k);
        return;
    }
    auto bin_of = [mx](unsigned c) {
        return static_cast<int>((static_cast<unsigned long long>(c) * kScheduleBins) /
                                (static_cast<unsigned long long>(mx) + 1));
    };
    for (int k = lo + tid; k < hi; k += kScheduleThreads)
        atomicAdd(&s_bin[bin_of(cost[k])], 1u);
    __syncthreads();
    // The tail is bins 0..cut, the lowest that hold `tail` units, laid out
    // after the other units from bin cut down to bin 0: inclusive prefix
    // sums over the bins (one bin per thread of the first kScheduleBins),
    // then each tail bin's cursor.
    static_assert(kScheduleBins % 64 == 0 && kScheduleBins <= kScheduleThreads,
                  "one bin per thread of whole wavefronts");
    unsigned cnt = 0, incl = 0;
    if (tid < kScheduleBins) {
        cnt = s_bin[tid];
        incl = cnt;
        for (int o = 1; o < 64; o <<= 1) {
            const unsigned y = static_cast<unsigned>(__shfl_up(static_cast<int>(incl), o));
            incl += lane >= o ? y : 0u;
        }
        if (lane == 63)
            s_wave[wave] = incl;
    }
    __syncthreads();
    if (tid < kScheduleBins) {
        for (int w = 0; w < wave; ++w)
            incl += s_wave[w];
        const unsigned want = static_cast<unsigned>(min(tail, hi - lo));
        const unsigned excl = incl - cnt;
        // the first bin whose prefix reaches `want` (the last bin if none)
        if ((incl >= want && excl < want) || (tid == kScheduleBins - 1 && incl < want)) {
            s_cut[0] = static_cast<unsigned>(tid);
            s_cut[1] = incl; // units in the tail
        }
    }
    __syncthreads();
    if (tid < kScheduleBins && tid <= static_cast<int>(s_cut[0])) {
        const unsigned head = static_cast<unsigned>(hi - lo) - s_cut[1];
        s_bin[tid] = head + (s_cut[1] - incl);
    }
    __syncthreads();
    const int cut = static_cast<int>(s_cut[0]);
    unsigned base = 0; // head units placed by earlier chunks
    for (int c0 = lo; c0 < hi; c0 += kScheduleThreads) {
        const int k = c0 + tid;
        const int b = k < hi ? bin_of(cost[k]) : -1;
        const bool first = b > cut; // before the tail, in curve order
        const unsigned long long m = __ballot(first);
        const unsigned before = static_cast<unsigned>(
            __popcll(m & ((1ull << lane) - 1ull)));
        if (lane == 0)
            s_wave[wave] = static_cast<unsigned>(__popcll(m));
        __syncthreads();
        unsigned off = base;
        unsigned total = 0;
        for (int w = 0; w < kScheduleThreads / 64; ++w) {
            const unsigned c = s_wave[w];
            off += w < wave ? c : 0u;
            total += c;
        }
        if (first)
            perm[lo + off + before] = static_cast<uint32_t>(k);
        else if (b >= 0)
            perm[lo + atomicAdd(&s_bin[b], 1u)] = static_cast<uint32_t>(k);
        base += total;
        __syncthreads();
    }
}

} // namespace

// Blocks of 4 persistent wavefronts: one wavefront per unit_per_wave units
// of work, at most waves_per_simd per SIMD (4 SIMDs per CU).  The cap is
// what is resident at once, so that no wavefront starts after others have
// emptied the queue.  gfx950 keeps 8 wavefronts per SIMD resident only for
// kernels of at most 64 VGPRs and 80 SGPRs (VCC included): kernels of 81,
// 85 and 89 SGPRs were measured at 7, of 97 SGPRs at 6.
// Resident wavefronts per SIMD of the whole-walk traversal kernels at their
// register counts (tests/test_kernel_resources.py checks the counts).
constexpr int kWalkMixedWavesPerSimd = 8;
constexpr int kWalkFp64WavesPerSimd = 7;      // reduced arithmetic, no containment test
constexpr int kWalkFp64WavesPerSimdOther = 6; // the other FP64 variants

inline int walk_persistent_blocks(long long units, int units_per_wave, int waves_per_simd)
{
    static int cus[64] = {};
    int dev = 0;
    NBH_HIP(hipGetDevice(&dev));
    int& c = cus[dev & 63];
    if (c == 0)
        NBH_HIP(hipDeviceGetAttribute(&c, hipDeviceAttributeMultiprocessorCount, dev));
    const long long want = (units + units_per_wave - 1) / units_per_wave;
    const long long cap = 4LL * waves_per_simd * c;
    const int waves = static_cast<int>(want < cap ? want : cap);
    return (waves + 3) / 4;
}

// Static grids (the FP64 walks): packets are Morton-consecutive; blocks are
// dealt to the 8 XCDs round
// robin, so block b takes packet range (b % 8) * per + b / 8: each XCD's L2
// then serves one contiguous eighth of space.
__device__ __forceinline__ int packet_of_block(int nblocks_padded)
{
    const int b = static_cast<int>(blockIdx.x);
    const int per = nblocks_padded / 8;
    return (b % 8) * per + b / 8;
}

// barnes_hut.cu:56-80 in FP64 with the contraction order of the oracle.
__device__ __forceinline__ void pm_fp64(double tx, double ty, double tz,
                                        double sx, double sy, double sz,
                                        double sm, double G, double eps2,
                                        double& ax, double& ay, double& az,
                                        bool& singular)
{
    if (sm == 0.0)
        return;
    const double dx = sx - tx;
    const double dy = sy - ty;
    const double dz = sz - tz;
    const double d2 = fma(dz, dz, fma(dx, dx, dy * dy)) + eps2;
    if (d2 <= 0.0) {
        singular = true;
        return;
    }
    const double inv = 1.0 / sqrt(d2);
    const double inv3 = inv * inv * inv;
    const double f = G * sm * inv3;
    ax = fma(f, dx, ax);
    ay = fma(f, dy, ay);
    az = fma(f, dz, az);
}

} // namespace nbh
