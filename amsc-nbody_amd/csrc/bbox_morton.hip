// bbox_morton.hip -- global bounding box, Morton encoding and the
// Morton-order particle gather.
//
// Bounding box: one pass, FP64 min/max per lane, wave64 shuffles, LDS across
// the 4 waves of a block and one record per block; a one-block pass folds the
// records and finishes the box on the device with the reference's arithmetic
// (globalbounding.cu:122-136: centre = (min+max)*0.5, side = 1.1*max range,
// side 0 -> 1).  No arrival counter: a fence and a same-address atomic per
// block serialise across thousands of blocks.  The kick-drift variant applies the first half kick and the
// drift of the leapfrog step (integrator.cu:51-83) in the same pass, so the
// new positions are reduced while they are in registers.
//
// Encode: normalise -> quantise -> interleave fused in one kernel
// (morton.cu:98-139 run as 7 launches in the reference), writing the key,
// the identity index and the digit histograms of every radix-sort pass.
#include "device_utils.hpp"
#include "nbh_internal.hpp"

#include <cmath>

namespace nbh {

namespace {

constexpr int kBoxThreads = 256;
constexpr int kBoxMaxBlocks = 2048;
static_assert(kBoxMaxBlocks % kBoxThreads == 0, "finish pass reads whole rounds");

struct Extent {
    double lo[3];
    double hi[3];
    bool bad;
    double mmax; // largest |m| (0 when masses are not read)
};

__device__ __forceinline__ void block_reduce_store(Extent e, double* partials)
{
    __shared__ double s_red[4][8];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    double v[8];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        v[a] = wave_min(e.lo[a]);
        v[3 + a] = wave_max(e.hi[a]);
    }
    v[6] = (__ballot(e.bad) != 0ull) ? 1.0 : 0.0;
    v[7] = wave_max(e.mmax);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            s_red[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            r[k] = s_red[0][k];
        for (int w = 1; w < 4; ++w) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                r[a] = fmin(r[a], s_red[w][a]);
                r[3 + a] = fmax(r[3 + a], s_red[w][3 + a]);
            }
            r[6] = fmax(r[6], s_red[w][6]);
            r[7] = fmax(r[7], s_red[w][7]);
        }
#pragma unroll
        for (int k = 0; k < 8; ++k)
            partials[blockIdx.x * 8 + k] = r[k];
    }
}

// One block: folds `nblocks` block records (all loaded before they are
// folded) and writes the box, plus box[4] = the largest |m| (1 if every
// mass is zero) when `with_mass`.
__global__ __launch_bounds__(kBoxThreads) void box_finish_kernel(
    const double* __restrict__ partials, unsigned nblocks, double* __restrict__ box,
    int* __restrict__ err, bool with_mass)
{
    constexpr int kPer = kBoxMaxBlocks / kBoxThreads;
    double rec[kPer][8];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const unsigned b = threadIdx.x + k * kBoxThreads;
        const double* p = partials + 8 * static_cast<size_t>(b < nblocks ? b : 0);
#pragma unroll
        for (int c = 0; c < 8; ++c)
            rec[k][c] = p[c];
    }
    const double inf = __builtin_inf();
    Extent e{{inf, inf, inf}, {-inf, -inf, -inf}, false, 0.0};
    double bad = 0.0;
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (threadIdx.x + k * kBoxThreads >= nblocks)
            continue;
#pragma unroll
        for (int a = 0; a < 3; ++a) {
            e.lo[a] = fmin(e.lo[a], rec[k][a]);
            e.hi[a] = fmax(e.hi[a], rec[k][3 + a]);
        }
        bad = fmax(bad, rec[k][6]);
        e.mmax = fmax(e.mmax, rec[k][7]);
    }
    e.bad = bad != 0.0;
    __shared__ double s_red[4][8];
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    double v[8];
#pragma unroll
    for (int a = 0; a < 3; ++a) {
        v[a] = wave_min(e.lo[a]);
        v[3 + a] = wave_max(e.hi[a]);
    }
    v[6] = (__ballot(e.bad) != 0ull) ? 1.0 : 0.0;
    v[7] = wave_max(e.mmax);
    if (lane == 0) {
#pragma unroll
        for (int k = 0; k < 8; ++k)
            s_red[wave][k] = v[k];
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double r[8];
#pragma unroll
        for (int k = 0; k < 8; ++k)
            r[k] = s_red[0][k];
        for (int w = 1; w < 4; ++w) {
#pragma unroll
            for (int a = 0; a < 3; ++a) {
                r[a] = fmin(r[a], s_red[w][a]);
                r[3 + a] = fmax(r[3 + a], s_red[w][3 + a]);
            }
            r[6] = fmax(r[6], s_red[w][6]);
            r[7] = fmax(r[7], s_red[w][7]);
        }
        // globalbounding.cu:122-136
        const double cx = (r[0] + r[3]) * 0.5;
        const double cy = (r[1] + r[4]) * 0.5;
        const double cz = (r[2] + r[5]) * 0.5;
        const double rx = r[3] - r[0];
        const double ry = r[4] - r[1];
        const double rz = r[5] - r[2];
        double m = rx;
        if (m < ry)
            m = ry;
        if (m < rz)
            m = rz;
        double side = 1.1 * m;
        if (side == 0.0)
            side = 1.0;
        box[0] = cx;
        box[1] = cy;
        box[2] = cz;
        box[3] = side;
        if (with_mass)
            box[4] = r[7] > 0.0 ? r[7] : 1.0;
        if (r[6] != 0.0)
            atomicCAS(err, 0, 1);
    }
}

__global__ __launch_bounds__(kBoxThreads) void bbox_kernel(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, size_t n, double* __restrict__ partials,
    const double* __restrict__ m, double4* __restrict__ pos4)
{
    const double inf = __builtin_inf();
    Extent e{{inf, inf, inf}, {-inf, -inf, -inf}, false, 0.0};
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < n; i += stride) {
        const double px = x[i], py = y[i], pz = z[i];
        if (pos4) {
            const double pm = m[i];
            pos4[i] = make_double4(px, py, pz, pm);
            e.mmax = fmax(e.mmax, fabs(pm));
        }
        e.bad = e.bad || !(isfinite(px) && isfinite(py) && isfinite(pz));
        e.lo[0] = fmin(e.lo[0], px);
        e.lo[1] = fmin(e.lo[1], py);
        e.lo[2] = fmin(e.lo[2], pz);
        e.hi[0] = fmax(e.hi[0], px);
        e.hi[1] = fmax(e.hi[1], py);
        e.hi[2] = fmax(e.hi[2], pz);
    }
    block_reduce_store(e, partials);
}

// First half kick + drift (integrator.cu:51-83: v += (dt/2) a; x += dt v,
// contracted to fma) fused with the bounding-box reduction of the new
// positions.
__global__ __launch_bounds__(kBoxThreads) void kick_drift_bbox_kernel(
    double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, size_t n, double dt,
    double* __restrict__ partials, const double* __restrict__ m,
    double4* __restrict__ pos4)
{
    const double half = 0.5 * dt;
    const double inf = __builtin_inf();
    Extent e{{inf, inf, inf}, {-inf, -inf, -inf}, false, 0.0};
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < n; i += stride) {
        const double wx = fma(half, ax[i], vx[i]);
        const double wy = fma(half, ay[i], vy[i]);
        const double wz = fma(half, az[i], vz[i]);
        const double px = fma(dt, wx, x[i]);
        const double py = fma(dt, wy, y[i]);
        const double pz = fma(dt, wz, z[i]);
        vx[i] = wx;
        vy[i] = wy;
        vz[i] = wz;
        x[i] = px;
        y[i] = py;
        z[i] = pz;
        if (pos4) {
            const double pm = m[i];
            pos4[i] = make_double4(px, py, pz, pm);
            e.mmax = fmax(e.mmax, fabs(pm));
        }
        e.bad = e.bad || !(isfinite(px) && isfinite(py) && isfinite(pz));
        e.lo[0] = fmin(e.lo[0], px);
        e.lo[1] = fmin(e.lo[1], py);
        e.lo[2] = fmin(e.lo[2], pz);
        e.hi[0] = fmax(e.hi[0], px);
        e.hi[1] = fmax(e.hi[1], py);
        e.hi[2] = fmax(e.hi[2], pz);
    }
    block_reduce_store(e, partials);
}

unsigned box_blocks(size_t n)
{
    unsigned b = blocks_for(n, kBoxThreads * 2);
    return b > kBoxMaxBlocks ? kBoxMaxBlocks : b;
}

// ------------------------------------------------------------- Morton
// 10-bit spread with the reference's masks (morton.cu:119-128).
__device__ __forceinline__ uint32_t spread10(uint32_t v)
{
    uint32_t x = v & 0x000003FFu;
    x = (x | (x << 16)) & 0x030000FFu;
    x = (x | (x << 8)) & 0x0300F00Fu;
    x = (x | (x << 4)) & 0x030C30C3u;
    x = (x | (x << 2)) & 0x09249249u;
    return x;
}

// 21-bit spread: bit k -> bit 3k, by halving shift distances.
__device__ __forceinline__ uint64_t spread21(uint32_t v)
{
    uint64_t x = v & 0x1FFFFFull;
    x = (x | (x << 32)) & 0x001F00000000FFFFull;
    x = (x | (x << 16)) & 0x001F0000FF0000FFull;
    x = (x | (x << 8)) & 0x100F00F00F00F00Full;
    x = (x | (x << 4)) & 0x10C30C30C30C30C3ull;
    x = (x | (x << 2)) & 0x1249249249249249ull;
    return x;
}

template <int BITS>
__device__ __forceinline__ typename Layout<BITS>::Key morton_key(uint32_t qx,
                                                                 uint32_t qy,
                                                                 uint32_t qz)
{
    if constexpr (BITS == 10)
        return (spread10(qx) << 2) | (spread10(qy) << 1) | spread10(qz);
    else
        return (spread21(qx) << 2) | (spread21(qy) << 1) | spread21(qz);
}

// morton.cu:98-117: (a - (c - s/2)) / s, * 2^bits, clamp, truncate.
template <int BITS>
__device__ __forceinline__ uint32_t quantise(double a, double origin,
                                             double side)
{
    constexpr double grid = static_cast<double>(1u << BITS);
    constexpr double maxq = static_cast<double>((1u << BITS) - 1u);
    const double normalised = (a - origin) / side;
    const double scaled = normalised * grid;
    const double clamped = fmin(fmax(scaled, 0.0), maxq);
    return static_cast<uint32_t>(clamped);
}

template <int BITS>
__global__ __launch_bounds__(256) void encode_kernel(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, size_t n, const double* __restrict__ box,
    typename Layout<BITS>::Key* __restrict__ keys, uint32_t* __restrict__ idx,
    uint32_t* __restrict__ hist)
{
    using Key = typename Layout<BITS>::Key;
    constexpr int P = Layout<BITS>::sort_passes;
    __shared__ uint32_t s_hist[P * 256];
    for (int j = threadIdx.x; j < P * 256; j += blockDim.x)
        s_hist[j] = 0;
    __syncthreads();
    const double side = box[3];
    const double ox = box[0] - side / 2.0;
    const double oy = box[1] - side / 2.0;
    const double oz = box[2] - side / 2.0;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    for (size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
         i < n; i += stride) {
        const uint32_t qx = quantise<BITS>(x[i], ox, side);
        const uint32_t qy = quantise<BITS>(y[i], oy, side);
        const uint32_t qz = quantise<BITS>(z[i], oz, side);
        const Key k = morton_key<BITS>(qx, qy, qz);
        keys[i] = k;
        idx[i] = static_cast<uint32_t>(i);
#pragma unroll
        for (int p = 0; p < P; ++p)
            atomicAdd(&s_hist[p * 256 + static_cast<uint32_t>((k >> (8 * p)) & 0xFFu)], 1u);
    }
    __syncthreads();
    for (int j = threadIdx.x; j < P * 256; j += blockDim.x) {
        const uint32_t c = s_hist[j];
        if (c)
            atomicAdd(&hist[j], c);
    }
}

// Morton-order gather (simulation.cu:33-57) into packed FP64 records
// {x, y, z, m}, plus the fixed-point records of the mixed traversal:
// coordinates on a 2^31 grid over the box, and m / box[4] in FP32, box[4]
// being the largest |m| (so that sum (m / box[4]) d / r^3, with d and r in
// grid units, times G box[4] / unit^2, unit = side / 2^31, is the
// acceleration in physical units; the FP32 sums then stay in range for any
// unit system).
__global__ __launch_bounds__(256) void gather_kernel(
    const uint32_t* __restrict__ idx, const double* __restrict__ m,
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, size_t n, double4* __restrict__ rec64,
    int4* __restrict__ rec32, const double* __restrict__ box,
    const double4* __restrict__ pos4)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    uint32_t p = idx[i];
    if (p >= n)
        p = 0; // unreachable for a permutation; keeps every read in bounds
    double px, py, pz, pm;
    if (pos4) {
        const double4 q = pos4[p]; // one 32-byte read instead of four
        px = q.x;
        py = q.y;
        pz = q.z;
        pm = q.w;
    } else {
        px = x[p];
        py = y[p];
        pz = z[p];
        pm = m[p];
    }
    rec64[i] = make_double4(px, py, pz, pm);
    if (rec32) {
        const double side = box[3];
        const double scale = 2147483648.0 / side;
        const double ox = box[0] - side / 2.0;
        const double oy = box[1] - side / 2.0;
        const double oz = box[2] - side / 2.0;
        const double lim = 2147483647.0;
        const int qx = static_cast<int>(fmin(fmax(rint((px - ox) * scale), 0.0), lim));
        const int qy = static_cast<int>(fmin(fmax(rint((py - oy) * scale), 0.0), lim));
        const int qz = static_cast<int>(fmin(fmax(rint((pz - oz) * scale), 0.0), lim));
        const float gm = static_cast<float>(pm / box[4]);
        rec32[i] = make_int4(qx, qy, qz, __float_as_int(gm));
    }
}

} // namespace

void launch_bbox(nbh_context* ctx, const double* x, const double* y,
                 const double* z, size_t n, double* d_box, double* partials,
                 int* err, const double* m, double4* pos4)
{
    const unsigned nb = box_blocks(n);
    bbox_kernel<<<nb, kBoxThreads, 0, ctx->stream>>>(x, y, z, n, partials, m, pos4);
    box_finish_kernel<<<1, kBoxThreads, 0, ctx->stream>>>(partials, nb, d_box, err,
                                                         pos4 != nullptr);
    NBH_HIP(hipGetLastError());
}

void launch_kick_drift_bbox(hipStream_t s, double* x, double* y, double* z,
                            double* vx, double* vy, double* vz,
                            const double* ax, const double* ay,
                            const double* az, size_t n, double dt,
                            double* d_box, double* partials, int* err,
                            const double* m, double4* pos4)
{
    const unsigned nb = box_blocks(n);
    kick_drift_bbox_kernel<<<nb, kBoxThreads, 0, s>>>(x, y, z, vx, vy, vz, ax, ay, az, n,
                                                       dt, partials, m, pos4);
    box_finish_kernel<<<1, kBoxThreads, 0, s>>>(partials, nb, d_box, err,
                                               pos4 != nullptr);
    NBH_HIP(hipGetLastError());
}

size_t bbox_partial_doubles() { return static_cast<size_t>(kBoxMaxBlocks) * 8; }

template <int BITS>
void launch_encode(hipStream_t s, const double* x, const double* y,
                   const double* z, size_t n, const double* d_box,
                   typename Layout<BITS>::Key* keys, uint32_t* idx,
                   uint32_t* hist)
{
    unsigned blocks = blocks_for(n, 256 * 8);
    if (blocks > 2048)
        blocks = 2048;
    encode_kernel<BITS><<<blocks, 256, 0, s>>>(x, y, z, n, d_box, keys, idx,
                                               hist);
    NBH_HIP(hipGetLastError());
}

template void launch_encode<10>(hipStream_t, const double*, const double*,
                                const double*, size_t, const double*,
                                uint32_t*, uint32_t*, uint32_t*);
template void launch_encode<21>(hipStream_t, const double*, const double*,
                                const double*, size_t, const double*,
                                uint64_t*, uint32_t*, uint32_t*);

void launch_gather(hipStream_t s, const uint32_t* idx, const double* m,
                   const double* x, const double* y, const double* z,
                   size_t n, double4* rec64, int4* rec32, const double* box,
                   const double4* pos4)
{
    gather_kernel<<<blocks_for(n, 256), 256, 0, s>>>(idx, m, x, y, z, n, rec64,
                                                     rec32, box, pos4);
    NBH_HIP(hipGetLastError());
}

} // namespace nbh
