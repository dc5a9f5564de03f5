// frame_capture.hip -- sampled, quantized position frames on the device
// (src/streaming/device_capture.cu:27-153 kernels, :155-330 DeviceFrameWriter).
//
// Sample j of S reads particle (j * N) / S (j when S == N).
//
// Bounds: the per-axis minimum and maximum of the finite sampled positions.
// Each value is reduced as an order-preserving 64-bit key, so the result does
// not depend on the reduction order.  The key order is total: -0 < +0, and
// NaN never enters it.  The bounds pass reduces a wave with shuffles, then a
// block through LDS, and writes one record per block: six keys and a flag
// for non-finite samples.  A one-block pass folds the block records.  No
// atomics: same-address atomics from thousands of blocks serialise in L2.
//
// Quantization:
//   q = 0                                               when max == min,
//   q = (uint16)fma(clamp((p - min)/(max - min), 0, 1), 65535, 0.5)   otherwise.
// The fma is the contraction nvcc applies to the reference's
// `normalized * 65535 + 0.5`; its --fmad default is on.
// The quantize pass reads the bounds from device memory, so the whole
// capture is three kernels and a single host synchronisation.  The full
// sample is processed at once: S <= N < 2^30, and 6 S bytes of staging fit
// any MI355X-sized run.
//
// Bytes per sampled particle: 24 read by the bounds pass, 24 read and 6
// written by the quantize pass.  Both passes are gathers when S < N and
// HBM-bound streams when S == N.
#include "nbh_internal.hpp"

#include <cstring>
#include <limits>

namespace nbh {

namespace {

constexpr unsigned kCaptureThreads = 256;
constexpr unsigned kBoundsMaxBlocks = 2048;
static_assert(kBoundsMaxBlocks % kCaptureThreads == 0, "finish pass reads whole rounds");
constexpr int kMaxParticleType = 4; // ParticleType::Asteroid (src/particle_type.hpp:7-18)

__device__ __forceinline__ unsigned long long order_key(double v)
{
    const auto b = static_cast<unsigned long long>(__double_as_longlong(v));
    return (b >> 63) ? ~b : (b | (1ull << 63));
}

__device__ __forceinline__ double key_value(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
    return __longlong_as_double(static_cast<long long>(b));
}

__device__ __forceinline__ size_t source_of(size_t j, size_t n, size_t s)
{
    return s == n ? j : (j * n) / s;
}

__device__ __forceinline__ uint16_t quantize(double p, double lo, double hi)
{
    if (hi == lo)
        return 0;
    double t = (p - lo) / (hi - lo);
    t = fmin(1.0, fmax(0.0, t));
    return static_cast<uint16_t>(fma(t, 65535.0, 0.5));
}

constexpr unsigned kWaves = kCaptureThreads / kWave;

// Block-level reduction of six keys and a flag.  red[c][wave] holds the wave
// results; returns the block result in threads 0..6 (key c, or the flag as 6).
__device__ __forceinline__ unsigned long long block_fold(unsigned long long v[6], bool bad,
                                                         unsigned long long (*red)[kWaves])
{
    for (int off = kWave / 2; off > 0; off >>= 1) {
#pragma unroll
        for (int c = 0; c < 6; ++c) {
            const unsigned long long o = __shfl_xor(v[c], off);
            v[c] = c < 3 ? min(v[c], o) : max(v[c], o);
        }
    }
    const bool wave_bad = __any(bad);
    const unsigned lane = threadIdx.x % kWave, wave = threadIdx.x / kWave;
    if (lane == 0) {
#pragma unroll
        for (int c = 0; c < 6; ++c)
            red[c][wave] = v[c];
        red[6][wave] = wave_bad ? 1ull : 0ull;
    }
    __syncthreads();
    unsigned long long r = 0;
    if (threadIdx.x < 7) {
        const unsigned c = threadIdx.x;
        r = red[c][0];
        for (unsigned w = 1; w < kWaves; ++w)
            r = c < 3 ? min(r, red[c][w]) : (c < 6 ? max(r, red[c][w]) : (r | red[c][w]));
    }
    return r;
}

// Per-block record: blocks[8 b + c] = minimum keys (c < 3), maximum keys
// (3 <= c < 6) and the non-finite flag (c = 6) of block b's samples.
__global__ __launch_bounds__(kCaptureThreads) void frame_bounds_kernel(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, size_t n, size_t s,
    unsigned long long* __restrict__ blocks)
{
    unsigned long long v[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
    bool invalid = false;
    const size_t stride = static_cast<size_t>(gridDim.x) * blockDim.x;
    // Four samples per lane per round, loads first, so that twelve loads
    // are in flight per lane.
    constexpr int kUnroll = 4;
    for (size_t j0 = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x; j0 < s;
         j0 += kUnroll * stride) {
        double px[kUnroll], py[kUnroll], pz[kUnroll];
        bool in[kUnroll];
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            const size_t j = j0 + u * stride;
            in[u] = j < s;
            const size_t i = in[u] ? source_of(j, n, s) : 0;
            px[u] = x[i], py[u] = y[i], pz[u] = z[i];
        }
#pragma unroll
        for (int u = 0; u < kUnroll; ++u) {
            if (!in[u])
                continue;
            if (isfinite(px[u]) && isfinite(py[u]) && isfinite(pz[u])) {
                const unsigned long long kx = order_key(px[u]), ky = order_key(py[u]),
                                         kz = order_key(pz[u]);
                v[0] = min(v[0], kx), v[1] = min(v[1], ky), v[2] = min(v[2], kz);
                v[3] = max(v[3], kx), v[4] = max(v[4], ky), v[5] = max(v[5], kz);
            } else {
                invalid = true;
            }
        }
    }
    __shared__ unsigned long long red[7][kWaves];
    const unsigned long long r = block_fold(v, invalid, red);
    if (threadIdx.x < 7)
        blocks[8 * static_cast<size_t>(blockIdx.x) + threadIdx.x] = r;
}

// One block: folds `nblocks` block records into keys[0..5] and keys[6].
__global__ __launch_bounds__(kCaptureThreads) void frame_bounds_finish_kernel(
    const unsigned long long* __restrict__ blocks, unsigned nblocks,
    unsigned long long* __restrict__ keys)
{
    unsigned long long v[6] = {~0ull, ~0ull, ~0ull, 0ull, 0ull, 0ull};
    bool invalid = false;
    // nblocks <= kBoundsMaxBlocks = 8 * kCaptureThreads: at most eight
    // records per thread, all loaded before they are folded.
    constexpr int kPer = kBoundsMaxBlocks / kCaptureThreads;
    unsigned long long rec[kPer][7];
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        const unsigned b = threadIdx.x + k * kCaptureThreads;
        const unsigned long long* r = blocks + 8 * static_cast<size_t>(b < nblocks ? b : 0);
#pragma unroll
        for (int c = 0; c < 7; ++c)
            rec[k][c] = r[c];
    }
#pragma unroll
    for (int k = 0; k < kPer; ++k) {
        if (threadIdx.x + k * kCaptureThreads >= nblocks)
            continue;
#pragma unroll
        for (int c = 0; c < 6; ++c)
            v[c] = c < 3 ? min(v[c], rec[k][c]) : max(v[c], rec[k][c]);
        invalid = invalid || rec[k][6] != 0ull;
    }
    __shared__ unsigned long long red[7][kWaves];
    const unsigned long long r = block_fold(v, invalid, red);
    if (threadIdx.x < 7)
        keys[threadIdx.x] = r;
}

// One thread per pair of samples.  WIDE: the six quantized values of a pair
// go out as three aligned 32-bit stores (positions buffer 4-byte aligned).
template <bool WIDE>
__global__ __launch_bounds__(kCaptureThreads) void frame_quantize_kernel(
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, size_t n, size_t s,
    const unsigned long long* __restrict__ keys, uint16_t* __restrict__ out)
{
    const size_t t = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    const size_t j0 = 2 * t;
    if (j0 >= s)
        return;
    const double lo[3] = {key_value(keys[0]), key_value(keys[1]), key_value(keys[2])};
    const double hi[3] = {key_value(keys[3]), key_value(keys[4]), key_value(keys[5])};
    uint16_t q[6] = {0, 0, 0, 0, 0, 0};
    const bool pair = j0 + 1 < s;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        if (k == 1 && !pair)
            break;
        const size_t i = source_of(j0 + k, n, s);
        q[3 * k + 0] = quantize(x[i], lo[0], hi[0]);
        q[3 * k + 1] = quantize(y[i], lo[1], hi[1]);
        q[3 * k + 2] = quantize(z[i], lo[2], hi[2]);
    }
    if (WIDE && pair) {
        uint32_t* o = reinterpret_cast<uint32_t*>(out) + 3 * t;
        o[0] = static_cast<uint32_t>(q[0]) | (static_cast<uint32_t>(q[1]) << 16);
        o[1] = static_cast<uint32_t>(q[2]) | (static_cast<uint32_t>(q[3]) << 16);
        o[2] = static_cast<uint32_t>(q[4]) | (static_cast<uint32_t>(q[5]) << 16);
    } else {
        const int m = pair ? 6 : 3;
        for (int c = 0; c < m; ++c)
            out[3 * j0 + c] = q[c];
    }
}

__global__ __launch_bounds__(kCaptureThreads) void frame_types_kernel(
    const uint8_t* __restrict__ type, size_t n, size_t s, uint8_t* __restrict__ out,
    int* __restrict__ bad)
{
    const size_t j = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    bool invalid = false;
    if (j < s) {
        const uint8_t v = type[source_of(j, n, s)];
        out[j] = v;
        invalid = v > kMaxParticleType;
    }
    if (__any(invalid) && threadIdx.x % kWave == 0)
        atomicOr(bad, 1);
}

void need(bool ok, int status, const char* message)
{
    if (!ok)
        fail(status, message);
}

void check_sample(size_t count, size_t sample_count)
{
    need(count != 0, NBH_INVALID_ARGUMENT,
         "Cannot capture an empty or incomplete device particle view");
    need(sample_count != 0 && sample_count <= count, NBH_INVALID_ARGUMENT,
         "frame sample count must be in [1, particle count]");
    need(count < (size_t{1} << 31), NBH_OVERFLOW_ERROR,
         "frame capture supports fewer than 2^31 particles");
}

} // namespace

// keys: 8 words (six keys, the flag, padding), then the block records.
size_t frame_capture_scratch_bytes() { return (8 + 8 * size_t{kBoundsMaxBlocks}) * 8; }

void launch_frame_capture(hipStream_t st, const double* x, const double* y,
                          const double* z, size_t n, size_t s, uint16_t* out,
                          unsigned long long* keys)
{
    const unsigned bb = static_cast<unsigned>(
        std::min<size_t>(kBoundsMaxBlocks, blocks_for(s, kCaptureThreads)));
    unsigned long long* blocks = keys + 8;
    hipLaunchKernelGGL(frame_bounds_kernel, dim3(bb), dim3(kCaptureThreads), 0, st, x, y,
                       z, n, s, blocks);
    hipLaunchKernelGGL(frame_bounds_finish_kernel, dim3(1), dim3(kCaptureThreads), 0, st,
                       blocks, bb, keys);
    const unsigned qb = blocks_for((s + 1) / 2, kCaptureThreads);
    if (reinterpret_cast<uintptr_t>(out) % 4 == 0)
        hipLaunchKernelGGL(frame_quantize_kernel<true>, dim3(qb), dim3(kCaptureThreads), 0,
                           st, x, y, z, n, s, keys, out);
    else
        hipLaunchKernelGGL(frame_quantize_kernel<false>, dim3(qb), dim3(kCaptureThreads), 0,
                           st, x, y, z, n, s, keys, out);
    NBH_HIP(hipGetLastError());
}

// Host decode of a bounds key (the inverse of order_key).
double frame_key_value(unsigned long long k)
{
    const unsigned long long b = (k >> 63) ? (k & ~(1ull << 63)) : ~k;
    double v;
    std::memcpy(&v, &b, sizeof(v));
    return v;
}

} // namespace nbh

using namespace nbh;

extern "C" int nbh_frame_capture(nbh_context* ctx, const double* d_x, const double* d_y,
                                 const double* d_z, size_t count, size_t sample_count,
                                 uint16_t* d_positions, double* h_bounds)
{
    return guard([&] {
        need(ctx != nullptr, NBH_INVALID_ARGUMENT, "context must not be null");
        need(d_x && d_y && d_z, NBH_INVALID_ARGUMENT,
             "Cannot capture an empty or incomplete device particle view");
        need(d_positions && h_bounds, NBH_INVALID_ARGUMENT,
             "frame output pointers must not be null");
        check_sample(count, sample_count);
        auto* keys = static_cast<unsigned long long*>(
            ctx->ws.get(SLOT_MISC, frame_capture_scratch_bytes()));
        launch_frame_capture(ctx->stream, d_x, d_y, d_z, count, sample_count, d_positions,
                             keys);
        // Readback: the six keys and the flag through the pinned words.
        NBH_HIP(hipMemcpyAsync(ctx->host_pinned, keys, 7 * sizeof(unsigned long long),
                               hipMemcpyDeviceToHost, ctx->stream));
        NBH_HIP(hipStreamSynchronize(ctx->stream));
        unsigned long long k[7];
        std::memcpy(k, ctx->host_pinned, sizeof(k));
        if (k[6] != 0)
            fail(NBH_RUNTIME_ERROR, "Cannot quantize a graphical frame with non-finite positions");
        for (int c = 0; c < 6; ++c)
            h_bounds[c] = frame_key_value(k[c]);
    });
}

extern "C" int nbh_frame_sample_types(nbh_context* ctx, const uint8_t* d_type, size_t count,
                                      size_t sample_count, uint8_t* d_types_out)
{
    return guard([&] {
        need(ctx != nullptr, NBH_INVALID_ARGUMENT, "context must not be null");
        need(d_type != nullptr, NBH_INVALID_ARGUMENT,
             "Cannot capture an empty or incomplete device particle view");
        need(d_types_out != nullptr, NBH_INVALID_ARGUMENT,
             "frame output pointers must not be null");
        check_sample(count, sample_count);
        int* bad = static_cast<int*>(ctx->ws.get(SLOT_MISC, frame_capture_scratch_bytes()));
        launch_fill_i32(ctx->stream, bad, 0, 1);
        hipLaunchKernelGGL(frame_types_kernel, dim3(blocks_for(sample_count, kCaptureThreads)),
                           dim3(kCaptureThreads), 0, ctx->stream, d_type, count, sample_count,
                           d_types_out, bad);
        NBH_HIP(hipGetLastError());
        int flag = 0;
        sync_read(ctx, bad, &flag, 1);
        if (flag != 0)
            fail(NBH_RUNTIME_ERROR, "Cannot capture a frame with invalid particle types");
    });
}
