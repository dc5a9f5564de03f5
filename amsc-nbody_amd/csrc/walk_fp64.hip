// walk_fp64.hip -- FP64 packet traversal (the reference arithmetic, bit-exact
// with the oracle) and the FP64 direct sum used as accuracy reference.
#include "walk_common.hpp"

namespace nbh {

namespace {

template <bool CONTAINS, bool COUNT>
__global__ __launch_bounds__(kWalkThreads) void walk_fp64_kernel(
    const double4* __restrict__ wnode, const double* __restrict__ wcrit,
    const int4* __restrict__ wlink, const double4* __restrict__ wcenter,
    const int* __restrict__ d_wcount, const double4* __restrict__ rec,
    const uint32_t* __restrict__ idx, int n, int nblocks_padded, double eps2,
    double G, double* __restrict__ ax_out, double* __restrict__ ay_out,
    double* __restrict__ az_out, int* err,
    unsigned long long* __restrict__ counters)
{
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int packet = packet_of_block(nblocks_padded) * 4 + wave;
    const int t = packet * 64 + lane;
    if (packet * 64 >= n)
        return;
    const bool valid = t < n;
    const int wcount = *d_wcount;
    double tx = 0.0, ty = 0.0, tz = 0.0;
    if (valid) {
        const double4 p = rec[t];
        tx = p.x;
        ty = p.y;
        tz = p.z;
    }
    double ax = 0.0, ay = 0.0, az = 0.0;
    unsigned long long npp = 0, npn = 0;
    int next = valid ? 0 : 0x7FFFFFFF;
    int code = 0;
    int i = 0;
    while (i < wcount) {
        const int4 lk = wlink[i];
        const int skip = lk.x;
        const int kind = lk.y;
        bool open = false;
        if (next <= i && code == 0) {
            if (kind == KIND_INTERNAL || kind == KIND_CHILDLESS) {
                const double4 nd = wnode[i];
                const double dx = nd.x - tx;
                const double dy = nd.y - ty;
                const double dz = nd.z - tz;
                const double d2 = fma(dz, dz, fma(dx, dx, dy * dy));
                bool accept = sqrt(d2) > wcrit[i];
                if (CONTAINS && accept) {
                    const double4 c = wcenter[i];
                    const bool inside = fabs(tx - c.x) <= c.w &&
                                        fabs(ty - c.y) <= c.w &&
                                        fabs(tz - c.z) <= c.w;
                    accept = !inside;
                }
                if (accept) {
                    bool sing = false;
                    pm_fp64(tx, ty, tz, nd.x, nd.y, nd.z, nd.w, G, eps2, ax,
                            ay, az, sing);
                    if (sing)
                        code = 6;
                    if (COUNT)
                        ++npn;
                    next = skip;
                } else if (kind == KIND_CHILDLESS) {
                    code = 5;
                } else {
                    open = true;
                    next = i + 1;
                }
            } else if (kind == KIND_LEAF) {
                const int first = lk.z;
                const int count = lk.w;
                bool sing = false;
                if (count == 1) {
                    if (first != t) {
                        const double4 s = wnode[i];
                        pm_fp64(tx, ty, tz, s.x, s.y, s.z, s.w, G, eps2, ax, ay,
                                az, sing);
                        if (COUNT && s.w != 0.0)
                            ++npp;
                    }
                } else {
                    for (int k = first; k < first + count; ++k) {
                        if (k == t)
                            continue;
                        const double4 s = rec[k];
                        pm_fp64(tx, ty, tz, s.x, s.y, s.z, s.w, G, eps2, ax, ay,
                                az, sing);
                        if (COUNT && s.w != 0.0)
                            ++npp;
                    }
                }
                if (sing)
                    code = 6;
                next = skip;
            } else if (kind == KIND_SKIP) {
                next = skip;
            } else {
                code = kind - KIND_ERROR;
            }
        }
        i = (__ballot(open) != 0ull) ? i + 1 : skip;
    }
    if (code != 0) {
        raise_error(err, code);
        return;
    }
    if (COUNT) {
        for (int o = 32; o > 0; o >>= 1) {
            npp += __shfl_xor(npp, o);
            npn += __shfl_xor(npn, o);
        }
        if (lane == 0) {
            atomicAdd(&counters[0], npp);
            atomicAdd(&counters[1], npn);
        }
        return;
    }
    if (valid) {
        const uint32_t o = idx[t];
        if (o >= static_cast<uint32_t>(n)) {
            raise_error(err, 3);
            return;
        }
        ax_out[o] = ax;
        ay_out[o] = ay;
        az_out[o] = az;
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

void launch_walk_fp64(hipStream_t s, const double4* wnode, const double* wcrit,
                      const int4* wlink, const double4* wcenter,
                      const int* d_wcount, const double4* rec,
                      const uint32_t* idx, int n, double eps2, double G,
                      bool contains, double* ax, double* ay, double* az,
                      int* err, unsigned long long* counters)
{
    const int nb = walk_blocks_padded(n);
#define NBH_ARGS                                                                 \
    wnode, wcrit, wlink, wcenter, d_wcount, rec, idx, n, nb, eps2, G, ax, ay, az,  \
        err, counters
    if (counters) {
        if (contains)
            walk_fp64_kernel<true, true><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS);
        else
            walk_fp64_kernel<false, true><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS);
    } else {
        if (contains)
            walk_fp64_kernel<true, false><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS);
        else
            walk_fp64_kernel<false, false><<<nb, kWalkThreads, 0, s>>>(NBH_ARGS);
    }
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
