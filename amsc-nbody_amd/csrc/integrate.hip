// integrate.hip -- leapfrog kernels and the four-array Morton gather.
//
// leapfrogKickDrift / leapfrogKick (integrator.cu:51-106): v += (dt/2) a and
// x += dt v, each contracted to one fma as the oracle states it.  The
// pipeline fuses the kick-drift into the bounding-box pass
// (bbox_morton.hip); these stand-alone kernels serve the C ABI entries and
// the closing half kick.
#include "nbh_internal.hpp"

namespace nbh {

namespace {

__global__ __launch_bounds__(256) void kick_drift_kernel(
    double* __restrict__ x, double* __restrict__ y, double* __restrict__ z,
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, size_t n, double dt)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double h = 0.5 * dt;
    const double wx = fma(h, ax[i], vx[i]);
    const double wy = fma(h, ay[i], vy[i]);
    const double wz = fma(h, az[i], vz[i]);
    vx[i] = wx;
    vy[i] = wy;
    vz[i] = wz;
    x[i] = fma(dt, wx, x[i]);
    y[i] = fma(dt, wy, y[i]);
    z[i] = fma(dt, wz, z[i]);
}

__global__ __launch_bounds__(256) void kick_kernel(
    double* __restrict__ vx, double* __restrict__ vy, double* __restrict__ vz,
    const double* __restrict__ ax, const double* __restrict__ ay,
    const double* __restrict__ az, size_t n, double dt)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const double h = 0.5 * dt;
    vx[i] = fma(h, ax[i], vx[i]);
    vy[i] = fma(h, ay[i], vy[i]);
    vz[i] = fma(h, az[i], vz[i]);
}

// gatherMortonOrderedParticleDataKernel (simulation.cu:33-57).
__global__ __launch_bounds__(256) void gather4_kernel(
    const uint32_t* __restrict__ idx, const double* __restrict__ m,
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, double* __restrict__ om,
    double* __restrict__ ox, double* __restrict__ oy, double* __restrict__ oz,
    size_t n)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t p = idx[i];
    om[i] = m[p];
    ox[i] = x[p];
    oy[i] = y[p];
    oz[i] = z[p];
}

// Packed Morton-order records for a caller-supplied particle order; an
// index outside the particle array is code 3 (barnes_hut.cu:175-179).
__global__ __launch_bounds__(256) void pack_records_kernel(
    const uint32_t* __restrict__ idx, const double* __restrict__ m,
    const double* __restrict__ x, const double* __restrict__ y,
    const double* __restrict__ z, size_t n, bool morton,
    double4* __restrict__ rec, int* err)
{
    const size_t i = static_cast<size_t>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (i >= n)
        return;
    const uint32_t p = idx[i];
    if (p >= n) {
        atomicCAS(err, 0, 3);
        rec[i] = make_double4(0.0, 0.0, 0.0, 0.0);
        return;
    }
    const size_t s = morton ? i : p;
    rec[i] = make_double4(x[s], y[s], z[s], m[s]);
}

} // namespace

void launch_kick_drift(hipStream_t s, double* x, double* y, double* z,
                       double* vx, double* vy, double* vz, const double* ax,
                       const double* ay, const double* az, size_t n, double dt)
{
    kick_drift_kernel<<<blocks_for(n, 256), 256, 0, s>>>(x, y, z, vx, vy, vz,
                                                         ax, ay, az, n, dt);
    NBH_HIP(hipGetLastError());
}

void launch_kick(hipStream_t s, double* vx, double* vy, double* vz,
                 const double* ax, const double* ay, const double* az,
                 size_t n, double dt)
{
    kick_kernel<<<blocks_for(n, 256), 256, 0, s>>>(vx, vy, vz, ax, ay, az, n,
                                                   dt);
    NBH_HIP(hipGetLastError());
}

void launch_gather4(hipStream_t s, const uint32_t* idx, const double* m,
                    const double* x, const double* y, const double* z,
                    double* om, double* ox, double* oy, double* oz, size_t n)
{
    gather4_kernel<<<blocks_for(n, 256), 256, 0, s>>>(idx, m, x, y, z, om, ox,
                                                      oy, oz, n);
    NBH_HIP(hipGetLastError());
}

void launch_pack_records(hipStream_t s, const uint32_t* idx, const double* m,
                         const double* x, const double* y, const double* z,
                         size_t n, bool morton, double4* rec, int* err)
{
    pack_records_kernel<<<blocks_for(n, 256), 256, 0, s>>>(idx, m, x, y, z, n,
                                                           morton, rec, err);
    NBH_HIP(hipGetLastError());
}

} // namespace nbh
