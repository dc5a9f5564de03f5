// walk_common.hpp -- definitions shared by the walk layout and the
// traversal kernels.
#pragma once

#include "device_utils.hpp"
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

constexpr int kWalkThreads = 256; // 4 independent wavefronts per block

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
