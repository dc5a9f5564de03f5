// octree_device.hpp -- device helpers shared by the octree kernels.
#pragma once

#include "nbh_internal.hpp"

namespace nbh {

// mortonPrefixAtLevel / octantAtLevel (octree_builder.cu:48-65) for either
// key layout.
template <int BITS>
__device__ __forceinline__ typename Layout<BITS>::Key prefix_at(
    typename Layout<BITS>::Key key, int level)
{
    using L = Layout<BITS>;
    using Key = typename L::Key;
    if (level == 0)
        return 0;
    const int remaining = L::spatial_bits - level * 3;
    const Key mask = L::spatial_mask & (~Key(0) << remaining);
    return key & mask;
}

template <int BITS>
__device__ __forceinline__ int octant_at(typename Layout<BITS>::Key key,
                                         int level)
{
    const int shift = Layout<BITS>::spatial_bits - level * 3;
    return static_cast<int>((key >> shift) & 0x7u);
}

// Cell centre and half size from the prefix path (octree_geometry.cu:39-102):
// validate level and prefix, then from the root centre halve the size per
// digit and step +-half along x/y/z by octant bits 4/2/1.  code = 1 (level)
// or 2 (prefix) on failure.
template <int BITS>
__device__ __forceinline__ bool node_geometry(typename Layout<BITS>::Key p,
                                              int lv, double rcx, double rcy,
                                              double rcz, double side,
                                              double& cx, double& cy,
                                              double& cz, double& half,
                                              int& code)
{
    using L = Layout<BITS>;
    using Key = typename L::Key;
    if (lv < 0 || lv > L::bits) {
        code = 1;
        return false;
    }
    bool valid = (p & ~L::spatial_mask) == 0;
    if (valid) {
        if (lv == 0) {
            valid = (p == 0);
        } else {
            const int remaining = L::spatial_bits - lv * 3;
            const Key mask = L::spatial_mask & (~Key(0) << remaining);
            valid = (p & ~mask) == 0;
        }
    }
    if (!valid) {
        code = 2;
        return false;
    }
    double x = rcx, y = rcy, z = rcz, h = side * 0.5;
    for (int depth = 1; depth <= lv; ++depth) {
        const int shift = L::spatial_bits - depth * 3;
        const int oct = static_cast<int>((p >> shift) & 0x7u);
        const double ch = h * 0.5;
        x += (oct & 4) ? ch : -ch;
        y += (oct & 2) ? ch : -ch;
        z += (oct & 1) ? ch : -ch;
        h = ch;
    }
    cx = x;
    cy = y;
    cz = z;
    half = h;
    return true;
}

} // namespace nbh
