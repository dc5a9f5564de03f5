// octree_geometry.hpp -- cell centre and half size of every octree node from
// its Morton prefix (src/octree_geometry.hpp:6-10).
#pragma once

#include <cmath>
#include <cstddef>
#include <stdexcept>

#include "globalbounding.hpp"
#include "nbody_amd_shim.hpp"
#include "octree_types.hpp"
#include "radix_to_octree.hpp"

inline void allocateOctreeGeometry(Octree& octree)
{
    if (octree.nNodes <= 0)
        throw std::invalid_argument("Octree topology must be allocated before geometry");
    if (octree.level.data() == nullptr || octree.prefix.data() == nullptr)
        throw std::logic_error("Octree topology arrays required for geometry are not allocated");
    if (octree.centerX.data() != nullptr)
        throw std::logic_error("Octree geometry is already allocated");
    const auto n = static_cast<std::size_t>(octree.nNodes);
    octree.centerX.allocate(n);
    octree.centerY.allocate(n);
    octree.centerZ.allocate(n);
    octree.halfSize.allocate(n);
}

inline void computeOctreeGeometry(Octree& octree, const Bbox& boundingBox)
{
    if (octree.nNodes <= 0 || octree.root != 0)
        throw std::logic_error("Octree topology is not valid");
    if (octree.level.data() == nullptr || octree.prefix.data() == nullptr)
        throw std::logic_error("Octree topology arrays required for geometry are not allocated");
    if (octree.centerX.data() == nullptr)
        throw std::logic_error("Octree geometry arrays are not allocated");
    const BboxValues b = boundingBox.values();
    if (!std::isfinite(b.center_x) || !std::isfinite(b.center_y) ||
        !std::isfinite(b.center_z) || !std::isfinite(b.side) || b.side <= 0.0)
        throw std::invalid_argument("Bounding box geometry is not valid");
    const double box[4] = {b.center_x, b.center_y, b.center_z, b.side};
    const nbh_octree_view v = nbody_amd::octree_abi_view(octree);
    nbody_amd::check(nbh_octree_geometry(nbody_amd::default_context(), &v,
                                         kMortonBitsPerAxis, box));
}
