// octree_physics.hpp -- mass and centre of mass of every octree node
// (src/octree_physics.hpp:7-9).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>

#include "nbody_amd_shim.hpp"
#include "octree_types.hpp"

inline void allocateOctreePhysicalData(Octree& octree)
{
    if (octree.nNodes <= 0)
        throw std::invalid_argument("Octree topology must be allocated first");
    if (octree.mass.data() != nullptr)
        throw std::logic_error("Octree physical data is already allocated");
    const auto n = static_cast<std::size_t>(octree.nNodes);
    octree.mass.allocate(n);
    octree.comX.allocate(n);
    octree.comY.allocate(n);
    octree.comZ.allocate(n);
    octree.pendingChildren.allocate(n);
    octree.childCount.allocate(n);
}

inline void computeOctreeMassAndCenterOfMass(Octree& octree,
                                             const std::uint32_t* d_sortedIndices,
                                             const double* d_mass,
                                             const double* d_positionX,
                                             const double* d_positionY,
                                             const double* d_positionZ)
{
    if (octree.nNodes <= 0 || octree.root != 0)
        throw std::logic_error("Octree topology is not valid");
    if (octree.children.data() == nullptr || octree.firstParticle.data() == nullptr)
        throw std::logic_error("Octree topology arrays are not allocated");
    if (octree.mass.data() == nullptr)
        throw std::logic_error("Octree physical arrays are not allocated");
    if (d_sortedIndices == nullptr)
        throw std::invalid_argument("d_sortedIndices must not be null");
    if (d_mass == nullptr || d_positionX == nullptr || d_positionY == nullptr ||
        d_positionZ == nullptr)
        throw std::invalid_argument("Particle arrays must not be null");
    const nbh_octree_view v = nbody_amd::octree_abi_view(octree);
    nbody_amd::check(nbh_octree_mass_com(nbody_amd::default_context(), &v, d_sortedIndices,
                                         d_mass, d_positionX, d_positionY, d_positionZ));
}
