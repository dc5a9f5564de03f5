// barnes_hut.hpp -- computeBarnesHutAcceleration (src/barnes_hut.hpp:8-21):
// FP64 tree walk with the reference's opening test and exact leaves;
// accelerations are written in original particle order.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>

#include "nbody_amd_shim.hpp"
#include "octree_types.hpp"

struct BarnesHutParameters {
    double theta = 0.5;
    double softening = 1.0e-6;
    double gravitationalConstant = 6.67430e-11;
};

enum class BarnesHutParticleOrder { original, morton };

inline void computeBarnesHutAcceleration(
    const Octree& octree, const std::uint32_t* d_sortedIndices, const double* d_mass,
    const double* d_positionX, const double* d_positionY, const double* d_positionZ,
    double* d_accelerationX, double* d_accelerationY, double* d_accelerationZ,
    std::size_t particleCount, const BarnesHutParameters& parameters,
    BarnesHutParticleOrder particleOrder = BarnesHutParticleOrder::original)
{
    if (octree.nNodes <= 0 || octree.root != 0)
        throw std::logic_error("Octree is not valid for Barnes-Hut traversal");
    if (particleCount != static_cast<std::size_t>(octree.nParticles))
        throw std::invalid_argument("Particle count does not match the octree");
    if (octree.children.data() == nullptr || octree.firstParticle.data() == nullptr)
        throw std::logic_error("Octree topology arrays required for traversal are not allocated");
    if (octree.mass.data() == nullptr)
        throw std::logic_error("Octree physical data is not allocated");
    if (octree.centerX.data() == nullptr)
        throw std::logic_error("Octree geometry is not allocated");
    int order = NBH_ORDER_ORIGINAL;
    if (particleOrder == BarnesHutParticleOrder::morton)
        order = NBH_ORDER_MORTON;
    else if (particleOrder != BarnesHutParticleOrder::original)
        throw std::invalid_argument("Unknown Barnes-Hut particle order");
    const nbh_octree_view v = nbody_amd::octree_abi_view(octree);
    nbody_amd::check(nbh_bh_acceleration(
        nbody_amd::default_context(), &v, d_sortedIndices, d_mass, d_positionX,
        d_positionY, d_positionZ, d_accelerationX, d_accelerationY, d_accelerationZ,
        particleCount, parameters.theta, parameters.softening,
        parameters.gravitationalConstant, order));
}
