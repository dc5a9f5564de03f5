// octree_types.hpp -- the sparse octree (src/octree_types.hpp:7-178): node 0 is
// the root, children are node-major (children[node * 8 + octant]), a leaf
// holds the sorted particle range [firstParticle, firstParticle +
// particleCount).
#pragma once

#include <cstdint>

#include "device_buffer.hpp"
#include "nbody_amd_shim.hpp"

struct OctreeView {
    int nParticles = 0;
    int nLeaves = 0;
    int nNodes = 0;
    int root = -1;
    int* children = nullptr;
    int* parent = nullptr;
    int* level = nullptr;
    std::uint32_t* prefix = nullptr;
    int* firstParticle = nullptr;
    int* particleCount = nullptr;
    double* mass = nullptr;
    double* comX = nullptr;
    double* comY = nullptr;
    double* comZ = nullptr;
    double* centerX = nullptr;
    double* centerY = nullptr;
    double* centerZ = nullptr;
    double* halfSize = nullptr;
    int* pendingChildren = nullptr;
    int* childCount = nullptr;
};

struct ConstOctreeView {
    int nParticles = 0;
    int nLeaves = 0;
    int nNodes = 0;
    int root = -1;
    const int* children = nullptr;
    const int* parent = nullptr;
    const int* level = nullptr;
    const std::uint32_t* prefix = nullptr;
    const int* firstParticle = nullptr;
    const int* particleCount = nullptr;
    const double* mass = nullptr;
    const double* comX = nullptr;
    const double* comY = nullptr;
    const double* comZ = nullptr;
    const double* centerX = nullptr;
    const double* centerY = nullptr;
    const double* centerZ = nullptr;
    const double* halfSize = nullptr;
    const int* pendingChildren = nullptr;
    const int* childCount = nullptr;
};

struct Octree {
    int nParticles = 0;
    int nLeaves = 0;
    int nNodes = 0;
    int root = -1;

    DeviceBuffer<int> children;
    DeviceBuffer<int> parent;
    DeviceBuffer<int> level;
    DeviceBuffer<std::uint32_t> prefix;
    DeviceBuffer<int> firstParticle;
    DeviceBuffer<int> particleCount;
    DeviceBuffer<double> mass;
    DeviceBuffer<double> comX;
    DeviceBuffer<double> comY;
    DeviceBuffer<double> comZ;
    DeviceBuffer<double> centerX;
    DeviceBuffer<double> centerY;
    DeviceBuffer<double> centerZ;
    DeviceBuffer<double> halfSize;
    DeviceBuffer<int> pendingChildren;
    DeviceBuffer<int> childCount;

    [[nodiscard]] OctreeView view() noexcept
    {
        return {nParticles, nLeaves, nNodes, root, children.data(), parent.data(),
                level.data(), prefix.data(), firstParticle.data(), particleCount.data(),
                mass.data(), comX.data(), comY.data(), comZ.data(), centerX.data(),
                centerY.data(), centerZ.data(), halfSize.data(), pendingChildren.data(),
                childCount.data()};
    }

    [[nodiscard]] ConstOctreeView view() const noexcept
    {
        return {nParticles, nLeaves, nNodes, root, children.data(), parent.data(),
                level.data(), prefix.data(), firstParticle.data(), particleCount.data(),
                mass.data(), comX.data(), comY.data(), comZ.data(), centerX.data(),
                centerY.data(), centerZ.data(), halfSize.data(), pendingChildren.data(),
                childCount.data()};
    }
};

namespace nbody_amd {

inline nbh_octree_view octree_abi_view(const Octree& o)
{
    nbh_octree_view v{};
    v.n_particles = o.nParticles;
    v.n_leaves = o.nLeaves;
    v.n_nodes = o.nNodes;
    v.root = o.root;
    v.children = const_cast<int*>(o.children.data());
    v.parent = const_cast<int*>(o.parent.data());
    v.level = const_cast<int*>(o.level.data());
    v.prefix = const_cast<std::uint32_t*>(o.prefix.data());
    v.first_particle = const_cast<int*>(o.firstParticle.data());
    v.particle_count = const_cast<int*>(o.particleCount.data());
    v.mass = const_cast<double*>(o.mass.data());
    v.com_x = const_cast<double*>(o.comX.data());
    v.com_y = const_cast<double*>(o.comY.data());
    v.com_z = const_cast<double*>(o.comZ.data());
    v.center_x = const_cast<double*>(o.centerX.data());
    v.center_y = const_cast<double*>(o.centerY.data());
    v.center_z = const_cast<double*>(o.centerZ.data());
    v.half_size = const_cast<double*>(o.halfSize.data());
    v.pending_children = const_cast<int*>(o.pendingChildren.data());
    v.child_count = const_cast<int*>(o.childCount.data());
    return v;
}

} // namespace nbody_amd
