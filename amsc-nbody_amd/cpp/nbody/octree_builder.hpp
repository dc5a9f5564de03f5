// octree_builder.hpp -- sparse octree topology from the radix tree and its
// plan (src/octree_builder.hpp:8-12).
#pragma once

#include <cstddef>
#include <stdexcept>

#include "morton_leaf_groups.hpp"
#include "nbody_amd_shim.hpp"
#include "octree_types.hpp"
#include "radix_to_octree.hpp"
#include "tree_types.hpp"

inline void allocateOctreeTopology(Octree& octree, int nNodes)
{
    if (nNodes <= 0)
        throw std::invalid_argument("Octree node count must be positive");
    if (octree.children.data() != nullptr)
        throw std::logic_error("Octree memory is already allocated");
    const auto n = static_cast<std::size_t>(nNodes);
    octree.children.allocate(8 * n);
    octree.parent.allocate(n);
    octree.level.allocate(n);
    octree.prefix.allocate(n);
    octree.firstParticle.allocate(n);
    octree.particleCount.allocate(n);
    octree.nNodes = nNodes;
}

inline void buildSparseOctreeTopology(Octree& octree, const Tree& radixTree,
                                      const MortonLeafGroups& groups,
                                      const RadixToOctreePlan& plan)
{
    if (groups.nGroups <= 0 || groups.nParticles <= 0)
        throw std::invalid_argument("Morton groups are empty");
    if (groups.nGroups != radixTree.nLeaves)
        throw std::invalid_argument("Morton groups and radix tree have different leaf counts");
    if (groups.uniqueKeys.data() == nullptr || groups.firstParticle.data() == nullptr ||
        groups.particleCount.data() == nullptr)
        throw std::logic_error("Morton group arrays are not allocated");
    if (plan.nRadixNodes != 2 * radixTree.nLeaves - 1)
        throw std::invalid_argument("Radix-to-octree plan does not match the radix tree");
    if (plan.nOctreeNodes <= 0)
        throw std::invalid_argument("Radix-to-octree plan has no octree nodes");
    if (octree.nNodes != plan.nOctreeNodes)
        throw std::invalid_argument("Allocated octree size does not match the construction plan");
    if (octree.children.data() == nullptr)
        throw std::logic_error("Octree topology memory is not allocated");
    if (radixTree.parent.data() == nullptr)
        throw std::logic_error("Radix parent array is not allocated");
    if (plan.radixLevel.data() == nullptr)
        throw std::logic_error("Radix-to-octree plan arrays are not allocated");
    const nbh_octree_view v = nbody_amd::octree_abi_view(octree);
    const nbh_tree_view t = nbody_amd::tree_abi_view(radixTree);
    nbody_amd::check(nbh_octree_build_topology(
        nbody_amd::default_context(), &v, &t, groups.uniqueKeys.data(),
        groups.firstParticle.data(), groups.particleCount.data(), groups.nGroups,
        groups.nParticles, kMortonBitsPerAxis, plan.radixLevel.data(),
        plan.edgeNodeCount.data(), plan.edgeNodeOffset.data()));
    octree.nParticles = groups.nParticles;
    octree.nLeaves = groups.nGroups;
    octree.root = 0;
}

inline void freeOctree(Octree& octree) noexcept
{
    octree.children.reset();
    octree.parent.reset();
    octree.level.reset();
    octree.prefix.reset();
    octree.firstParticle.reset();
    octree.particleCount.reset();
    octree.mass.reset();
    octree.comX.reset();
    octree.comY.reset();
    octree.comZ.reset();
    octree.centerX.reset();
    octree.centerY.reset();
    octree.centerZ.reset();
    octree.halfSize.reset();
    octree.pendingChildren.reset();
    octree.childCount.reset();
    octree.nParticles = 0;
    octree.nLeaves = 0;
    octree.nNodes = 0;
    octree.root = -1;
}
