// tree_builder.hpp -- allocation, Karras construction and the legacy
// binary-tree centre of mass (src/tree_builder.hpp:8-22).
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>

#include "morton_leaf_groups.hpp"
#include "nbody_amd_shim.hpp"
#include "tree_types.hpp"

inline void allocateTreeTopology(Tree& tree, int nLeaves)
{
    if (nLeaves <= 0)
        throw std::invalid_argument("N must be positive");
    if (tree.parent.data() != nullptr)
        throw std::logic_error("Tree memory is already allocated");
    const auto n = static_cast<std::size_t>(nLeaves);
    if (n > 1) {
        tree.left.allocate(n - 1);
        tree.right.allocate(n - 1);
        tree.rangeFirst.allocate(n - 1);
        tree.rangeLast.allocate(n - 1);
        tree.prefixLength.allocate(n - 1);
    }
    tree.parent.allocate(2 * n - 1);
    tree.nLeaves = nLeaves;
    tree.nInternalNodes = nLeaves - 1;
}

inline void allocateTree(Tree& tree, int nLeaves)
{
    allocateTreeTopology(tree, nLeaves);
    const auto n = static_cast<std::size_t>(nLeaves);
    if (n > 1)
        tree.visitCount.allocate(n - 1);
    tree.mass.allocate(2 * n - 1);
    tree.comX.allocate(2 * n - 1);
    tree.comY.allocate(2 * n - 1);
    tree.comZ.allocate(2 * n - 1);
}

inline void buildTree(Tree& tree, const std::uint32_t* d_sortedKeys, int nLeaves)
{
    if (nLeaves < 1)
        throw std::invalid_argument("buildTree requires N >= 1");
    if (nLeaves != tree.nLeaves)
        throw std::invalid_argument("buildTree N does not match allocated tree size");
    if (tree.parent.data() == nullptr)
        throw std::logic_error("Tree memory has not been allocated");
    if (d_sortedKeys == nullptr)
        throw std::invalid_argument("d_sortedKeys must not be null");
    const nbh_tree_view v = nbody_amd::tree_abi_view(tree);
    nbody_amd::check(nbh_tree_build(nbody_amd::default_context(), &v, d_sortedKeys, 4,
                                    nLeaves));
}

inline void buildTreeFromMortonGroups(Tree& tree, const MortonLeafGroups& groups)
{
    if (groups.nParticles <= 0)
        throw std::invalid_argument("MortonLeafGroups has no particles");
    if (groups.nGroups <= 0)
        throw std::invalid_argument("MortonLeafGroups has no groups");
    if (groups.nGroups > groups.nParticles || groups.nGroups > groups.capacity)
        throw std::logic_error("MortonLeafGroups metadata is inconsistent");
    if (groups.uniqueKeys.data() == nullptr)
        throw std::logic_error("MortonLeafGroups uniqueKeys is not allocated");
    if (tree.nLeaves != groups.nGroups)
        throw std::invalid_argument("Tree leaf count does not match Morton group count");
    buildTree(tree, groups.uniqueKeys.data(), groups.nGroups);
}

// One particle per leaf (the tests' legacy path).
inline void computeCenterOfMass(Tree& tree, const std::uint32_t* d_sortedIndices,
                                const double* d_mass, const double* d_positionX,
                                const double* d_positionY, const double* d_positionZ,
                                int nLeaves)
{
    if (nLeaves < 1)
        throw std::invalid_argument("computeCentersOfMass requires N >= 1");
    if (nLeaves != tree.nLeaves)
        throw std::invalid_argument("computeCentersOfMass N does not match tree size");
    if (tree.mass.data() == nullptr)
        throw std::logic_error("Tree memory has not been fully allocated");
    const nbh_tree_view v = nbody_amd::tree_abi_view(tree);
    nbody_amd::check(nbh_tree_center_of_mass(nbody_amd::default_context(), &v, nullptr,
                                             nullptr, d_sortedIndices, d_mass,
                                             d_positionX, d_positionY, d_positionZ));
}

// One Morton group (one or more particles) per leaf.
inline void computeGroupedCenterOfMass(Tree& tree, const MortonLeafGroups& groups,
                                       const std::uint32_t* d_sortedIndices,
                                       const double* d_mass, const double* d_positionX,
                                       const double* d_positionY,
                                       const double* d_positionZ)
{
    if (groups.nParticles <= 0)
        throw std::invalid_argument("computeGroupedCenterOfMass requires at least one particle");
    if (groups.nGroups <= 0)
        throw std::invalid_argument("computeGroupedCenterOfMass requires at least one Morton group");
    if (tree.nLeaves != groups.nGroups)
        throw std::invalid_argument("Tree leaf count does not match Morton group count");
    if (groups.firstParticle.data() == nullptr || groups.particleCount.data() == nullptr)
        throw std::logic_error("MortonLeafGroups arrays are not allocated");
    if (tree.mass.data() == nullptr)
        throw std::logic_error("Tree memory has not been fully allocated");
    const nbh_tree_view v = nbody_amd::tree_abi_view(tree);
    nbody_amd::check(nbh_tree_center_of_mass(nbody_amd::default_context(), &v,
                                             groups.firstParticle.data(),
                                             groups.particleCount.data(), d_sortedIndices,
                                             d_mass, d_positionX, d_positionY, d_positionZ));
}

inline void freeTree(Tree& tree) noexcept
{
    tree.left.reset();
    tree.right.reset();
    tree.parent.reset();
    tree.rangeFirst.reset();
    tree.rangeLast.reset();
    tree.prefixLength.reset();
    tree.mass.reset();
    tree.comX.reset();
    tree.comY.reset();
    tree.comZ.reset();
    tree.visitCount.reset();
    tree.nLeaves = 0;
    tree.nInternalNodes = 0;
}
