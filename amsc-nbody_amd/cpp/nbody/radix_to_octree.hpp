// radix_to_octree.hpp -- the radix-tree -> octree plan
// (src/radix_to_octree.hpp:10-71): octree level of every radix node, the
// number of octree nodes each radix edge spawns and their exclusive scan.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>

#include "device_buffer.hpp"
#include "morton_leaf_groups.hpp"
#include "nbody_amd_shim.hpp"
#include "tree_types.hpp"

constexpr int kMortonDimensions = 3;
constexpr int kMortonBitsPerAxis = 10;
constexpr int kMortonSpatialBits = kMortonDimensions * kMortonBitsPerAxis;
constexpr int kMortonStorageBits = 32;
constexpr int kMortonPaddingBits = kMortonStorageBits - kMortonSpatialBits;
constexpr int kOctreeMaxLevel = kMortonBitsPerAxis;
constexpr std::uint32_t kMortonSpatialMask = 0x3FFFFFFFu;

struct RadixToOctreePlanView {
    int nRadixNodes = 0;
    int nOctreeNodes = 0;
    const int* radixLevel = nullptr;
    const int* edgeNodeCount = nullptr;
    const int* edgeNodeOffset = nullptr;
};

struct RadixToOctreePlan {
    int nRadixNodes = 0;  // 2 * nLeaves - 1
    int nOctreeNodes = 0; // including the root at index 0
    DeviceBuffer<int> radixLevel;
    DeviceBuffer<int> edgeNodeCount;
    DeviceBuffer<int> edgeNodeOffset;
    // Kept for layout compatibility (errors travel through the library).
    DeviceBuffer<int> errorCodeDevice;
    DeviceBuffer<std::byte> temporaryStorage;

    [[nodiscard]] RadixToOctreePlanView view() const noexcept
    {
        return {nRadixNodes, nOctreeNodes, radixLevel.data(), edgeNodeCount.data(),
                edgeNodeOffset.data()};
    }
};

inline void allocateRadixToOctreePlan(RadixToOctreePlan& plan, int nRadixNodes)
{
    if (nRadixNodes <= 0)
        throw std::invalid_argument("nRadixNodes must be positive");
    if (plan.radixLevel.data() != nullptr)
        throw std::logic_error("RadixToOctreePlan memory is already allocated");
    const auto n = static_cast<std::size_t>(nRadixNodes);
    plan.radixLevel.allocate(n);
    plan.edgeNodeCount.allocate(n);
    plan.edgeNodeOffset.allocate(n);
    plan.nRadixNodes = nRadixNodes;
    plan.nOctreeNodes = 0;
}

inline void buildRadixToOctreePlan(RadixToOctreePlan& plan, const Tree& radixTree,
                                   const MortonLeafGroups& groups)
{
    if (radixTree.nLeaves <= 0)
        throw std::invalid_argument("The radix tree has no leaves");
    if (groups.nGroups != radixTree.nLeaves)
        throw std::invalid_argument("Morton group count does not match radix-tree leaf count");
    if (groups.uniqueKeys.data() == nullptr)
        throw std::logic_error("Morton unique keys are not allocated");
    if (radixTree.parent.data() == nullptr)
        throw std::logic_error("Radix-tree parent array is not allocated");
    if (radixTree.nLeaves > 1 && radixTree.prefixLength.data() == nullptr)
        throw std::logic_error("Radix-tree prefix metadata is not allocated");
    if (plan.nRadixNodes != 2 * radixTree.nLeaves - 1)
        throw std::invalid_argument("RadixToOctreePlan size does not match the radix tree");
    if (plan.radixLevel.data() == nullptr)
        throw std::logic_error("RadixToOctreePlan memory is not allocated");
    plan.nOctreeNodes = 0;
    const nbh_tree_view v = nbody_amd::tree_abi_view(radixTree);
    int n_octree = 0;
    nbody_amd::check(nbh_radix_to_octree_plan(
        nbody_amd::default_context(), &v, groups.uniqueKeys.data(), kMortonBitsPerAxis,
        plan.radixLevel.data(), plan.edgeNodeCount.data(), plan.edgeNodeOffset.data(),
        &n_octree));
    plan.nOctreeNodes = n_octree;
}

inline void freeRadixToOctreePlan(RadixToOctreePlan& plan) noexcept
{
    plan.radixLevel.reset();
    plan.edgeNodeCount.reset();
    plan.edgeNodeOffset.reset();
    plan.errorCodeDevice.reset();
    plan.temporaryStorage.reset();
    plan.nRadixNodes = 0;
    plan.nOctreeNodes = 0;
}
