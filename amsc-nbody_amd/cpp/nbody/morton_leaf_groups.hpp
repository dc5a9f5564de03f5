// morton_leaf_groups.hpp -- MortonLeafGroups (src/morton_leaf_groups.hpp:8-47):
// one group per run of equal sorted keys -- the unique key, the first sorted
// position and the particle count of the run.
#pragma once

#include <cstddef>
#include <cstdint>
#include <stdexcept>

#include "device_buffer.hpp"
#include "nbody_amd_shim.hpp"

struct MortonLeafGroupsView {
    int nParticles = 0;
    int nGroups = 0;
    int capacity = 0;
    const std::uint32_t* uniqueKeys = nullptr;
    const int* firstParticle = nullptr;
    const int* particleCount = nullptr;
};

struct MortonLeafGroups {
    int nParticles = 0;
    int nGroups = 0;
    int capacity = 0;

    DeviceBuffer<std::uint32_t> uniqueKeys;
    DeviceBuffer<int> firstParticle;
    DeviceBuffer<int> particleCount;
    // Kept for layout compatibility; the group count is returned by the
    // library and no temporary storage is needed here.
    DeviceBuffer<int> numberOfGroupsDevice;
    DeviceBuffer<std::byte> temporaryStorage;

    [[nodiscard]] MortonLeafGroupsView view() const noexcept
    {
        return {nParticles, nGroups, capacity, uniqueKeys.data(),
                firstParticle.data(), particleCount.data()};
    }
};

inline void allocateMortonLeafGroups(MortonLeafGroups& groups, int capacity)
{
    if (capacity <= 0)
        throw std::invalid_argument("MortonLeafGroups capacity must be positive");
    if (groups.uniqueKeys.data() != nullptr)
        throw std::logic_error("MortonLeafGroups memory is already allocated");
    const auto n = static_cast<std::size_t>(capacity);
    groups.uniqueKeys.allocate(n);
    groups.firstParticle.allocate(n);
    groups.particleCount.allocate(n);
    groups.capacity = capacity;
    groups.nParticles = 0;
    groups.nGroups = 0;
}

inline void buildMortonLeafGroups(MortonLeafGroups& groups,
                                  const std::uint32_t* d_sortedKeys, int nParticles)
{
    if (nParticles < 1)
        throw std::invalid_argument("buildMortonLeafGroups requires nParticles >= 1");
    if (nParticles > groups.capacity)
        throw std::invalid_argument("nParticles exceeds MortonLeafGroups capacity");
    if (d_sortedKeys == nullptr)
        throw std::invalid_argument("d_sortedKeys must not be null");
    if (groups.uniqueKeys.data() == nullptr)
        throw std::logic_error("MortonLeafGroups memory has not been allocated");
    groups.nParticles = 0;
    groups.nGroups = 0;
    int n_groups = 0;
    nbody_amd::check(nbh_leaf_groups_build(nbody_amd::default_context(), d_sortedKeys, 4,
                                           nParticles, groups.capacity,
                                           groups.uniqueKeys.data(),
                                           groups.firstParticle.data(),
                                           groups.particleCount.data(), &n_groups));
    groups.nParticles = nParticles;
    groups.nGroups = n_groups;
}

inline void freeMortonLeafGroups(MortonLeafGroups& groups) noexcept
{
    groups.uniqueKeys.reset();
    groups.firstParticle.reset();
    groups.particleCount.reset();
    groups.numberOfGroupsDevice.reset();
    groups.temporaryStorage.reset();
    groups.nParticles = 0;
    groups.nGroups = 0;
    groups.capacity = 0;
}
