// morton.hpp -- MortonKeys (src/morton.hpp:8-49): 30-bit Morton keys (10 bits
// per axis, the reference layout) of the particles, sorted ascending, with
// the matching original indices (a stable sort: equal keys keep index order).
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <stdexcept>

#include "device_buffer.hpp"
#include "globalbounding.hpp"
#include "nbody_amd_shim.hpp"

using MortonKey = std::uint32_t;
using QuantisedCoordinate = std::uint32_t;
using ParticleIndex = std::uint32_t;

class MortonKeys {
public:
    // Coordinate pointers must each reference at least count doubles in
    // device memory.  Keys and particle indices are stored in ascending
    // Morton order.
    MortonKeys(const double* x, const double* y, const double* z, std::size_t count)
    {
        validate(x, y, z, count);
        const Bbox box(x, y, z, count);
        build(x, y, z, count, box);
    }

    MortonKeys(const double* x, const double* y, const double* z, std::size_t count,
               const Bbox& boundingBox)
    {
        validate(x, y, z, count);
        build(x, y, z, count, boundingBox);
    }

    [[nodiscard]] const MortonKey* keys_device_data() const noexcept { return keys_.data(); }
    [[nodiscard]] MortonKey* keys_device_data() noexcept { return keys_.data(); }
    [[nodiscard]] const ParticleIndex* indices_device_data() const noexcept { return indices_.data(); }
    [[nodiscard]] ParticleIndex* indices_device_data() noexcept { return indices_.data(); }
    [[nodiscard]] std::size_t size() const noexcept { return count_; }

private:
    DeviceBuffer<MortonKey> keys_;
    DeviceBuffer<ParticleIndex> indices_;
    std::size_t count_ = 0;

    static void validate(const double* x, const double* y, const double* z,
                         std::size_t count)
    {
        if (count == 0)
            throw std::invalid_argument("Cannot compute Morton keys for an empty particle set");
        if (x == nullptr || y == nullptr || z == nullptr)
            throw std::invalid_argument("MortonKeys coordinate pointers must not be null");
    }

    void build(const double* x, const double* y, const double* z, std::size_t count,
               const Bbox& boundingBox)
    {
        const BboxValues b = boundingBox.values();
        if (!std::isfinite(b.center_x) || !std::isfinite(b.center_y) ||
            !std::isfinite(b.center_z) || !std::isfinite(b.side) || b.side <= 0.0)
            throw std::invalid_argument("Morton bounding box is not valid");
        keys_.allocate(count);
        indices_.allocate(count);
        nbody_amd::check(nbh_morton_keys(nbody_amd::default_context(), x, y, z, count,
                                         boundingBox.device_data(), 10, keys_.data(),
                                         indices_.data()));
        count_ = count;
    }
};
