// globalbounding.hpp -- Bbox (include/globalbounding.hpp:7-36): the cubic
// bounding box [centre_x, centre_y, centre_z, side] kept in device memory.
// A default-constructed box holds side -1 ("not computed").
#pragma once

#include <array>
#include <cstddef>

#include "device_buffer.hpp"
#include "nbody_amd_shim.hpp"

struct BboxValues {
    double center_x;
    double center_y;
    double center_z;
    double side;
};

class Bbox {
public:
    static constexpr std::size_t value_count = 4;

    Bbox() : centre_and_side_(value_count)
    {
        const std::array<double, value_count> init{0.0, 0.0, 0.0, -1.0};
        nbody_amd::check_hip(hipMemcpy(centre_and_side_.data(), init.data(),
                                       sizeof(init), hipMemcpyHostToDevice),
                             "Bbox init");
    }

    // Coordinate pointers must each reference at least count doubles in
    // device memory.
    Bbox(const double* d_x, const double* d_y, const double* d_z, std::size_t count)
        : centre_and_side_(value_count)
    {
        recompute(d_x, d_y, d_z, count);
    }

    void recompute(const double* d_x, const double* d_y, const double* d_z,
                   std::size_t count)
    {
        nbody_amd::check(nbh_bbox_compute(nbody_amd::default_context(), d_x, d_y,
                                          d_z, count, centre_and_side_.data(),
                                          nullptr));
    }

    [[nodiscard]] BboxValues values() const
    {
        std::array<double, value_count> h{};
        nbody_amd::check(nbh_copy_to_host(h.data(), centre_and_side_.data(), sizeof(h)));
        return {h[0], h[1], h[2], h[3]};
    }

    [[nodiscard]] const double* device_data() const noexcept { return centre_and_side_.data(); }
    [[nodiscard]] double* device_data() noexcept { return centre_and_side_.data(); }

private:
    DeviceBuffer<double> centre_and_side_;
};
