// particle.hpp -- Particles (src/particle.hpp:85-226): the FP64 structure of
// arrays of mass, position, velocity and type in device memory, loaded from
// the planar binary file format (particle_file.hpp) or, as an addition, from
// host arrays.
#pragma once

#include <hip/hip_runtime_api.h>

#include <cstddef>
#include <cstdint>
#include <fstream>
#include <stdexcept>
#include <utility>
#include <vector>

#include "device_buffer.hpp"
#include "nbody_amd_shim.hpp"
#include "particle_file.hpp"
#include "particle_type.hpp"

inline constexpr double G = 6.67430e-11; // gravitational constant
inline constexpr double K = 8.98755e9;   // Coulomb constant

struct DeviceParticlesView {
    std::size_t count;
    double* mass;
    double* x;
    double* y;
    double* z;
    double* vx;
    double* vy;
    double* vz;
    std::uint8_t* type;
};

struct ConstDeviceParticlesView {
    std::size_t count;
    const double* mass;
    const double* x;
    const double* y;
    const double* z;
    const double* vx;
    const double* vy;
    const double* vz;
    const std::uint8_t* type;
};

class Particles {
public:
    explicit Particles(std::ifstream& inFile)
    {
        if (!inFile.is_open() || !inFile.good())
            throw std::runtime_error("Invalid or unreadable file stream provided.");
        upload(nbody_amd::read_particle_file(inFile));
    }

    // From host arrays of equal length; types default to Unknown.
    Particles(const std::vector<double>& mass, const std::vector<double>& x,
              const std::vector<double>& y, const std::vector<double>& z,
              const std::vector<double>& vx, const std::vector<double>& vy,
              const std::vector<double>& vz, const std::vector<std::uint8_t>& type = {})
    {
        const std::size_t n = mass.size();
        if (n == 0)
            throw std::runtime_error("Particle count is zero.");
        if (x.size() != n || y.size() != n || z.size() != n || vx.size() != n ||
            vy.size() != n || vz.size() != n || (!type.empty() && type.size() != n))
            throw std::invalid_argument("Particle arrays must have equal lengths");
        nbody_amd::ParticleFileData d{mass, x, y, z, vx, vy, vz, type};
        if (d.type.empty())
            d.type.assign(n, static_cast<std::uint8_t>(ParticleType::Unknown));
        upload(d);
    }

    Particles(Particles&&) noexcept = default;
    Particles& operator=(Particles&&) noexcept = default;

    [[nodiscard]] DeviceParticlesView device_view() noexcept
    {
        return {count_, mass_.data(), x_.data(), y_.data(), z_.data(),
                vx_.data(), vy_.data(), vz_.data(), type_.data()};
    }

    [[nodiscard]] ConstDeviceParticlesView device_view() const noexcept
    {
        return {count_, mass_.data(), x_.data(), y_.data(), z_.data(),
                vx_.data(), vy_.data(), vz_.data(), type_.data()};
    }

    // Not in the reference: hands the type codes over to a new owner (the
    // Simulation, which copies the FP64 state into its own arrays and keeps
    // only the types of the set it was built from).
    [[nodiscard]] DeviceBuffer<std::uint8_t> release_types() noexcept
    {
        return std::move(type_);
    }

private:
    std::size_t count_ = 0;
    DeviceBuffer<double> mass_, x_, y_, z_, vx_, vy_, vz_;
    DeviceBuffer<std::uint8_t> type_;

    void upload(const nbody_amd::ParticleFileData& d)
    {
        const std::size_t n = d.size();
        const std::vector<double>* src[7] = {&d.mass, &d.x, &d.y, &d.z, &d.vx, &d.vy, &d.vz};
        DeviceBuffer<double>* dst[7] = {&mass_, &x_, &y_, &z_, &vx_, &vy_, &vz_};
        for (int k = 0; k < 7; ++k) {
            dst[k]->allocate(n);
            nbody_amd::check_hip(hipMemcpy(dst[k]->data(), src[k]->data(),
                                           n * sizeof(double), hipMemcpyHostToDevice),
                                 "Particles upload");
        }
        type_.allocate(n);
        nbody_amd::check_hip(hipMemcpy(type_.data(), d.type.data(), n,
                                       hipMemcpyHostToDevice),
                             "Particles upload");
        count_ = n;
    }
};
