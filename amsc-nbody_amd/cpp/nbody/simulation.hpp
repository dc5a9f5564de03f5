// simulation.hpp -- Simulation and SimulationConfig (src/simulation.hpp:18-66):
// the Barnes-Hut leapfrog loop over the particles it owns.
//
// SimulationConfig keeps the reference's three fields; the extra fields have
// defaults that reproduce the reference's choices (FP64, 30-bit Morton keys,
// the SI gravitational constant) and select the mixed-precision traversal or
// 63-bit keys when set.  The step runs on one stream with two host
// synchronisations; particles() views the simulation's own device arrays,
// which the steps update in place.
#pragma once

#include <cstddef>
#include <stdexcept>
#include <utility>

#include "nbody_amd_shim.hpp"
#include "particle.hpp"

struct SimulationConfig {
    double timeStep;
    double theta;
    double softening;
    // additions (defaults = reference behaviour)
    double gravitationalConstant = G;
    int precision = NBH_PRECISION_FP64; // or NBH_PRECISION_MIXED
    int bitsPerAxis = 10;                // 10 (30-bit keys) or 21 (63-bit keys)
};

class Simulation {
public:
    // Copies the FP64 state of the moved-in set into the simulation's own
    // device arrays and frees it, keeping only the type codes: the
    // simulation holds one device copy of the state.
    Simulation(Particles particles, SimulationConfig config) : config_(config)
    {
        const ConstDeviceParticlesView p = std::as_const(particles).device_view();
        nbh_sim_config c{};
        c.time_step = config.timeStep;
        c.theta = config.theta;
        c.softening = config.softening;
        c.gravitational_constant = config.gravitationalConstant;
        c.precision = config.precision;
        c.bits_per_axis = config.bitsPerAxis;
        nbody_amd::check(nbh_sim_create_device(&sim_, &c, p.count, p.mass, p.x, p.y, p.z,
                                               p.vx, p.vy, p.vz, nullptr));
        types_ = particles.release_types();
    }

    ~Simulation() noexcept
    {
        if (sim_ != nullptr)
            (void)nbh_sim_destroy(sim_);
    }

    Simulation(const Simulation&) = delete;
    Simulation& operator=(const Simulation&) = delete;

    void initialize() { nbody_amd::check(nbh_sim_initialize(sim_)); }
    void step() { nbody_amd::check(nbh_sim_step(sim_)); }
    void run(std::size_t numberOfSteps) { nbody_amd::check(nbh_sim_run(sim_, numberOfSteps)); }

    [[nodiscard]] double time() const noexcept { return nbh_sim_time(sim_); }
    [[nodiscard]] std::size_t stepNumber() const noexcept { return nbh_sim_step_number(sim_); }

    [[nodiscard]] DeviceParticlesView particles() noexcept
    {
        nbh_particles_view v{};
        (void)nbh_sim_particles(sim_, &v);
        return {v.count, v.mass, v.x, v.y, v.z, v.vx, v.vy, v.vz, types_.data()};
    }

    // Not in the reference: the underlying C ABI handle (phase timings,
    // interaction counts, energy, structure views).
    [[nodiscard]] nbh_sim* handle() noexcept { return sim_; }

private:
    SimulationConfig config_;
    DeviceBuffer<std::uint8_t> types_; // the type codes; the state lives in sim_
    nbh_sim* sim_ = nullptr;
};
