// The reference's system tests through the C++ drop-in layer, written the
// way the reference's callers write them:
// - the HORIZONS one-day ephemeris: Particles(std::ifstream&) from the
//   epoch-A particle file, Simulation with the default configuration
//   {900 s, theta 0, softening 0} (FP64, 30-bit keys, SI G), 96 steps,
//   per-body bounds against the epoch-B file
//   (tests/integration/cuda/solar_system_ephemeris_test.cu:141-196);
// - the SI circular two-body orbit through a Simulation loaded from a
//   particle file (tests/integration/cuda/simulation_test.cu:195-283);
// - the two-body leapfrog convergence and one-period closure through the
//   staged functions Bbox ... computeBarnesHutAcceleration and the leapfrog
//   kernels (tests/integration/cuda/two_body_orbit_test.cu:101-880).
//
// Usage: system_test <dir>, where <dir> holds ephemeris_epoch_a.bin,
// ephemeris_epoch_b.bin (the reference's planar particle format) and
// ephemeris_tolerances.txt (one line per body: name, position bound in m,
// velocity bound in m/s).  Prints one line per case; exit status 0 when
// every case passes.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <fstream>
#include <functional>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "barnes_hut.hpp"
#include "globalbounding.hpp"
#include "integrator.hpp"
#include "morton.hpp"
#include "morton_leaf_groups.hpp"
#include "octree_builder.hpp"
#include "octree_geometry.hpp"
#include "octree_physics.hpp"
#include "particle_file.hpp"
#include "radix_to_octree.hpp"
#include "simulation.hpp"
#include "tree_builder.hpp"

namespace {

std::vector<double> host(const double* d, std::size_t n)
{
    std::vector<double> h(n);
    nbody_amd::check(nbh_copy_to_host(h.data(), d, n * sizeof(double)));
    return h;
}

DeviceBuffer<double> device(const std::vector<double>& h)
{
    DeviceBuffer<double> b(h.size());
    nbody_amd::check_hip(hipMemcpy(b.data(), h.data(), h.size() * sizeof(double),
                                   hipMemcpyHostToDevice),
                         "upload");
    return b;
}

struct HostState {
    std::vector<double> m, x, y, z, vx, vy, vz;
};

HostState copy_state(const DeviceParticlesView& v)
{
    return {host(v.mass, v.count), host(v.x, v.count),  host(v.y, v.count),
            host(v.z, v.count),    host(v.vx, v.count), host(v.vy, v.count),
            host(v.vz, v.count)};
}

void write_particle_file(const std::filesystem::path& path, const HostState& s)
{
    std::ofstream out(path, std::ios::binary | std::ios::trunc);
    if (!out.is_open())
        throw std::runtime_error("cannot create " + path.string());
    const auto n = static_cast<std::uint64_t>(s.m.size());
    out.write(reinterpret_cast<const char*>(&n), sizeof(n));
    for (const auto* b : {&s.m, &s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz})
        out.write(reinterpret_cast<const char*>(b->data()),
                  static_cast<std::streamsize>(b->size() * sizeof(double)));
    if (!out.good())
        throw std::runtime_error("cannot write " + path.string());
}

double norm3(double a, double b, double c) { return std::sqrt(a * a + b * b + c * c); }

// ---------------------------------------------------------- ephemeris
bool ephemeris(const std::filesystem::path& dir)
{
    struct Bound {
        std::string name;
        double pos, vel;
    };
    std::vector<Bound> bounds;
    {
        std::ifstream t(dir / "ephemeris_tolerances.txt");
        Bound b;
        while (t >> b.name >> b.pos >> b.vel)
            bounds.push_back(b);
    }
    std::ifstream ref_in(dir / "ephemeris_epoch_b.bin", std::ios::binary);
    const nbody_amd::ParticleFileData ref = nbody_amd::read_particle_file(ref_in);
    if (bounds.size() != ref.size() || bounds.empty())
        return false;

    constexpr std::size_t steps = 96;
    constexpr double duration = 86'400.0;
    std::ifstream input(dir / "ephemeris_epoch_a.bin", std::ios::binary);
    Particles particles(input);
    Simulation simulation(std::move(particles), {duration / static_cast<double>(steps), 0.0, 0.0});
    simulation.initialize();
    simulation.run(steps);
    bool ok = simulation.stepNumber() == steps && simulation.time() == duration;
    const HostState a = copy_state(simulation.particles());
    for (std::size_t i = 0; i < ref.size(); ++i) {
        const double pe = norm3(a.x[i] - ref.x[i], a.y[i] - ref.y[i], a.z[i] - ref.z[i]);
        const double ve = norm3(a.vx[i] - ref.vx[i], a.vy[i] - ref.vy[i], a.vz[i] - ref.vz[i]);
        const bool body_ok = a.m[i] == ref.mass[i] && pe <= bounds[i].pos && ve <= bounds[i].vel;
        std::printf("    %-8s position error %12.4f m (bound %g), velocity error %.3e m/s "
                    "(bound %g)%s\n",
                    bounds[i].name.c_str(), pe, bounds[i].pos, ve, bounds[i].vel,
                    body_ok ? "" : "  <-- exceeds");
        ok = ok && body_ok;
    }
    return ok;
}

// ------------------------------------------------ SI circular two-body
bool si_two_body(const std::filesystem::path& dir)
{
    constexpr double mass = 1.0e10, separation = 1.0;
    constexpr std::size_t steps = 128;
    const double v = std::sqrt(G * mass / (2.0 * separation));
    const double period = 2.0 * std::acos(-1.0) / (2.0 * v / separation);
    const HostState initial{{mass, mass}, {-0.5, 0.5}, {0.0, 0.0}, {0.0, 0.0},
                            {0.0, 0.0},   {-v, v},     {0.0, 0.0}};
    const auto path = dir / "si_two_body.bin";
    write_particle_file(path, initial);
    std::ifstream input(path, std::ios::binary);
    Particles particles(input);
    Simulation simulation(std::move(particles), {period / static_cast<double>(steps), 0.5, 0.0});
    simulation.initialize();
    const double e0 = mass * v * v - G * mass * mass / separation;
    double max_e = 0.0, max_sep = 0.0, max_com = 0.0;
    for (std::size_t k = 0; k < steps; ++k) {
        simulation.step();
        const HostState s = copy_state(simulation.particles());
        const double r = norm3(s.x[1] - s.x[0], s.y[1] - s.y[0], s.z[1] - s.z[0]);
        const double ke = 0.5 * mass *
                          (s.vx[0] * s.vx[0] + s.vy[0] * s.vy[0] + s.vz[0] * s.vz[0] +
                           s.vx[1] * s.vx[1] + s.vy[1] * s.vy[1] + s.vz[1] * s.vz[1]);
        const double pe = -G * mass * mass / r;
        max_e = std::max(max_e, std::abs((ke + pe - e0) / e0));
        max_sep = std::max(max_sep, std::abs(r - separation));
        max_com = std::max(max_com, norm3(0.5 * (s.x[0] + s.x[1]), 0.5 * (s.y[0] + s.y[1]),
                                          0.5 * (s.z[0] + s.z[1])));
    }
    const HostState f = copy_state(simulation.particles());
    std::printf("    energy %.3e, separation %.3e, COM drift %.3e\n", max_e, max_sep, max_com);
    bool ok = simulation.stepNumber() == steps &&
              std::abs(simulation.time() - period) <= period * 1.0e-13 && max_e < 5.0e-6 &&
              max_sep < 2.0e-3 && max_com < 1.0e-12;
    for (int i = 0; i < 2; ++i)
        ok = ok && std::abs(f.x[i] - initial.x[i]) <= 3e-3 &&
             std::abs(f.y[i] - initial.y[i]) <= 3e-3 &&
             std::abs(f.vx[i] - initial.vx[i]) <= 3e-3 &&
             std::abs(f.vy[i] - initial.vy[i]) <= 3e-3;
    return ok;
}

// --------------------------------- two bodies through the staged functions
// computeAcceleration of two_body_orbit_test.cu:101-220: the whole build
// and the traversal at theta 0.5, softening 0, G = 1.
void compute_acceleration(const DeviceBuffer<double>& m, const DeviceBuffer<double>& x,
                          const DeviceBuffer<double>& y, const DeviceBuffer<double>& z,
                          DeviceBuffer<double>& ax, DeviceBuffer<double>& ay,
                          DeviceBuffer<double>& az)
{
    const std::size_t n = m.size();
    const int count = static_cast<int>(n);
    const Bbox box(x.data(), y.data(), z.data(), n);
    MortonKeys morton(x.data(), y.data(), z.data(), n, box);
    MortonLeafGroups groups;
    allocateMortonLeafGroups(groups, count);
    buildMortonLeafGroups(groups, morton.keys_device_data(), count);
    Tree tree;
    allocateTreeTopology(tree, groups.nGroups);
    buildTreeFromMortonGroups(tree, groups);
    RadixToOctreePlan plan;
    allocateRadixToOctreePlan(plan, 2 * groups.nGroups - 1);
    buildRadixToOctreePlan(plan, tree, groups);
    Octree octree;
    allocateOctreeTopology(octree, plan.nOctreeNodes);
    buildSparseOctreeTopology(octree, tree, groups, plan);
    allocateOctreePhysicalData(octree);
    allocateOctreeGeometry(octree);
    computeOctreeMassAndCenterOfMass(octree, morton.indices_device_data(), m.data(), x.data(),
                                     y.data(), z.data());
    computeOctreeGeometry(octree, box);
    BarnesHutParameters parameters;
    parameters.theta = 0.5;
    parameters.softening = 0.0;
    parameters.gravitationalConstant = 1.0;
    computeBarnesHutAcceleration(octree, morton.indices_device_data(), m.data(), x.data(),
                                 y.data(), z.data(), ax.data(), ay.data(), az.data(), n,
                                 parameters);
}

struct Orbit {
    DeviceBuffer<double> m, x, y, z, vx, vy, vz, ax, ay, az;
    Orbit()
    {
        const double v = 1.0 / std::sqrt(2.0);
        m = device({1.0, 1.0});
        x = device({-0.5, 0.5});
        y = device({0.0, 0.0});
        z = device({0.0, 0.0});
        vx = device({0.0, 0.0});
        vy = device({-v, v});
        vz = device({0.0, 0.0});
        ax.allocate(2);
        ay.allocate(2);
        az.allocate(2);
        compute_acceleration(m, x, y, z, ax, ay, az);
    }
    void step(double dt)
    {
        leapfrogKickDrift(x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(),
                          ax.data(), ay.data(), az.data(), 2, dt);
        compute_acceleration(m, x, y, z, ax, ay, az);
        leapfrogKick(vx.data(), vy.data(), vz.data(), ax.data(), ay.data(), az.data(), 2, dt);
    }
    HostState state() const
    {
        return {host(m.data(), 2),  host(x.data(), 2),  host(y.data(), 2), host(z.data(), 2),
                host(vx.data(), 2), host(vy.data(), 2), host(vz.data(), 2)};
    }
};

const double kOrbitV = 1.0 / std::sqrt(2.0);
const double kPeriod = 2.0 * std::acos(-1.0) / std::sqrt(2.0);

bool leapfrog_convergence()
{
    constexpr std::array<int, 4> counts{32, 64, 128, 256};
    double prev_p = 0.0, prev_v = 0.0;
    bool ok = true;
    for (std::size_t k = 0; k < counts.size(); ++k) {
        Orbit o;
        const double dt = (kPeriod / 4.0) / counts[k];
        for (int s = 0; s < counts[k]; ++s)
            o.step(dt);
        const HostState s = o.state();
        const double pe = std::sqrt(s.x[0] * s.x[0] + (s.y[0] + 0.5) * (s.y[0] + 0.5) +
                                    s.z[0] * s.z[0] + s.x[1] * s.x[1] +
                                    (s.y[1] - 0.5) * (s.y[1] - 0.5) + s.z[1] * s.z[1]);
        const double ve = std::sqrt((s.vx[0] - kOrbitV) * (s.vx[0] - kOrbitV) +
                                    s.vy[0] * s.vy[0] + s.vz[0] * s.vz[0] +
                                    (s.vx[1] + kOrbitV) * (s.vx[1] + kOrbitV) +
                                    s.vy[1] * s.vy[1] + s.vz[1] * s.vz[1]);
        double po = 0.0, vo = 0.0;
        if (k > 0) {
            po = std::log2(prev_p / pe);
            vo = std::log2(prev_v / ve);
            ok = ok && pe < prev_p && ve < prev_v && po > 1.8 && po < 2.2 && vo > 1.8 &&
                 vo < 2.2;
        }
        std::printf("    %4d steps: position error %.6e, velocity error %.6e, orders %.4f %.4f\n",
                    counts[k], pe, ve, po, vo);
        ok = ok && std::isfinite(pe) && std::isfinite(ve);
        prev_p = pe;
        prev_v = ve;
    }
    return ok;
}

bool one_period()
{
    Orbit o;
    const HostState s0 = o.state();
    const auto energy = [](const HostState& s, double& r) {
        r = norm3(s.x[1] - s.x[0], s.y[1] - s.y[0], s.z[1] - s.z[0]);
        return 0.5 * (s.vx[0] * s.vx[0] + s.vy[0] * s.vy[0] + s.vz[0] * s.vz[0] +
                      s.vx[1] * s.vx[1] + s.vy[1] * s.vy[1] + s.vz[1] * s.vz[1]) -
               1.0 / r;
    };
    double r = 0.0;
    const double e0 = energy(s0, r);
    double max_e = 0.0, max_sep = 0.0;
    const double dt = kPeriod / 128.0;
    for (int k = 0; k < 128; ++k) {
        o.step(dt);
        const double e = energy(o.state(), r);
        max_e = std::max(max_e, std::abs((e - e0) / e0));
        max_sep = std::max(max_sep, std::abs(r - 1.0));
    }
    const HostState f = o.state();
    std::printf("    energy %.3e, separation %.3e\n", max_e, max_sep);
    bool ok = max_sep < 2.0e-3 && max_e < 5.0e-6;
    for (int i = 0; i < 2; ++i)
        ok = ok && std::abs(f.x[i] - s0.x[i]) <= 3e-3 && std::abs(f.y[i] - s0.y[i]) <= 3e-3 &&
             std::abs(f.vx[i] - s0.vx[i]) <= 4e-3 && std::abs(f.vy[i] - s0.vy[i]) <= 4e-3;
    return ok;
}

} // namespace

int main(int argc, char** argv)
{
    if (argc != 2) {
        std::fprintf(stderr, "usage: system_test <fixture directory>\n");
        return 2;
    }
    const std::filesystem::path dir(argv[1]);
    int failed = 0;
    const auto run = [&](const char* name, const std::function<bool()>& f) {
        std::printf("---- %s\n", name);
        std::fflush(stdout);
        bool ok = false;
        try {
            ok = f();
        } catch (const std::exception& e) {
            std::printf("    exception: %s\n", e.what());
        }
        std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
        std::fflush(stdout);
        failed += ok ? 0 : 1;
    };
    run("HORIZONS one-day ephemeris from a particle file", [&] { return ephemeris(dir); });
    run("SI circular two-body orbit through a file-loaded Simulation",
        [&] { return si_two_body(dir); });
    run("two-body leapfrog convergence through the staged functions", leapfrog_convergence);
    run("two-body one-period closure through the staged functions", one_period);
    std::printf("%d failed\n", failed);
    return failed == 0 ? 0 : 1;
}
