// The C++ drop-in layer on the device: reference-style callers of Bbox,
// MortonKeys, the staged tree/octree functions, computeBarnesHutAcceleration,
// the leapfrog functions and Simulation, with the exception classes the
// reference throws.  The staged pipeline's accelerations must equal those of
// the library's own Simulation (FP64, 30-bit keys) bit for bit.
// Prints one line per case; exit status 0 when every case passes.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <filesystem>
#include <functional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "barnes_hut.hpp"
#include "globalbounding.hpp"
#include "integrator.hpp"
#include "morton.hpp"
#include "morton_leaf_groups.hpp"
#include "octree_builder.hpp"
#include "octree_geometry.hpp"
#include "octree_physics.hpp"
#include "radix_to_octree.hpp"
#include "simulation.hpp"
#include "streaming/device_capture.hpp"
#include "tree_builder.hpp"

namespace {

template <class T>
std::vector<T> host(const T* d, std::size_t n)
{
    std::vector<T> h(n);
    nbody_amd::check(nbh_copy_to_host(h.data(), d, n * sizeof(T)));
    return h;
}

template <class T>
DeviceBuffer<T> device(const std::vector<T>& h)
{
    DeviceBuffer<T> b(h.size());
    nbody_amd::check_hip(hipMemcpy(b.data(), h.data(), h.size() * sizeof(T),
                                   hipMemcpyHostToDevice),
                         "upload");
    return b;
}

template <class E>
bool throws(const std::function<void()>& f)
{
    try {
        f();
    } catch (const E&) {
        return true;
    } catch (const std::exception& e) {
        std::printf("    other exception: %s\n", e.what());
        return false;
    }
    std::printf("    no exception\n");
    return false;
}

struct Set {
    std::vector<double> m, x, y, z, vx, vy, vz;
};

Set uniform(std::size_t n, unsigned seed)
{
    std::mt19937_64 rng(seed);
    std::uniform_real_distribution<double> u(-1.0, 1.0);
    Set s;
    for (std::size_t i = 0; i < n; ++i) {
        s.m.push_back(1.0 / static_cast<double>(n));
        s.x.push_back(u(rng));
        s.y.push_back(u(rng));
        s.z.push_back(u(rng));
        s.vx.push_back(0.1 * u(rng));
        s.vy.push_back(0.1 * u(rng));
        s.vz.push_back(0.1 * u(rng));
    }
    return s;
}

} // namespace

int main()
{
    int failed = 0;
    const auto run = [&](const char* name, const std::function<bool()>& f) {
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

    run("DeviceBuffer allocates once and moves", [] {
        DeviceBuffer<int> a;
        a.allocate(0);
        if (!a.empty())
            return false;
        a.allocate(16);
        if (!throws<std::logic_error>([&] { a.allocate(4); }))
            return false;
        DeviceBuffer<int> b(std::move(a));
        return a.data() == nullptr && b.size() == 16;
    });

    run("Bbox sentinel, known answer and input errors", [] {
        Bbox empty_box;
        if (empty_box.values().side != -1.0)
            return false;
        auto x = device<double>({-2.0, 4.0, 1.0});
        auto y = device<double>({10.0, 14.0, -2.0});
        auto z = device<double>({-1.0, 0.0, 5.0});
        const BboxValues v = Bbox(x.data(), y.data(), z.data(), 3).values();
        if (std::fabs(v.center_x - 1.0) > 1e-15 || std::fabs(v.center_y - 6.0) > 1e-15 ||
            std::fabs(v.center_z - 2.0) > 1e-15 || std::fabs(v.side - 17.6) > 1e-12)
            return false;
        auto bad = device<double>({0.0, NAN, 1.0});
        return throws<std::domain_error>([&] { Bbox(bad.data(), y.data(), z.data(), 3); }) &&
               throws<std::invalid_argument>([&] { Bbox(x.data(), y.data(), z.data(), 0); });
    });

    run("MortonKeys: sorted keys, stable ties, errors", [] {
        auto x = device<double>({0.5, -0.5, 0.5, 0.25});
        auto y = device<double>({0.5, -0.5, 0.5, 0.25});
        auto z = device<double>({0.5, -0.5, 0.5, 0.25});
        MortonKeys k(x.data(), y.data(), z.data(), 4);
        const auto keys = host(k.keys_device_data(), 4);
        const auto idx = host(k.indices_device_data(), 4);
        for (int i = 1; i < 4; ++i)
            if (keys[i] < keys[i - 1])
                return false;
        // particles 0 and 2 coincide: the stable sort keeps 0 before 2
        int p0 = -1, p2 = -1;
        for (int i = 0; i < 4; ++i) {
            if (idx[i] == 0)
                p0 = i;
            if (idx[i] == 2)
                p2 = i;
        }
        return p0 >= 0 && p2 == p0 + 1 && k.size() == 4 &&
               throws<std::invalid_argument>([&] { MortonKeys(x.data(), y.data(), z.data(), 0); });
    });

    run("staged pipeline equals the Simulation's forces bit for bit", [] {
        const std::size_t n = 3000;
        const Set s = uniform(n, 7);
        auto m = device(s.m);
        auto x = device(s.x);
        auto y = device(s.y);
        auto z = device(s.z);
        const int N = static_cast<int>(n);
        const Bbox box(x.data(), y.data(), z.data(), n);
        MortonKeys keys(x.data(), y.data(), z.data(), n, box);
        MortonLeafGroups groups;
        allocateMortonLeafGroups(groups, N);
        buildMortonLeafGroups(groups, keys.keys_device_data(), N);
        Tree tree;
        allocateTreeTopology(tree, groups.nGroups);
        buildTreeFromMortonGroups(tree, groups);
        RadixToOctreePlan plan;
        allocateRadixToOctreePlan(plan, 2 * groups.nGroups - 1);
        buildRadixToOctreePlan(plan, tree, groups);
        Octree oct;
        allocateOctreeTopology(oct, plan.nOctreeNodes);
        buildSparseOctreeTopology(oct, tree, groups, plan);
        allocateOctreeGeometry(oct);
        computeOctreeGeometry(oct, box);
        allocateOctreePhysicalData(oct);
        computeOctreeMassAndCenterOfMass(oct, keys.indices_device_data(), m.data(),
                                         x.data(), y.data(), z.data());
        DeviceBuffer<double> ax(n), ay(n), az(n);
        const BarnesHutParameters p{0.5, 0.01, 1.0};
        computeBarnesHutAcceleration(oct, keys.indices_device_data(), m.data(), x.data(),
                                     y.data(), z.data(), ax.data(), ay.data(), az.data(), n,
                                     p);
        Simulation sim(Particles(s.m, s.x, s.y, s.z, s.vx, s.vy, s.vz),
                       SimulationConfig{1e-3, 0.5, 0.01, 1.0});
        sim.initialize();
        double *sx = nullptr, *sy = nullptr, *sz = nullptr;
        nbody_amd::check(nbh_sim_accelerations(sim.handle(), &sx, &sy, &sz));
        return host(ax.data(), n) == host(sx, n) && host(ay.data(), n) == host(sy, n) &&
               host(az.data(), n) == host(sz, n);
    });

    run("stage functions throw the reference's exception classes", [] {
        Tree t;
        allocateTreeTopology(t, 8);
        MortonLeafGroups g;
        RadixToOctreePlan plan;
        Octree o;
        return throws<std::logic_error>([&] { allocateTreeTopology(t, 8); }) &&
               throws<std::invalid_argument>([&] { allocateTreeTopology(t, 0); }) &&
               throws<std::invalid_argument>([&] { buildTree(t, nullptr, 7); }) &&
               throws<std::invalid_argument>([&] { allocateMortonLeafGroups(g, 0); }) &&
               throws<std::logic_error>([&] {
                   std::uint32_t* k = nullptr;
                   DeviceBuffer<std::uint32_t> kb(4);
                   k = kb.data();
                   g.capacity = 4;
                   buildMortonLeafGroups(g, k, 4);
               }) &&
               throws<std::invalid_argument>([&] { allocateRadixToOctreePlan(plan, 0); }) &&
               throws<std::invalid_argument>([&] { allocateOctreeTopology(o, 0); }) &&
               throws<std::logic_error>([&] {
                   allocateOctreeTopology(o, 4);
                   allocateOctreeTopology(o, 4);
               });
    });

    run("leapfrog known answers", [] {
        auto x = device<double>({1.0});
        auto y = device<double>({2.0});
        auto z = device<double>({3.0});
        auto vx = device<double>({1.0});
        auto vy = device<double>({0.0});
        auto vz = device<double>({-1.0});
        auto ax = device<double>({2.0});
        auto ay = device<double>({4.0});
        auto az = device<double>({0.0});
        leapfrogKickDrift(x.data(), y.data(), z.data(), vx.data(), vy.data(), vz.data(),
                          ax.data(), ay.data(), az.data(), 1, 0.5);
        // v = 1 + 0.25 * 2 = 1.5, x = 1 + 0.5 * 1.5 = 1.75; vy = 1, y = 2.5
        const bool kd = host(x.data(), 1)[0] == 1.75 && host(vx.data(), 1)[0] == 1.5 &&
                        host(y.data(), 1)[0] == 2.5 && host(vz.data(), 1)[0] == -1.0;
        leapfrogKick(vx.data(), vy.data(), vz.data(), ax.data(), ay.data(), az.data(), 1, 0.5);
        return kd && host(vx.data(), 1)[0] == 2.0 && host(vy.data(), 1)[0] == 2.0 &&
               throws<std::invalid_argument>([&] {
                   leapfrogKick(vx.data(), vy.data(), vz.data(), ax.data(), ay.data(),
                                az.data(), 1, 0.0);
               });
    });

    run("Simulation lifecycle and configuration errors", [] {
        const Set s = uniform(500, 3);
        const auto make = [&](SimulationConfig c) {
            return Simulation(Particles(s.m, s.x, s.y, s.z, s.vx, s.vy, s.vz), c);
        };
        if (!throws<std::invalid_argument>([&] { make({0.0, 0.5, 0.01}); }) ||
            !throws<std::invalid_argument>([&] { make({1e-3, -1.0, 0.01}); }) ||
            !throws<std::invalid_argument>([&] { make({1e-3, 0.5, -0.01}); }))
            return false;
        Simulation sim(Particles(s.m, s.x, s.y, s.z, s.vx, s.vy, s.vz),
                       SimulationConfig{1e-3, 0.5, 0.01, 1.0});
        if (!throws<std::logic_error>([&] { sim.step(); }) ||
            !throws<std::logic_error>([&] { sim.run(1); }))
            return false;
        sim.initialize();
        if (!throws<std::logic_error>([&] { sim.initialize(); }))
            return false;
        sim.run(5);
        const DeviceParticlesView v = sim.particles();
        const auto px = host(v.x, v.count);
        return sim.stepNumber() == 5 && std::fabs(sim.time() - 5e-3) < 1e-15 &&
               v.count == 500 && px != s.x;
    });

    run("Simulation in SI units: circular two-body orbit", [] {
        // Sun-Earth-like pair: one period of a circular orbit at 1 AU
        const double M = 1.989e30, m = 5.972e24, r = 1.496e11;
        const double Gs = 6.67430e-11;
        const double v = std::sqrt(Gs * (M + m) / r);
        const double xs = -r * m / (M + m), xe = r * M / (M + m);
        const double vs = -v * m / (M + m), ve = v * M / (M + m);
        const double pi = 3.14159265358979323846;
        const double period = 2.0 * pi * std::sqrt(r * r * r / (Gs * (M + m)));
        Simulation sim(Particles({M, m}, {xs, xe}, {0.0, 0.0}, {0.0, 0.0}, {0.0, 0.0},
                                 {vs, ve}, {0.0, 0.0}),
                       SimulationConfig{period / 2000.0, 0.5, 0.0});
        sim.initialize();
        sim.run(2000);
        const DeviceParticlesView pv = sim.particles();
        const auto px = host(pv.x, 2), py = host(pv.y, 2);
        const double sep = std::hypot(px[1] - px[0], py[1] - py[0]);
        const double com_x = (M * px[0] + m * px[1]) / (M + m);
        return std::fabs(sep - r) < 2e-3 * r && std::fabs(px[1] - xe) < 2e-2 * r &&
               std::fabs(com_x) < 1e-6 * r;
    });

    run("DeviceFrameWriter on a Simulation: sampled, quantized, typed frames", [] {
        const std::size_t n = 3000, S = 700;
        const Set s = uniform(n, 21);
        std::vector<std::uint8_t> types(n);
        for (std::size_t i = 0; i < n; ++i)
            types[i] = static_cast<std::uint8_t>(i % 5);
        SimulationConfig cfg{1e-3, 0.5, 1e-2};
        cfg.gravitationalConstant = 1.0;
        Simulation sim(Particles(s.m, s.x, s.y, s.z, s.vx, s.vy, s.vz, types), cfg);
        sim.initialize();
        sim.run(2);
        const DeviceParticlesView v = sim.particles();
        const ConstDeviceParticlesView cv{v.count, v.mass, v.x, v.y, v.z,
                                          v.vx, v.vy, v.vz, v.type};
        nbody::streaming::DeviceFrameWriter writer(cv, S);
        const auto dir = std::filesystem::temp_directory_path() / "nbody-amd-shim-frames";
        std::filesystem::remove_all(dir);
        nbody::streaming::FrameSpool spool(dir, "shim");
        const auto path = writer.write(spool, 4, sim.stepNumber(), sim.time());
        const auto h = nbody::streaming::readFrameHeader(path);
        const auto pos = nbody::streaming::readFramePositions(path);
        const auto t = nbody::streaming::readFrameParticleTypes(path);
        const auto px = host(v.x, n), py = host(v.y, n), pz = host(v.z, n);
        bool ok = h.version == 3 && h.sequence == 4 && h.simulationStep == 2 &&
                  h.particleCount == S && h.sourceParticleCount == n &&
                  writer.sampleParticleCount() == S && writer.sourceParticleCount() == n;
        double lo[3] = {1e300, 1e300, 1e300}, hi[3] = {-1e300, -1e300, -1e300};
        for (std::size_t j = 0; j < S; ++j) {
            const std::size_t i = j * n / S;
            const double p[3] = {px[i], py[i], pz[i]};
            for (int c = 0; c < 3; ++c)
                lo[c] = std::min(lo[c], p[c]), hi[c] = std::max(hi[c], p[c]);
        }
        for (int c = 0; c < 3; ++c)
            ok = ok && h.bounds.minimum[c] == lo[c] && h.bounds.maximum[c] == hi[c];
        for (std::size_t j = 0; ok && j < S; ++j) {
            const std::size_t i = j * n / S;
            const double p[3] = {px[i], py[i], pz[i]};
            ok = t[j] == i % 5;
            for (int c = 0; c < 3 && ok; ++c)
                ok = std::fabs(pos[3 * j + c] - p[c]) <= (hi[c] - lo[c]) / 65535.0 + 1e-6;
        }
        const bool rejects = throws<std::invalid_argument>([&] {
            nbody::streaming::DeviceFrameWriter w(ConstDeviceParticlesView{0}, 10);
        });
        std::filesystem::remove_all(dir);
        return ok && rejects;
    });

    std::printf("%d failed\n", failed);
    return failed == 0 ? 0 : 1;
}
