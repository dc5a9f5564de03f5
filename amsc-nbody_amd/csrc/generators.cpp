// generators.cpp -- seeded synthetic particle sets for the benchmark and
// parity workloads.
//
// Every particle draws from its own counter-based stream (splitmix64 of
// seed, particle index and draw number), so a set is identical for any
// thread count.  Equal masses are 1/N and G = 1 throughout.
#include "nbody_amd.h"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <numeric>
#include <thread>
#include <vector>

namespace {

inline uint64_t mix64(uint64_t z)
{
    z += 0x9E3779B97F4A7C15ull;
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

// Stream of uniform doubles in (0, 1) for one particle.
struct Stream {
    uint64_t key;
    uint64_t counter = 0;
    Stream(uint64_t seed, uint64_t id) : key(mix64(seed ^ mix64(id + 0x632BE59BD9B4E019ull))) {}
    double uniform()
    {
        const uint64_t r = mix64(key + 0xD1B54A32D192ED03ull * (++counter));
        return (static_cast<double>(r >> 11) + 0.5) * (1.0 / 9007199254740992.0);
    }
    double normal()
    {
        const double u1 = uniform();
        const double u2 = uniform();
        return std::sqrt(-2.0 * std::log(u1)) * std::cos(6.283185307179586 * u2);
    }
    void direction(double& x, double& y, double& z)
    {
        const double c = 2.0 * uniform() - 1.0;
        const double phi = 6.283185307179586 * uniform();
        const double s = std::sqrt(std::max(0.0, 1.0 - c * c));
        x = s * std::cos(phi);
        y = s * std::sin(phi);
        z = c;
    }
};

template <typename F>
void parallel_for(size_t n, F&& f)
{
    unsigned t = std::thread::hardware_concurrency();
    if (t == 0)
        t = 1;
    t = std::min<unsigned>(t, 32);
    if (n < 65536)
        t = 1;
    std::vector<std::thread> pool;
    const size_t chunk = (n + t - 1) / t;
    for (unsigned k = 0; k < t; ++k) {
        const size_t b = k * chunk;
        const size_t e = std::min(n, b + chunk);
        if (b >= e)
            break;
        pool.emplace_back([&f, b, e] {
            for (size_t i = b; i < e; ++i)
                f(i);
        });
    }
    for (auto& th : pool)
        th.join();
}

struct Out {
    double *m, *x, *y, *z, *vx, *vy, *vz;
};

// Plummer sphere, a = M = G = 1: r = 1/sqrt(X^(-2/3) - 1) rejected above 10,
// isotropic; speed q v_esc with q from von Neumann rejection on
// g(q) = q^2 (1 - q^2)^(7/2) <= 0.1 and v_esc = sqrt(2) (1 + r^2)^(-1/4).
void plummer(size_t n, uint64_t seed, Out o)
{
    parallel_for(n, [&](size_t i) {
        Stream s(seed, i);
        double r;
        do {
            const double X = s.uniform();
            r = 1.0 / std::sqrt(std::pow(X, -2.0 / 3.0) - 1.0);
        } while (!(r <= 10.0));
        double ux, uy, uz;
        s.direction(ux, uy, uz);
        o.x[i] = r * ux;
        o.y[i] = r * uy;
        o.z[i] = r * uz;
        double q;
        while (true) {
            q = s.uniform();
            const double y = 0.1 * s.uniform();
            if (y < q * q * std::pow(1.0 - q * q, 3.5))
                break;
        }
        const double v = q * std::sqrt(2.0) * std::pow(1.0 + r * r, -0.25);
        s.direction(ux, uy, uz);
        o.vx[i] = v * ux;
        o.vy[i] = v * uy;
        o.vz[i] = v * uz;
        o.m[i] = 1.0 / static_cast<double>(n);
    });
}

// Uniform cube [-1,1]^3, lognormal(0, 0.5) masses normalised to 1,
// Gaussian velocities with sigma 0.1 per component.
void uniform_cube(size_t n, uint64_t seed, Out o)
{
    parallel_for(n, [&](size_t i) {
        Stream s(seed, i);
        o.x[i] = 2.0 * s.uniform() - 1.0;
        o.y[i] = 2.0 * s.uniform() - 1.0;
        o.z[i] = 2.0 * s.uniform() - 1.0;
        o.m[i] = std::exp(0.5 * s.normal());
        o.vx[i] = 0.1 * s.normal();
        o.vy[i] = 0.1 * s.normal();
        o.vz[i] = 0.1 * s.normal();
    });
    long double total = 0.0L;
    for (size_t i = 0; i < n; ++i)
        total += o.m[i];
    const double inv = static_cast<double>(1.0L / total);
    for (size_t i = 0; i < n; ++i)
        o.m[i] *= inv;
}

// Exponential disk (80% of particles, scale length 1, sech^2 vertical
// profile of scale height 0.1) and Hernquist bulge (20%, a = 0.2).  Disk
// velocities: circular speed from the enclosed mass (spherical estimate)
// plus 10% Gaussian dispersion; bulge velocities: isotropic Gaussian with
// sigma^2 = M(<r) / (3 r).
void disk_bulge(size_t n, uint64_t seed, Out o)
{
    const size_t nd = (n * 4) / 5;
    std::vector<double> rad(n);
    parallel_for(n, [&](size_t i) {
        Stream s(seed, i);
        if (i < nd) {
            double R;
            do {
                R = -std::log(s.uniform() * s.uniform());
            } while (R > 20.0);
            const double phi = 6.283185307179586 * s.uniform();
            const double zz = 0.1 * std::atanh(2.0 * s.uniform() - 1.0);
            o.x[i] = R * std::cos(phi);
            o.y[i] = R * std::sin(phi);
            o.z[i] = zz;
        } else {
            double r;
            do {
                const double su = std::sqrt(s.uniform());
                r = 0.2 * su / (1.0 - su);
            } while (r > 10.0);
            double ux, uy, uz;
            s.direction(ux, uy, uz);
            o.x[i] = r * ux;
            o.y[i] = r * uy;
            o.z[i] = r * uz;
        }
        o.m[i] = 1.0 / static_cast<double>(n);
        rad[i] = std::sqrt(o.x[i] * o.x[i] + o.y[i] * o.y[i] + o.z[i] * o.z[i]);
    });
    std::vector<double> sorted(rad);
    std::sort(sorted.begin(), sorted.end());
    auto enclosed = [&](double r) {
        const size_t k = std::upper_bound(sorted.begin(), sorted.end(), r) - sorted.begin();
        return static_cast<double>(k) / static_cast<double>(n);
    };
    parallel_for(n, [&](size_t i) {
        Stream s(seed ^ 0xA5A5A5A5A5A5A5A5ull, i);
        const double r = std::max(rad[i], 1e-6);
        const double menc = enclosed(r);
        if (i < nd) {
            const double R = std::max(std::sqrt(o.x[i] * o.x[i] + o.y[i] * o.y[i]), 1e-6);
            const double vc = std::sqrt(menc / r);
            const double sig = 0.1 * vc;
            o.vx[i] = -vc * o.y[i] / R + sig * s.normal();
            o.vy[i] = vc * o.x[i] / R + sig * s.normal();
            o.vz[i] = sig * s.normal();
        } else {
            const double sig = std::sqrt(menc / (3.0 * r));
            o.vx[i] = sig * s.normal();
            o.vy[i] = sig * s.normal();
            o.vz[i] = sig * s.normal();
        }
    });
}

// 256 Gaussian clumps: centres uniform in [-10,10]^3, sigma log-uniform in
// [0.05, 0.5], Poisson populations around N/256 renormalised to N, equal
// masses, zero velocities.
void clumps(size_t n, uint64_t seed, Out o)
{
    constexpr int K = 256;
    Stream g(seed, 0xC10C10ull);
    double cen[K][3], sig[K];
    std::vector<double> pop(K);
    const double lambda = static_cast<double>(n) / K;
    for (int k = 0; k < K; ++k) {
        for (int a = 0; a < 3; ++a)
            cen[k][a] = -10.0 + 20.0 * g.uniform();
        sig[k] = 0.05 * std::pow(10.0, g.uniform());
        double c;
        if (lambda < 30.0) {
            const double L = std::exp(-lambda);
            double p = 1.0;
            int kk = 0;
            do {
                ++kk;
                p *= g.uniform();
            } while (p > L);
            c = kk - 1;
        } else {
            c = std::floor(lambda + std::sqrt(lambda) * g.normal() + 0.5);
        }
        pop[k] = std::max(c, 0.0);
    }
    const double tot = std::accumulate(pop.begin(), pop.end(), 0.0);
    std::vector<size_t> cnt(K);
    size_t sum = 0;
    for (int k = 0; k < K; ++k) {
        cnt[k] = tot > 0 ? static_cast<size_t>(std::floor(pop[k] * n / tot)) : n / K;
        sum += cnt[k];
    }
    for (size_t k = 0; sum < n; k = (k + 1) % K, ++sum)
        ++cnt[k];
    std::vector<int> owner(n);
    size_t p = 0;
    for (int k = 0; k < K; ++k)
        for (size_t j = 0; j < cnt[k] && p < n; ++j)
            owner[p++] = k;
    parallel_for(n, [&](size_t i) {
        Stream s(seed, i);
        const int k = owner[i];
        o.x[i] = cen[k][0] + sig[k] * s.normal();
        o.y[i] = cen[k][1] + sig[k] * s.normal();
        o.z[i] = cen[k][2] + sig[k] * s.normal();
        o.vx[i] = o.vy[i] = o.vz[i] = 0.0;
        o.m[i] = 1.0 / static_cast<double>(n);
    });
}

// Gaussian N(0,1)^3 blob with velocities N(0, 0.01^2), equal masses: the
// shape of the reference benchmark input (data/generate_galaxy.py:96-107).
void blob(size_t n, uint64_t seed, Out o)
{
    parallel_for(n, [&](size_t i) {
        Stream s(seed, i);
        o.x[i] = s.normal();
        o.y[i] = s.normal();
        o.z[i] = s.normal();
        o.vx[i] = 0.01 * s.normal();
        o.vy[i] = 0.01 * s.normal();
        o.vz[i] = 0.01 * s.normal();
        o.m[i] = 1.0 / static_cast<double>(n);
    });
}

} // namespace

extern "C" int nbh_generate(int kind, size_t count, uint64_t seed,
                            double* h_mass, double* h_x, double* h_y,
                            double* h_z, double* h_vx, double* h_vy,
                            double* h_vz)
{
    if (count == 0 || !h_mass || !h_x || !h_y || !h_z || !h_vx || !h_vy || !h_vz)
        return NBH_INVALID_ARGUMENT;
    const Out o{h_mass, h_x, h_y, h_z, h_vx, h_vy, h_vz};
    switch (kind) {
    case 0: plummer(count, seed, o); break;
    case 1: uniform_cube(count, seed, o); break;
    case 2: disk_bulge(count, seed, o); break;
    case 3: clumps(count, seed, o); break;
    case 4: blob(count, seed, o); break;
    default: return NBH_INVALID_ARGUMENT;
    }
    return NBH_OK;
}
