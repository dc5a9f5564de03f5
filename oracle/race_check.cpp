// Sanitizer driver of the oracle (test infrastructure): runs the oracle's
// parallel paths -- Morton keys, the Simulation loop (bbox, tree, mass/COM,
// the Barnes-Hut forces over an OpenMP loop, leapfrog), the direct sum and
// frame capture -- at 1 and 4 threads on a seeded particle set and requires
// identical results.  Built under AddressSanitizer + UndefinedBehavior-
// Sanitizer and under ThreadSanitizer (with LLVM's OpenMP runtime and its
// Archer tool, so the runtime's synchronisation is visible to the race
// detector) by the sanitize targets of oracle/Makefile.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <vector>

extern "C" {
const char* orc_last_error_message(void);
void orc_set_num_threads(int n);
int orc_bbox(const double* x, const double* y, const double* z, std::size_t n, double* out);
int orc_morton_keys(const double* x, const double* y, const double* z, std::size_t n,
                    const double* box, int bits, void* keys, std::uint32_t* idx);
void* orc_sim_create(int bits, std::size_t n, const double* m, const double* x,
                     const double* y, const double* z, const double* vx, const double* vy,
                     const double* vz, double dt, double theta, double softening, double G);
void orc_sim_destroy(void* h);
int orc_sim_initialize(void* h);
int orc_sim_step(void* h, int steps);
int orc_sim_partial_forces(void* h, int t_begin, int t_end);
int orc_sim_state(void* h, double* x, double* y, double* z, double* vx, double* vy,
                  double* vz, double* ax, double* ay, double* az);
int orc_direct(const double* m, const double* x, const double* y, const double* z,
               std::size_t n, const std::uint32_t* targets, std::size_t n_targets,
               double softening, double G, double* ax, double* ay, double* az);
int orc_frame_capture(const double* x, const double* y, const double* z,
                      const std::uint8_t* type, std::size_t n, std::size_t s, double* bounds,
                      std::uint16_t* q, std::uint8_t* types);
}

namespace {

struct Set {
    std::vector<double> m, x, y, z, vx, vy, vz;
};

Set clustered(std::size_t n, unsigned seed)
{
    std::mt19937_64 rng(seed);
    std::normal_distribution<double> g(0.0, 1.0);
    std::uniform_int_distribution<int> c(0, 7);
    Set s;
    for (std::size_t i = 0; i < n; ++i) {
        const int k = c(rng);
        const double cx = (k & 1) ? 3.0 : -3.0, cy = (k & 2) ? 3.0 : -3.0,
                     cz = (k & 4) ? 3.0 : -3.0;
        s.m.push_back(1.0 / static_cast<double>(n));
        s.x.push_back(cx + 0.5 * g(rng));
        s.y.push_back(cy + 0.5 * g(rng));
        s.z.push_back(cz + 0.5 * g(rng));
        s.vx.push_back(0.05 * g(rng));
        s.vy.push_back(0.05 * g(rng));
        s.vz.push_back(0.05 * g(rng));
    }
    return s;
}

struct Result {
    std::vector<std::uint64_t> keys;
    std::vector<std::uint32_t> idx;
    std::vector<double> state, direct, bounds;
    std::vector<std::uint16_t> frame;
};

bool ok(int status, const char* what)
{
    if (status != 0)
        std::printf("FAIL %s: %s\n", what, orc_last_error_message());
    return status == 0;
}

bool run(const Set& s, int threads, Result& r)
{
    orc_set_num_threads(threads);
    const std::size_t n = s.m.size();
    double box[4];
    if (!ok(orc_bbox(s.x.data(), s.y.data(), s.z.data(), n, box), "bbox"))
        return false;
    r.keys.assign(n, 0);
    r.idx.assign(n, 0);
    if (!ok(orc_morton_keys(s.x.data(), s.y.data(), s.z.data(), n, box, 21, r.keys.data(),
                            r.idx.data()),
            "morton keys"))
        return false;
    void* h = orc_sim_create(21, n, s.m.data(), s.x.data(), s.y.data(), s.z.data(),
                             s.vx.data(), s.vy.data(), s.vz.data(), 1e-3, 0.5, 0.01, 1.0);
    bool good = ok(orc_sim_initialize(h), "initialize") && ok(orc_sim_step(h, 3), "steps") &&
                ok(orc_sim_partial_forces(h, 0, static_cast<int>(n / 3)), "partial forces");
    r.state.assign(9 * n, 0.0);
    good = good && ok(orc_sim_state(h, &r.state[0], &r.state[n], &r.state[2 * n],
                                    &r.state[3 * n], &r.state[4 * n], &r.state[5 * n],
                                    &r.state[6 * n], &r.state[7 * n], &r.state[8 * n]),
                      "state");
    orc_sim_destroy(h);
    const std::size_t nt = 500;
    std::vector<std::uint32_t> targets(nt);
    for (std::size_t k = 0; k < nt; ++k)
        targets[k] = static_cast<std::uint32_t>((k * 7919) % n);
    r.direct.assign(3 * nt, 0.0);
    good = good && ok(orc_direct(s.m.data(), s.x.data(), s.y.data(), s.z.data(), n,
                                 targets.data(), nt, 0.01, 1.0, &r.direct[0], &r.direct[nt],
                                 &r.direct[2 * nt]),
                      "direct");
    const std::size_t fs = n / 2;
    r.bounds.assign(6, 0.0);
    r.frame.assign(3 * fs, 0);
    good = good && ok(orc_frame_capture(s.x.data(), s.y.data(), s.z.data(), nullptr, n, fs,
                                        r.bounds.data(), r.frame.data(), nullptr),
                      "frame capture");
    return good;
}

} // namespace

int main()
{
    const Set s = clustered(12000, 2026);
    Result one, four;
    if (!run(s, 1, one) || !run(s, 4, four))
        return 1;
    const bool same = one.keys == four.keys && one.idx == four.idx &&
                      std::memcmp(one.state.data(), four.state.data(),
                                  one.state.size() * sizeof(double)) == 0 &&
                      one.direct == four.direct && one.bounds == four.bounds &&
                      one.frame == four.frame;
    if (!same) {
        std::printf("FAIL results differ between 1 and 4 threads\n");
        return 1;
    }
    std::printf("PASS oracle parallel paths, 1 and 4 threads, identical results\n");
    return 0;
}
