// The host-side C ABI without a device: the workload generators, the
// particle-file reader and the .nbsnap frame spool and readers, plus the
// Hilbert-curve table derivation.  Built from the library's own host sources
// (csrc/generators.cpp, particle_io.cpp, frame_io.cpp, hilbert.cpp,
// error_state.cpp) by the sanitize target of the Makefile, under
// AddressSanitizer + UndefinedBehaviorSanitizer.  Prints one line per case;
// exit status 0 when every case passes.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <string>
#include <vector>

#include <unistd.h>

#include "../../csrc/hilbert.hpp"
#include "nbody_amd.h"

namespace fs = std::filesystem;

namespace {

struct Set {
    std::vector<double> m, x, y, z, vx, vy, vz;
    explicit Set(std::size_t n) : m(n), x(n), y(n), z(n), vx(n), vy(n), vz(n) {}
};

void write_file(const fs::path& p, const Set& s, bool wide, const std::vector<uint8_t>* types,
                std::size_t drop_bytes = 0)
{
    std::vector<char> buf;
    const auto put = [&](const void* d, std::size_t b) {
        buf.insert(buf.end(), static_cast<const char*>(d), static_cast<const char*>(d) + b);
    };
    const std::uint64_t n64 = s.m.size();
    const std::uint32_t n32 = static_cast<std::uint32_t>(s.m.size());
    if (wide)
        put(&n64, 8);
    else
        put(&n32, 4);
    for (const auto* v : {&s.m, &s.x, &s.y, &s.z, &s.vx, &s.vy, &s.vz})
        put(v->data(), v->size() * sizeof(double));
    if (types)
        put(types->data(), types->size());
    buf.resize(buf.size() - drop_bytes);
    std::ofstream(p, std::ios::binary).write(buf.data(), static_cast<std::streamsize>(buf.size()));
}

bool read_back(const fs::path& p, const Set& s, bool with_types, const std::vector<uint8_t>* t)
{
    std::size_t count = 0;
    int has_types = -1;
    if (nbh_particle_file_info(p.c_str(), &count, &has_types) != NBH_OK)
        return false;
    if (count != s.m.size() || has_types != (with_types ? 1 : 0))
        return false;
    Set r(count);
    std::vector<uint8_t> types(count, 99);
    std::size_t got = 0;
    if (nbh_particle_file_read(p.c_str(), count, r.m.data(), r.x.data(), r.y.data(), r.z.data(),
                               r.vx.data(), r.vy.data(), r.vz.data(), types.data(), &got) != NBH_OK)
        return false;
    return got == count && r.m == s.m && r.x == s.x && r.z == s.z && r.vy == s.vy &&
           (!with_types || types == *t) && (with_types || types[0] == 0);
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
    const fs::path dir = fs::temp_directory_path() / ("nbody-amd-host-abi-" + std::to_string(::getpid()));
    fs::remove_all(dir);
    fs::create_directories(dir);

    run("generators: every kind, several sizes, finite and seeded", [] {
        for (int kind = 0; kind < 5; ++kind) {
            for (std::size_t n : {std::size_t{1}, std::size_t{2}, std::size_t{1000}, std::size_t{65537}}) {
                Set a(n), b(n);
                if (nbh_generate(kind, n, 7, a.m.data(), a.x.data(), a.y.data(), a.z.data(),
                                 a.vx.data(), a.vy.data(), a.vz.data()) != NBH_OK ||
                    nbh_generate(kind, n, 7, b.m.data(), b.x.data(), b.y.data(), b.z.data(),
                                 b.vx.data(), b.vy.data(), b.vz.data()) != NBH_OK)
                    return false;
                double total = 0.0;
                for (std::size_t i = 0; i < n; ++i) {
                    if (!std::isfinite(a.x[i]) || !std::isfinite(a.vz[i]) || !(a.m[i] > 0.0))
                        return false;
                    total += a.m[i];
                }
                if (a.x != b.x || a.vy != b.vy || !(total > 0.0))
                    return false;
            }
        }
        Set s(4);
        return nbh_generate(9, 4, 1, s.m.data(), s.x.data(), s.y.data(), s.z.data(), s.vx.data(),
                            s.vy.data(), s.vz.data()) == NBH_INVALID_ARGUMENT &&
               nbh_generate(0, 0, 1, s.m.data(), s.x.data(), s.y.data(), s.z.data(),
                            s.vx.data(), s.vy.data(), s.vz.data()) == NBH_INVALID_ARGUMENT;
    });

    run("particle files: both header widths, type block, rejections", [&] {
        Set s(257);
        nbh_generate(1, 257, 3, s.m.data(), s.x.data(), s.y.data(), s.z.data(), s.vx.data(),
                     s.vy.data(), s.vz.data());
        std::vector<uint8_t> t(257);
        for (int i = 0; i < 257; ++i)
            t[i] = static_cast<uint8_t>(i % 5);
        for (bool wide : {false, true}) {
            write_file(dir / "a.bin", s, wide, nullptr);
            write_file(dir / "b.bin", s, wide, &t);
            if (!read_back(dir / "a.bin", s, false, nullptr) || !read_back(dir / "b.bin", s, true, &t))
                return false;
        }
        Set bad = s;
        bad.x[5] = NAN;
        write_file(dir / "nan.bin", bad, true, nullptr);
        Set neg = s;
        neg.m[3] = -1.0;
        write_file(dir / "neg.bin", neg, true, nullptr);
        write_file(dir / "short.bin", s, true, nullptr, 5);
        std::size_t n = 0, got = 0;
        int ht = 0;
        Set r(257);
        const auto read = [&](const fs::path& p) {
            return nbh_particle_file_read(p.c_str(), 257, r.m.data(), r.x.data(), r.y.data(),
                                          r.z.data(), r.vx.data(), r.vy.data(), r.vz.data(),
                                          nullptr, &got);
        };
        return read(dir / "nan.bin") == NBH_RUNTIME_ERROR &&
               std::strlen(nbh_last_error_message()) > 0 &&
               read(dir / "neg.bin") == NBH_RUNTIME_ERROR &&
               nbh_particle_file_info((dir / "short.bin").c_str(), &n, &ht) == NBH_RUNTIME_ERROR &&
               nbh_particle_file_info((dir / "missing.bin").c_str(), &n, &ht) != NBH_OK;
    });

    run("frames: quantized and float layouts, types, completion marker", [&] {
        const std::size_t n = 1000;
        std::vector<uint16_t> q(3 * n);
        std::vector<float> f(3 * n);
        std::vector<uint8_t> t(n);
        for (std::size_t i = 0; i < 3 * n; ++i) {
            q[i] = static_cast<uint16_t>((i * 37) & 0xFFFF);
            f[i] = static_cast<float>(i) * 0.5f;
        }
        for (std::size_t i = 0; i < n; ++i)
            t[i] = static_cast<uint8_t>(i % 5);
        const double bounds[6] = {-1.0, -2.0, -3.0, 1.0, 2.0, 3.0};
        char path[4096];
        if (nbh_frame_write_quantized(dir.c_str(), "s1", 3, 10, 0.5, n, 2 * n, bounds, q.data(),
                                      t.data(), path, sizeof(path)) != NBH_OK)
            return false;
        nbh_frame_header h{};
        if (nbh_frame_read_header(path, &h) != NBH_OK || h.version != 3 || h.particle_count != n ||
            h.source_particle_count != 2 * n || h.sequence != 3)
            return false;
        std::vector<float> pos(3 * n);
        std::vector<uint8_t> tt(n);
        std::size_t got = 0;
        if (nbh_frame_read(path, n, pos.data(), tt.data(), &got) != NBH_OK || got != n || tt != t)
            return false;
        // a published frame is never overwritten
        if (nbh_frame_write_quantized(dir.c_str(), "s1", 3, 10, 0.5, n, 0, bounds, q.data(),
                                      t.data(), nullptr, 0) == NBH_OK)
            return false;
        if (nbh_frame_write_float(dir.c_str(), "s1", 4, 11, 0.75, 20, n, 0, f.data(), nullptr,
                                  path, sizeof(path)) != NBH_OK ||
            nbh_frame_read_header(path, &h) != NBH_OK || h.version != 1)
            return false;
        if (nbh_frame_read(path, n, pos.data(), tt.data(), &got) != NBH_OK ||
            std::memcmp(pos.data(), f.data(), 3 * n * sizeof(float)) != 0)
            return false;
        return nbh_frame_mark_complete(dir.c_str(), "s1", path, sizeof(path)) == NBH_OK &&
               nbh_frame_mark_complete(dir.c_str(), "s1", nullptr, 0) == NBH_OK &&
               nbh_frame_write_float(dir.c_str(), "bad id!", 5, 0, 0.0, 0, n, 0, f.data(),
                                     nullptr, nullptr, 0) == NBH_INVALID_ARGUMENT;
    });

    run("Hilbert tables: derived and checked against the index transform", [] {
        const nbh::HilbertTables& t = nbh::hilbert_tables();
        for (int s = 0; s < nbh::kHilbertStates; ++s) {
            int seen = 0;
            for (int oc = 0; oc < 8; ++oc) {
                if (t.rank[s][oc] > 7 || t.next[s][oc] >= nbh::kHilbertStates)
                    return false;
                seen |= 1 << t.rank[s][oc];
            }
            if (seen != 0xFF)
                return false;
        }
        return true;
    });

    fs::remove_all(dir);
    std::printf("%d failed\n", failed);
    return failed == 0 ? 0 : 1;
}
