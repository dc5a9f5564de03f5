// Particle file reader cases: the format and the failures the reference's
// loader defines (tests/unit/cpu/particle_file_test.cpp), on the host-only
// reader.  Prints one line per case; exit status 0 when every case passes.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include "particle_file.hpp"

namespace {

struct Data {
    std::vector<double> mass{1.5e30, 2.5e30};
    std::vector<double> x{1.0, 2.0}, y{3.0, 4.0}, z{5.0, 6.0};
    std::vector<double> vx{0.1, 0.2}, vy{0.3, 0.4}, vz{0.5, 0.6};
    std::vector<std::uint8_t> type{1, 4};
};

std::string bytes_of(const Data& d, bool wide, bool types, std::uint64_t count)
{
    std::ostringstream o(std::ios::binary);
    if (wide) {
        o.write(reinterpret_cast<const char*>(&count), sizeof(count));
    } else {
        const auto c = static_cast<std::uint32_t>(count);
        o.write(reinterpret_cast<const char*>(&c), sizeof(c));
    }
    for (const auto* b : {&d.mass, &d.x, &d.y, &d.z, &d.vx, &d.vy, &d.vz})
        o.write(reinterpret_cast<const char*>(b->data()),
                static_cast<std::streamsize>(b->size() * sizeof(double)));
    if (types)
        o.write(reinterpret_cast<const char*>(d.type.data()),
                static_cast<std::streamsize>(d.type.size()));
    return o.str();
}

std::string bytes_of(const Data& d, bool wide, bool types = false)
{
    return bytes_of(d, wide, types, d.mass.size());
}

const std::filesystem::path kFile =
    std::filesystem::temp_directory_path() / "nbody_amd_particle_file_test.bin";

void write_file(const std::string& b)
{
    std::ofstream o(kFile, std::ios::binary | std::ios::trunc);
    o.write(b.data(), static_cast<std::streamsize>(b.size()));
}

nbody_amd::ParticleFileData load()
{
    std::ifstream in(kFile, std::ios::binary);
    return nbody_amd::read_particle_file(in);
}

bool throws_containing(const std::function<void()>& f, const std::string& text)
{
    try {
        f();
    } catch (const std::runtime_error& e) {
        if (std::string(e.what()).find(text) != std::string::npos)
            return true;
        std::printf("    unexpected message: %s\n", e.what());
        return false;
    }
    std::printf("    no exception\n");
    return false;
}

bool same(const Data& d, const nbody_amd::ParticleFileData& r, bool types)
{
    if (r.mass != d.mass || r.x != d.x || r.y != d.y || r.z != d.z || r.vx != d.vx ||
        r.vy != d.vy || r.vz != d.vz)
        return false;
    if (types)
        return r.type == d.type;
    for (auto t : r.type)
        if (t != 0)
            return false;
    return r.type.size() == d.mass.size();
}

} // namespace

int main()
{
    const Data d;
    int failed = 0;
    const auto run = [&](const char* name, const std::function<bool()>& f) {
        bool ok = false;
        try {
            ok = f();
        } catch (const std::exception& e) {
            std::printf("    exception: %s\n", e.what());
        }
        std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
        failed += ok ? 0 : 1;
    };
    run("loads every block, 32-bit header", [&] {
        write_file(bytes_of(d, false));
        return same(d, load(), false);
    });
    run("loads every block, 64-bit header", [&] {
        write_file(bytes_of(d, true));
        return same(d, load(), false);
    });
    run("loads particle types, 32-bit header", [&] {
        write_file(bytes_of(d, false, true));
        return same(d, load(), true);
    });
    run("loads particle types, 64-bit header", [&] {
        write_file(bytes_of(d, true, true));
        return same(d, load(), true);
    });
    run("loads regardless of initial stream position", [&] {
        write_file(bytes_of(d, true));
        std::ifstream in(kFile, std::ios::binary);
        in.seekg(5);
        return same(d, nbody_amd::read_particle_file(in), false);
    });
    run("rejects a stream in an error state", [&] {
        write_file(bytes_of(d, false));
        return throws_containing([&] {
            std::ifstream in(kFile, std::ios::binary);
            in.setstate(std::ios::badbit);
            nbody_amd::read_particle_file(in);
        }, "Invalid or unreadable file stream");
    });
    run("rejects an empty file", [&] {
        write_file("");
        return throws_containing([] { load(); }, "Unsupported particle file layout");
    });
    run("rejects a zero particle count", [&] {
        const std::uint32_t zero = 0;
        write_file(std::string(reinterpret_cast<const char*>(&zero), sizeof(zero)));
        return throws_containing([] { load(); }, "Unsupported particle file layout");
    });
    run("rejects truncated particle data", [&] {
        std::string b = bytes_of(d, true);
        b.resize(b.size() - sizeof(double));
        write_file(b);
        return throws_containing([] { load(); }, "Unsupported particle file layout");
    });
    run("rejects trailing particle data", [&] {
        write_file(bytes_of(d, false) + std::string(1, '\0'));
        return throws_containing([] { load(); }, "Unsupported particle file layout");
    });
    run("rejects an unsupported particle type", [&] {
        std::string b = bytes_of(d, true, true);
        b.back() = static_cast<char>(255);
        write_file(b);
        return throws_containing([] { load(); }, "Unsupported particle type value");
    });
    run("rejects a 64-bit count beyond the supported range", [&] {
        const std::uint64_t n =
            static_cast<std::uint64_t>(std::numeric_limits<std::uint32_t>::max()) + 1;
        write_file(std::string(reinterpret_cast<const char*>(&n), sizeof(n)));
        return throws_containing([] { load(); }, "Unsupported particle file layout");
    });
    const auto bad_value = [&](std::vector<double> Data::*block, std::size_t i, double v,
                               const char* text) {
        Data e = d;
        (e.*block)[i] = v;
        write_file(bytes_of(e, true));
        return throws_containing([] { load(); }, text);
    };
    run("rejects a negative mass", [&] { return bad_value(&Data::mass, 0, -1.0, "Negative mass value"); });
    run("rejects a non-finite mass", [&] {
        return bad_value(&Data::mass, 0, std::nan(""), "Non-finite mass value");
    });
    run("rejects a non-finite position", [&] {
        return bad_value(&Data::x, 1, std::numeric_limits<double>::infinity(),
                         "Non-finite x position value");
    });
    run("rejects a non-finite velocity", [&] {
        return bad_value(&Data::vz, 0, std::nan(""), "Non-finite z velocity value");
    });
    run("accepts a zero-mass particle", [&] {
        Data e = d;
        e.mass[0] = 0.0;
        write_file(bytes_of(e, true));
        return load().mass[0] == 0.0;
    });
    std::filesystem::remove(kFile);
    std::printf("%d failed\n", failed);
    return failed == 0 ? 0 : 1;
}
