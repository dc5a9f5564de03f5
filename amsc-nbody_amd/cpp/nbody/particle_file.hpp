// particle_file.hpp -- reader of the planar binary particle file
// (src/particle.hpp:19-77,145-224), host code only.
//
// Layout: a particle count N as uint64 or uint32, then seven blocks of N
// doubles (mass, x, y, z, vx, vy, vz), optionally followed by N uint8
// particle types.  The header width and the type block are recognised from
// the file size: the 64-bit reading is tried first (count in 1..2^32-1),
// then the 32-bit one, each with and without types; the size must match
// exactly.  Every value must be finite and masses non-negative.  Failures
// throw std::runtime_error with the reference's message texts.
#pragma once

#include <cmath>
#include <cstddef>
#include <cstdint>
#include <istream>
#include <limits>
#include <stdexcept>
#include <string>
#include <vector>

#include "particle_type.hpp"

namespace nbody_amd {

struct ParticleFileData {
    std::vector<double> mass, x, y, z, vx, vy, vz;
    std::vector<std::uint8_t> type;
    std::size_t size() const noexcept { return mass.size(); }
};

struct ParticleFileLayout {
    std::uint32_t count;
    std::streamoff header_bytes;
    bool has_types;
};

inline ParticleFileLayout detect_particle_file_layout(std::istream& in)
{
    in.clear();
    in.seekg(0, std::ios::end);
    const std::streamoff size = in.tellg();
    if (size < 0)
        throw std::runtime_error("Failed to determine particle file size.");
    const auto fits = [&](std::uint64_t n, std::streamoff header, bool types) {
        const std::streamoff body = static_cast<std::streamoff>(n) *
                                    static_cast<std::streamoff>(7 * sizeof(double));
        return size == header + body + (types ? static_cast<std::streamoff>(n) : 0);
    };
    in.clear();
    in.seekg(0, std::ios::beg);
    std::uint64_t n64 = 0;
    in.read(reinterpret_cast<char*>(&n64), sizeof(n64));
    if (in && n64 > 0 && n64 <= std::numeric_limits<std::uint32_t>::max()) {
        const auto h = static_cast<std::streamoff>(sizeof(n64));
        if (fits(n64, h, true))
            return {static_cast<std::uint32_t>(n64), h, true};
        if (fits(n64, h, false))
            return {static_cast<std::uint32_t>(n64), h, false};
    }
    in.clear();
    in.seekg(0, std::ios::beg);
    std::uint32_t n32 = 0;
    in.read(reinterpret_cast<char*>(&n32), sizeof(n32));
    if (in && n32 > 0) {
        const auto h = static_cast<std::streamoff>(sizeof(n32));
        if (fits(n32, h, true))
            return {n32, h, true};
        if (fits(n32, h, false))
            return {n32, h, false};
    }
    throw std::runtime_error(
        "Unsupported particle file layout. Expected planar doubles with a 32-bit "
        "or 64-bit particle count header, optionally followed by one uint8 "
        "particle type per particle.");
}

// Reads and validates a whole file.  `in` must be open and in a good state;
// its initial read position does not matter.
inline ParticleFileData read_particle_file(std::istream& in)
{
    if (!in.good())
        throw std::runtime_error("Invalid or unreadable file stream provided.");
    const ParticleFileLayout layout = detect_particle_file_layout(in);
    if (layout.count == 0)
        throw std::runtime_error("Particle count is zero.");
    const std::size_t n = layout.count;
    in.clear();
    in.seekg(layout.header_bytes, std::ios::beg);
    ParticleFileData d;
    const auto bytes = static_cast<std::streamsize>(n * sizeof(double));
    const auto block = [&](std::vector<double>& out, const char* name, bool non_negative) {
        out.resize(n);
        in.read(reinterpret_cast<char*>(out.data()), bytes);
        if (in.gcount() != bytes)
            throw std::runtime_error(std::string("Incomplete ") + name + " block");
        for (std::size_t i = 0; i < n; ++i) {
            if (!std::isfinite(out[i]))
                throw std::runtime_error(std::string("Non-finite ") + name +
                                         " value at particle " + std::to_string(i));
            if (non_negative && out[i] < 0.0)
                throw std::runtime_error(std::string("Negative ") + name +
                                         " value at particle " + std::to_string(i));
        }
    };
    block(d.mass, "mass", true);
    block(d.x, "x position", false);
    block(d.y, "y position", false);
    block(d.z, "z position", false);
    block(d.vx, "x velocity", false);
    block(d.vy, "y velocity", false);
    block(d.vz, "z velocity", false);
    d.type.assign(n, static_cast<std::uint8_t>(ParticleType::Unknown));
    if (layout.has_types) {
        const auto tb = static_cast<std::streamsize>(n);
        in.read(reinterpret_cast<char*>(d.type.data()), tb);
        if (in.gcount() != tb)
            throw std::runtime_error("Incomplete particle type block");
        for (const std::uint8_t t : d.type)
            if (!isValidParticleType(t))
                throw std::runtime_error("Unsupported particle type value");
    }
    if (!in)
        throw std::runtime_error(
            "Failed to read all expected particle data. File may be truncated.");
    return d;
}

} // namespace nbody_amd
