// streaming/frame.hpp -- the .nbsnap frame format and its durable spool
// (src/streaming/frame.hpp:13-129), host code only.
//
// A frame is a little-endian header followed by interleaved xyz positions and
// optional per-particle types:
//   "NBSNAP01", u32 version, u32 headerBytes, u64 sequence, u64 step,
//   f64 time, u64 sourceParticleCount, u64 particleCount, u32 scalarBytes,
//   u32 components (3), u64 payloadBytes                          -- 72 bytes
//   v1 (72 B):  float32 positions
//   v2 (80 B):  + u64 totalSteps; float32 positions, then u8 types
//   v2 (120 B): + f64 bounds min[3], max[3]; uint16 quantized positions
//   v3 (128 B): + bounds + u64 particleTypeBytes; uint16 positions, u8 types
// A quantized value q decodes to min + (max - min) q / 65535 (min when the
// axis collapsed).  Frames are written to "<name>.part" and renamed into
// place, named run-<session>-frame-<20-digit sequence>.nbsnap; a completed
// run publishes run-<session>.complete the same way.
#pragma once

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <iomanip>
#include <limits>
#include <sstream>
#include <stdexcept>
#include <string>
#include <vector>

#include <unistd.h>

#include "../particle_type.hpp"

namespace nbody::streaming {

inline constexpr std::uint32_t kLegacyFrameVersion = 1;
inline constexpr std::uint32_t kLegacyTypedFrameVersion = 2;
inline constexpr std::uint32_t kQuantizedFrameVersion = 2;
inline constexpr std::uint32_t kFrameVersion = 3;
inline constexpr std::size_t kLegacyFrameHeaderBytes = 72;
inline constexpr std::size_t kLegacyTypedFrameHeaderBytes = 80;
inline constexpr std::size_t kQuantizedFrameHeaderBytes = 120;
inline constexpr std::size_t kFrameHeaderBytes = 128;
inline constexpr std::size_t kPositionComponents = 3;
inline constexpr std::size_t kFrameChunkParticles = 1024 * 1024;
inline constexpr std::uint16_t kQuantizedPositionMaximum = 65535;

struct PositionBounds {
    std::array<double, kPositionComponents> minimum{};
    std::array<double, kPositionComponents> maximum{};
};

struct FrameHeader {
    std::uint32_t version = kFrameVersion;
    std::uint32_t headerBytes = static_cast<std::uint32_t>(kFrameHeaderBytes);
    std::uint64_t sequence = 0;
    std::uint64_t simulationStep = 0;
    double simulationTime = 0.0;
    std::uint64_t sourceParticleCount = 0;
    std::uint64_t particleCount = 0;
    std::uint32_t scalarBytes = sizeof(std::uint16_t);
    std::uint32_t components = kPositionComponents;
    std::uint64_t payloadBytes = 0;
    std::uint64_t totalSteps = 0;
    PositionBounds bounds{};
    std::uint64_t particleTypeBytes = 0;
};

// Chunk callbacks: fill `count` particles starting at particle `offset`.
using PositionChunkReader = std::function<void(std::size_t, std::size_t, float*)>;
using QuantizedPositionChunkReader =
    std::function<void(std::size_t, std::size_t, std::uint16_t*)>;
using ParticleTypeChunkReader =
    std::function<void(std::size_t, std::size_t, std::uint8_t*)>;

namespace detail {

inline constexpr char kMagic[8] = {'N', 'B', 'S', 'N', 'A', 'P', '0', '1'};

template <class U>
void put_le(std::ostream& out, U v)
{
    for (std::size_t b = 0; b < sizeof(U); ++b)
        out.put(static_cast<char>((v >> (8 * b)) & 0xFFu));
}

template <class U>
U get_le(std::istream& in)
{
    U v = 0;
    for (std::size_t b = 0; b < sizeof(U); ++b) {
        const int c = in.get();
        if (c == std::char_traits<char>::eof())
            throw std::runtime_error("Truncated N-body frame header");
        v |= static_cast<U>(static_cast<unsigned char>(c)) << (8 * b);
    }
    return v;
}

inline void put_f64(std::ostream& out, double v)
{
    std::uint64_t bits;
    std::memcpy(&bits, &v, sizeof(bits));
    put_le(out, bits);
}

inline double get_f64(std::istream& in)
{
    const auto bits = get_le<std::uint64_t>(in);
    double v;
    std::memcpy(&v, &bits, sizeof(v));
    return v;
}

inline void write_header(std::ostream& out, const FrameHeader& h)
{
    out.write(kMagic, sizeof(kMagic));
    put_le(out, h.version);
    put_le(out, h.headerBytes);
    put_le(out, h.sequence);
    put_le(out, h.simulationStep);
    put_f64(out, h.simulationTime);
    put_le(out, h.sourceParticleCount);
    put_le(out, h.particleCount);
    put_le(out, h.scalarBytes);
    put_le(out, h.components);
    put_le(out, h.payloadBytes);
    if (h.headerBytes == kLegacyTypedFrameHeaderBytes) {
        put_le(out, h.totalSteps);
    } else if (h.headerBytes == kQuantizedFrameHeaderBytes ||
               h.headerBytes == kFrameHeaderBytes) {
        for (double v : h.bounds.minimum)
            put_f64(out, v);
        for (double v : h.bounds.maximum)
            put_f64(out, v);
        if (h.headerBytes == kFrameHeaderBytes)
            put_le(out, h.particleTypeBytes);
    }
}

inline void validate_bounds(const PositionBounds& b)
{
    for (std::size_t c = 0; c < kPositionComponents; ++c)
        if (!std::isfinite(b.minimum[c]) || !std::isfinite(b.maximum[c]) ||
            b.minimum[c] > b.maximum[c] || !std::isfinite(b.maximum[c] - b.minimum[c]))
            throw std::invalid_argument("Position frame bounds must be finite and ordered");
}

inline std::string default_session_id()
{
    const auto us = std::chrono::duration_cast<std::chrono::microseconds>(
                        std::chrono::system_clock::now().time_since_epoch())
                        .count();
    return std::to_string(us) + "-" + std::to_string(static_cast<unsigned long long>(::getpid()));
}

inline void validate_session_id(const std::string& id)
{
    if (id.empty())
        throw std::invalid_argument("Frame session id cannot be empty");
    for (const char c : id)
        if (!((c >= 'a' && c <= 'z') || (c >= 'A' && c <= 'Z') || (c >= '0' && c <= '9') ||
              c == '-' || c == '_'))
            throw std::invalid_argument(
                "Frame session id may contain only letters, numbers, '-' and '_'");
}

} // namespace detail

// Wall-clock sampler: due at 0, then at most once per 1/rate seconds; a long
// gap yields one frame, not a burst.
class FrameSampler {
public:
    explicit FrameSampler(double samplesPerSecond = 60.0)
        : rate_(samplesPerSecond), interval_(1.0 / samplesPerSecond)
    {
        if (!std::isfinite(samplesPerSecond) || samplesPerSecond <= 0.0)
            throw std::invalid_argument("Frame sample rate must be finite and positive");
    }

    [[nodiscard]] bool isDue(double elapsed)
    {
        if (!std::isfinite(elapsed) || elapsed < 0.0)
            throw std::invalid_argument("Frame sampling time must be finite and non-negative");
        const double slack =
            std::numeric_limits<double>::epsilon() * std::max(1.0, std::abs(elapsed)) * 8.0;
        if (elapsed + slack < next_)
            return false;
        const double behind = std::max(0.0, elapsed - next_);
        next_ += (std::floor(behind / interval_) + 1.0) * interval_;
        return true;
    }

    [[nodiscard]] std::uint64_t takeSequence() noexcept { return sequence_++; }
    [[nodiscard]] double rate() const noexcept { return rate_; }

private:
    double rate_;
    double interval_;
    double next_ = 0.0;
    std::uint64_t sequence_ = 0;
};

class FrameSpool {
public:
    explicit FrameSpool(std::filesystem::path directory, std::string sessionId = {})
        : directory_(std::move(directory)),
          session_(sessionId.empty() ? detail::default_session_id() : std::move(sessionId))
    {
        if (directory_.empty())
            throw std::invalid_argument("Frame spool directory cannot be empty");
        detail::validate_session_id(session_);
        std::error_code e;
        std::filesystem::create_directories(directory_, e);
        if (e)
            throw std::runtime_error("Could not create frame spool directory '" +
                                     directory_.string() + "': " + e.message());
    }

    [[nodiscard]] const std::filesystem::path& directory() const noexcept { return directory_; }
    [[nodiscard]] const std::string& sessionId() const noexcept { return session_; }

    // Version 1: float32 positions.
    std::filesystem::path writePositions(std::uint64_t sequence, std::uint64_t step, double time,
                                         std::size_t count, const PositionChunkReader& reader,
                                         std::size_t sourceCount = 0) const
    {
        return writeTypedPositions(sequence, step, time, 0, count, reader, {}, sourceCount);
    }

    // Float32 positions: version 2 (80-byte header, total step count, then
    // one type byte per particle) with a type reader, version 1 without.
    std::filesystem::path writeTypedPositions(std::uint64_t sequence, std::uint64_t step,
                                              double time, std::uint64_t totalSteps,
                                              std::size_t count,
                                              const PositionChunkReader& reader,
                                              const ParticleTypeChunkReader& typeReader,
                                              std::size_t sourceCount = 0) const
    {
        check_time(time);
        if (typeReader && totalSteps < step)
            throw std::invalid_argument(
                "Frame total step count cannot be smaller than its current step");
        check_count(count, static_cast<bool>(reader));
        sourceCount = check_source(count, sourceCount);
        const std::uint64_t types = typeReader ? count : 0;
        FrameHeader h = header(sequence, step, time, count, sourceCount);
        h.version = typeReader ? kLegacyTypedFrameVersion : kLegacyFrameVersion;
        h.headerBytes = static_cast<std::uint32_t>(typeReader ? kLegacyTypedFrameHeaderBytes
                                                              : kLegacyFrameHeaderBytes);
        h.scalarBytes = sizeof(float);
        h.totalSteps = typeReader ? totalSteps : 0;
        h.particleTypeBytes = types;
        h.payloadBytes = payload_bytes(count, sizeof(float), types);
        return publish(sequence, h, [&](std::ostream& out) {
            write_chunks<float>(out, count, reader);
            if (typeReader)
                write_types(out, count, typeReader);
        });
    }

    // Quantized positions: version 3 with types, version 2 (120-byte
    // header) without.
    std::filesystem::path writeQuantizedPositions(
        std::uint64_t sequence, std::uint64_t step, double time, std::size_t count,
        const PositionBounds& bounds, const QuantizedPositionChunkReader& reader,
        std::size_t sourceCount = 0, const ParticleTypeChunkReader& typeReader = {}) const
    {
        check_time(time);
        check_count(count, static_cast<bool>(reader));
        detail::validate_bounds(bounds);
        sourceCount = check_source(count, sourceCount);
        FrameHeader h = header(sequence, step, time, count, sourceCount);
        if (!typeReader) {
            h.version = kQuantizedFrameVersion;
            h.headerBytes = static_cast<std::uint32_t>(kQuantizedFrameHeaderBytes);
        }
        h.bounds = bounds;
        h.particleTypeBytes = typeReader ? count : 0;
        h.payloadBytes = payload_bytes(count, sizeof(std::uint16_t), h.particleTypeBytes);
        return publish(sequence, h, [&](std::ostream& out) {
            write_chunks<std::uint16_t>(out, count, reader);
            if (typeReader)
                write_types(out, count, typeReader);
        });
    }

    // Publishes run-<session>.complete (idempotent).
    [[nodiscard]] std::filesystem::path markComplete() const
    {
        const auto final_path = directory_ / ("run-" + session_ + ".complete");
        auto part = final_path;
        part += ".part";
        std::error_code e;
        if (std::filesystem::exists(final_path, e) && !e)
            return final_path;
        try {
            {
                std::ofstream out(part, std::ios::trunc | std::ios::binary);
                if (!out.is_open())
                    throw std::runtime_error("Could not create completion marker: " + part.string());
                out << session_ << '\n';
                out.flush();
                if (!out)
                    throw std::runtime_error("Could not flush completion marker: " + part.string());
            }
            std::filesystem::rename(part, final_path, e);
            if (e)
                throw std::runtime_error("Could not publish completion marker '" +
                                         final_path.string() + "': " + e.message());
        } catch (...) {
            std::error_code ignored;
            std::filesystem::remove(part, ignored);
            throw;
        }
        return final_path;
    }

private:
    std::filesystem::path directory_;
    std::string session_;

    static void check_time(double time)
    {
        if (!std::isfinite(time) || time < 0.0)
            throw std::invalid_argument("Frame simulation time is invalid");
    }

    static void check_count(std::size_t count, bool callable)
    {
        if (count == 0)
            throw std::invalid_argument("Cannot write an empty position frame");
        if (!callable)
            throw std::invalid_argument("Position frame reader is not callable");
    }

    // A zero source count means "the sample is the whole population".
    static std::size_t check_source(std::size_t count, std::size_t sourceCount)
    {
        if (sourceCount == 0)
            return count;
        if (sourceCount < count)
            throw std::invalid_argument(
                "Frame source particle count cannot be smaller than its sample");
        return sourceCount;
    }

    static std::uint64_t payload_bytes(std::size_t count, std::size_t scalar,
                                       std::uint64_t typeBytes)
    {
        const auto per = static_cast<std::uint64_t>(kPositionComponents * scalar);
        if (count > std::numeric_limits<std::uint64_t>::max() / per)
            throw std::overflow_error("Position frame payload is too large");
        const std::uint64_t pos = static_cast<std::uint64_t>(count) * per;
        if (pos > std::numeric_limits<std::uint64_t>::max() - typeBytes)
            throw std::overflow_error("Position frame payload is too large");
        return pos + typeBytes;
    }

    static FrameHeader header(std::uint64_t sequence, std::uint64_t step, double time,
                              std::size_t count, std::size_t sourceCount)
    {
        FrameHeader h;
        h.sequence = sequence;
        h.simulationStep = step;
        h.simulationTime = time;
        h.particleCount = count;
        h.sourceParticleCount = sourceCount == 0 ? count : sourceCount;
        return h;
    }

    template <class T, class Reader>
    static void write_chunks(std::ostream& out, std::size_t count, const Reader& reader)
    {
        std::vector<T> buf(std::min(kFrameChunkParticles, count) * kPositionComponents);
        for (std::size_t off = 0; off < count;) {
            const std::size_t n = std::min(kFrameChunkParticles, count - off);
            reader(off, n, buf.data());
            out.write(reinterpret_cast<const char*>(buf.data()),
                      static_cast<std::streamsize>(n * kPositionComponents * sizeof(T)));
            if (!out)
                throw std::runtime_error("Failed while writing frame payload");
            off += n;
        }
    }

    static void write_types(std::ostream& out, std::size_t count,
                            const ParticleTypeChunkReader& reader)
    {
        std::vector<std::uint8_t> buf(std::min(kFrameChunkParticles, count));
        for (std::size_t off = 0; off < count;) {
            const std::size_t n = std::min(kFrameChunkParticles, count - off);
            reader(off, n, buf.data());
            for (std::size_t i = 0; i < n; ++i)
                if (!isValidParticleType(buf[i]))
                    throw std::runtime_error("Particle type reader returned an unsupported value");
            out.write(reinterpret_cast<const char*>(buf.data()), static_cast<std::streamsize>(n));
            if (!out)
                throw std::runtime_error("Failed while writing frame particle types");
            off += n;
        }
    }

    template <class Body>
    std::filesystem::path publish(std::uint64_t sequence, const FrameHeader& h,
                                  const Body& body) const
    {
        std::ostringstream name;
        name << "run-" << session_ << "-frame-" << std::setw(20) << std::setfill('0')
             << sequence << ".nbsnap";
        const auto final_path = directory_ / name.str();
        auto part = final_path;
        part += ".part";
        std::error_code e;
        if (std::filesystem::exists(final_path, e) && !e)
            throw std::runtime_error("Refusing to overwrite queued frame: " + final_path.string());
        try {
            {
                std::ofstream out(part, std::ios::binary | std::ios::trunc);
                if (!out.is_open())
                    throw std::runtime_error("Could not create temporary frame: " + part.string());
                detail::write_header(out, h);
                body(out);
                out.flush();
                if (!out)
                    throw std::runtime_error("Failed to flush frame: " + part.string());
            }
            std::filesystem::rename(part, final_path, e);
            if (e)
                throw std::runtime_error("Could not publish completed frame '" +
                                         final_path.string() + "': " + e.message());
        } catch (...) {
            std::error_code ignored;
            std::filesystem::remove(part, ignored);
            throw;
        }
        return final_path;
    }
};

// Reads and validates a frame header against the four layouts and the file
// size.
[[nodiscard]] inline FrameHeader readFrameHeader(const std::filesystem::path& path)
{
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open())
        throw std::runtime_error("Could not open N-body frame: " + path.string());
    char magic[8] = {};
    in.read(magic, sizeof(magic));
    if (!in || std::memcmp(magic, detail::kMagic, sizeof(magic)) != 0)
        throw std::runtime_error("Invalid N-body frame magic: " + path.string());
    FrameHeader h;
    h.version = detail::get_le<std::uint32_t>(in);
    h.headerBytes = detail::get_le<std::uint32_t>(in);
    h.sequence = detail::get_le<std::uint64_t>(in);
    h.simulationStep = detail::get_le<std::uint64_t>(in);
    h.simulationTime = detail::get_f64(in);
    h.sourceParticleCount = detail::get_le<std::uint64_t>(in);
    h.particleCount = detail::get_le<std::uint64_t>(in);
    h.scalarBytes = detail::get_le<std::uint32_t>(in);
    h.components = detail::get_le<std::uint32_t>(in);
    h.payloadBytes = detail::get_le<std::uint64_t>(in);
    const bool v1 = h.version == kLegacyFrameVersion && h.headerBytes == kLegacyFrameHeaderBytes;
    const bool v2t = h.version == kLegacyTypedFrameVersion &&
                     h.headerBytes == kLegacyTypedFrameHeaderBytes;
    const bool v2q = h.version == kQuantizedFrameVersion &&
                     h.headerBytes == kQuantizedFrameHeaderBytes;
    const bool v3 = h.version == kFrameVersion && h.headerBytes == kFrameHeaderBytes;
    if (v2t) {
        h.totalSteps = detail::get_le<std::uint64_t>(in);
        h.particleTypeBytes = h.particleCount;
    } else if (v2q || v3) {
        for (double& v : h.bounds.minimum)
            v = detail::get_f64(in);
        for (double& v : h.bounds.maximum)
            v = detail::get_f64(in);
        if (v3)
            h.particleTypeBytes = detail::get_le<std::uint64_t>(in);
    }
    const bool ok_layout = (v1 && h.scalarBytes == sizeof(float)) ||
                           (v2t && h.scalarBytes == sizeof(float) &&
                            h.totalSteps >= h.simulationStep) ||
                           (v2q && h.scalarBytes == sizeof(std::uint16_t)) ||
                           (v3 && h.scalarBytes == sizeof(std::uint16_t) &&
                            h.particleTypeBytes == h.particleCount);
    bool ok = ok_layout && h.components == kPositionComponents && h.particleCount > 0 &&
              h.sourceParticleCount >= h.particleCount && std::isfinite(h.simulationTime) &&
              h.simulationTime >= 0.0;
    if (ok) {
        const std::uint64_t per = kPositionComponents * static_cast<std::uint64_t>(h.scalarBytes);
        ok = h.particleCount <= std::numeric_limits<std::uint64_t>::max() / per;
        if (ok) {
            const std::uint64_t pos = h.particleCount * per;
            ok = pos <= std::numeric_limits<std::uint64_t>::max() - h.particleTypeBytes &&
                 h.payloadBytes == pos + h.particleTypeBytes &&
                 h.payloadBytes <= std::numeric_limits<std::uint64_t>::max() - h.headerBytes;
        }
    }
    if (ok && (v2q || v3)) {
        try {
            detail::validate_bounds(h.bounds);
        } catch (const std::invalid_argument&) {
            ok = false;
        }
    }
    if (!ok)
        throw std::runtime_error("Unsupported or inconsistent N-body frame header: " +
                                 path.string());
    std::error_code e;
    const auto size = std::filesystem::file_size(path, e);
    if (e || size != h.headerBytes + h.payloadBytes)
        throw std::runtime_error("N-body frame size does not match its header: " + path.string());
    return h;
}

// Interleaved float32 xyz; quantized payloads are decoded with the bounds.
[[nodiscard]] inline std::vector<float> readFramePositions(const std::filesystem::path& path)
{
    const FrameHeader h = readFrameHeader(path);
    const std::size_t n = static_cast<std::size_t>(h.particleCount) * kPositionComponents;
    std::vector<float> out(n);
    std::ifstream in(path, std::ios::binary);
    in.seekg(static_cast<std::streamoff>(h.headerBytes));
    if (h.scalarBytes == sizeof(float)) {
        in.read(reinterpret_cast<char*>(out.data()), static_cast<std::streamsize>(n * sizeof(float)));
        if (!in)
            throw std::runtime_error("Failed while reading N-body frame payload: " + path.string());
        return out;
    }
    std::vector<std::uint16_t> q(n);
    in.read(reinterpret_cast<char*>(q.data()), static_cast<std::streamsize>(n * 2));
    if (!in)
        throw std::runtime_error("Failed while reading N-body frame payload: " + path.string());
    for (std::size_t i = 0; i < n; ++i) {
        const std::size_t c = i % kPositionComponents;
        const double lo = h.bounds.minimum[c];
        const double range = h.bounds.maximum[c] - lo;
        out[i] = static_cast<float>(range == 0.0
                                        ? lo
                                        : lo + range * static_cast<double>(q[i]) /
                                                   static_cast<double>(kQuantizedPositionMaximum));
    }
    return out;
}

// One type per particle; frames without a type block read as Unknown.
[[nodiscard]] inline std::vector<std::uint8_t>
readFrameParticleTypes(const std::filesystem::path& path)
{
    const FrameHeader h = readFrameHeader(path);
    std::vector<std::uint8_t> types(static_cast<std::size_t>(h.particleCount),
                                    static_cast<std::uint8_t>(ParticleType::Unknown));
    if (h.particleTypeBytes == 0)
        return types;
    std::ifstream in(path, std::ios::binary);
    in.seekg(static_cast<std::streamoff>(h.headerBytes + h.particleCount *
                                                             kPositionComponents * h.scalarBytes));
    in.read(reinterpret_cast<char*>(types.data()), static_cast<std::streamsize>(types.size()));
    if (!in)
        throw std::runtime_error("Failed while reading frame particle types: " + path.string());
    for (const std::uint8_t t : types)
        if (!isValidParticleType(t))
            throw std::runtime_error("N-body frame contains an invalid particle type: " +
                                     path.string());
    return types;
}

} // namespace nbody::streaming
