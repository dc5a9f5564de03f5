// frame_io.cpp -- the .nbsnap frame spool and readers behind the C ABI (host
// code; the format itself is cpp/nbody/streaming/frame.hpp, the one the C++
// DeviceFrameWriter and the command-line driver use).
#include "nbody_amd.h"

#include <cstring>
#include <stdexcept>
#include <string>

#include "streaming/frame.hpp"

namespace nbh {
void set_last_error(int code, const std::string& message);
}

namespace {

namespace fs = nbody::streaming;

template <class F>
int guarded_frame(F&& f)
{
    try {
        f();
        nbh::set_last_error(0, "");
        return NBH_OK;
    } catch (const std::invalid_argument& e) {
        nbh::set_last_error(0, e.what());
        return NBH_INVALID_ARGUMENT;
    } catch (const std::overflow_error& e) {
        nbh::set_last_error(0, e.what());
        return NBH_OVERFLOW_ERROR;
    } catch (const std::bad_alloc&) {
        nbh::set_last_error(0, "host allocation failed");
        return NBH_OUT_OF_MEMORY;
    } catch (const std::exception& e) {
        nbh::set_last_error(0, e.what());
        return NBH_RUNTIME_ERROR;
    }
}

fs::FrameSpool open_spool(const char* directory, const char* session_id)
{
    if (directory == nullptr)
        throw std::invalid_argument("Frame spool directory cannot be empty");
    return fs::FrameSpool(directory, session_id ? std::string(session_id) : std::string());
}

// Copies the published path; a buffer too small for it is an error only
// after the frame is on disk, so the caller can still find it.
void put_path(const std::filesystem::path& p, char* out, size_t capacity)
{
    if (out == nullptr || capacity == 0)
        return;
    const std::string s = p.string();
    if (s.size() + 1 > capacity)
        throw std::invalid_argument("path buffer is too small for the published frame path");
    std::memcpy(out, s.c_str(), s.size() + 1);
}

fs::ParticleTypeChunkReader type_reader(const uint8_t* types)
{
    if (types == nullptr)
        return {};
    return [types](std::size_t off, std::size_t n, std::uint8_t* dst) {
        std::memcpy(dst, types + off, n);
    };
}

} // namespace

extern "C" int nbh_frame_write_quantized(const char* directory, const char* session_id,
                                         uint64_t sequence, uint64_t step, double time,
                                         size_t count, size_t source_count,
                                         const double* bounds, const uint16_t* h_positions,
                                         const uint8_t* h_types, char* path_out,
                                         size_t path_capacity)
{
    return guarded_frame([&] {
        if (bounds == nullptr || h_positions == nullptr)
            throw std::invalid_argument("frame bounds and positions must not be null");
        const fs::FrameSpool spool = open_spool(directory, session_id);
        fs::PositionBounds b;
        for (int c = 0; c < 3; ++c) {
            b.minimum[c] = bounds[c];
            b.maximum[c] = bounds[3 + c];
        }
        const fs::QuantizedPositionChunkReader reader =
            [h_positions](std::size_t off, std::size_t n, std::uint16_t* dst) {
                std::memcpy(dst, h_positions + 3 * off, 3 * n * sizeof(std::uint16_t));
            };
        put_path(spool.writeQuantizedPositions(sequence, step, time, count, b, reader,
                                               source_count, type_reader(h_types)),
                 path_out, path_capacity);
    });
}

extern "C" int nbh_frame_write_float(const char* directory, const char* session_id,
                                     uint64_t sequence, uint64_t step, double time,
                                     uint64_t total_steps, size_t count, size_t source_count,
                                     const float* h_positions, const uint8_t* h_types,
                                     char* path_out, size_t path_capacity)
{
    return guarded_frame([&] {
        if (h_positions == nullptr)
            throw std::invalid_argument("frame positions must not be null");
        const fs::FrameSpool spool = open_spool(directory, session_id);
        const fs::PositionChunkReader reader = [h_positions](std::size_t off, std::size_t n,
                                                             float* dst) {
            std::memcpy(dst, h_positions + 3 * off, 3 * n * sizeof(float));
        };
        put_path(spool.writeTypedPositions(sequence, step, time, total_steps, count, reader,
                                           type_reader(h_types), source_count),
                 path_out, path_capacity);
    });
}

extern "C" int nbh_frame_mark_complete(const char* directory, const char* session_id,
                                       char* path_out, size_t path_capacity)
{
    return guarded_frame([&] {
        if (session_id == nullptr)
            throw std::invalid_argument("Frame session id cannot be empty");
        const fs::FrameSpool spool = open_spool(directory, session_id);
        put_path(spool.markComplete(), path_out, path_capacity);
    });
}

extern "C" int nbh_frame_read_header(const char* path, nbh_frame_header* out)
{
    return guarded_frame([&] {
        if (path == nullptr || out == nullptr)
            throw std::invalid_argument("frame path and header must not be null");
        const fs::FrameHeader h = fs::readFrameHeader(path);
        out->version = h.version;
        out->header_bytes = h.headerBytes;
        out->sequence = h.sequence;
        out->simulation_step = h.simulationStep;
        out->simulation_time = h.simulationTime;
        out->source_particle_count = h.sourceParticleCount;
        out->particle_count = h.particleCount;
        out->scalar_bytes = h.scalarBytes;
        out->components = h.components;
        out->payload_bytes = h.payloadBytes;
        out->total_steps = h.totalSteps;
        for (int c = 0; c < 3; ++c) {
            out->bounds_min[c] = h.bounds.minimum[c];
            out->bounds_max[c] = h.bounds.maximum[c];
        }
        out->particle_type_bytes = h.particleTypeBytes;
    });
}

extern "C" int nbh_frame_read(const char* path, size_t capacity, float* h_positions,
                              uint8_t* h_types, size_t* count)
{
    return guarded_frame([&] {
        if (path == nullptr || h_positions == nullptr || count == nullptr)
            throw std::invalid_argument("frame path, positions and count must not be null");
        const std::vector<float> p = fs::readFramePositions(path);
        const std::size_t n = p.size() / fs::kPositionComponents;
        if (n > capacity)
            throw std::invalid_argument("frame holds more particles than capacity");
        std::memcpy(h_positions, p.data(), p.size() * sizeof(float));
        if (h_types) {
            const std::vector<std::uint8_t> t = fs::readFrameParticleTypes(path);
            std::memcpy(h_types, t.data(), t.size());
        }
        *count = n;
    });
}
