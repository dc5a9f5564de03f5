// Frame format, spool and sampler cases (tests/unit/cpu/streaming_frame_test.cpp)
// on the host-only streaming/frame.hpp, plus header-corruption and
// session-id cases that the reader and the spool define.  Prints one line per
// case; exit status 0 when every case passes.
#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <functional>
#include <random>
#include <stdexcept>
#include <string>
#include <vector>

#include "streaming/frame.hpp"

namespace fsx = nbody::streaming;
namespace fs = std::filesystem;

namespace {

template <class U>
void le(std::ostream& o, U v)
{
    for (std::size_t b = 0; b < sizeof(U); ++b)
        o.put(static_cast<char>((v >> (8 * b)) & 0xFFu));
}

void le_f64(std::ostream& o, double v)
{
    std::uint64_t b;
    std::memcpy(&b, &v, 8);
    le(o, b);
}

// A version-1 frame written byte by byte: one particle at (1.5, -2, 4.25),
// sequence 3, step 9, time 0.25.
void write_v1(const fs::path& p)
{
    std::ofstream o(p, std::ios::binary);
    o.write("NBSNAP01", 8);
    le(o, std::uint32_t{1});
    le(o, std::uint32_t{72});
    le(o, std::uint64_t{3});
    le(o, std::uint64_t{9});
    le_f64(o, 0.25);
    le(o, std::uint64_t{1});
    le(o, std::uint64_t{1});
    le(o, std::uint32_t{4});
    le(o, std::uint32_t{3});
    le(o, std::uint64_t{12});
    for (float f : {1.5F, -2.0F, 4.25F}) {
        std::uint32_t b;
        std::memcpy(&b, &f, 4);
        le(o, b);
    }
}

fs::path fresh_dir()
{
    std::random_device rd;
    const auto token = (static_cast<std::uint64_t>(rd()) << 32) | rd();
    return fs::temp_directory_path() / ("nbody-amd-frame-test-" + std::to_string(token));
}

template <class E>
bool throws(const std::function<void()>& f, const std::string& text = {})
{
    try {
        f();
    } catch (const E& e) {
        if (text.empty() || std::string(e.what()).find(text) != std::string::npos)
            return true;
        std::printf("    unexpected message: %s\n", e.what());
        return false;
    } catch (const std::exception& e) {
        std::printf("    wrong exception: %s\n", e.what());
        return false;
    }
    std::printf("    no exception\n");
    return false;
}

fsx::QuantizedPositionChunkReader from(const std::vector<std::uint16_t>& q)
{
    return [&q](std::size_t off, std::size_t n, std::uint16_t* out) {
        std::copy_n(q.data() + 3 * off, 3 * n, out);
    };
}

std::string read_all(const fs::path& p)
{
    std::ifstream in(p, std::ios::binary);
    return std::string(std::istreambuf_iterator<char>(in), {});
}

void write_all(const fs::path& p, const std::string& s)
{
    std::ofstream o(p, std::ios::binary | std::ios::trunc);
    o.write(s.data(), static_cast<std::streamsize>(s.size()));
}

} // namespace

int main()
{
    int failed = 0;
    int index = 0;
    const auto run = [&](const char* name, const std::function<bool(const fs::path&)>& f) {
        const fs::path dir = fresh_dir() += "-" + std::to_string(index++);
        bool ok = false;
        try {
            ok = f(dir);
        } catch (const std::exception& e) {
            std::printf("    exception: %s\n", e.what());
        }
        std::error_code ignored;
        fs::remove_all(dir, ignored);
        std::printf("%s %s\n", ok ? "PASS" : "FAIL", name);
        failed += ok ? 0 : 1;
    };

    run("sampler: initial state at 60 wall-clock Hz", [](const fs::path&) {
        fsx::FrameSampler s(60.0);
        bool ok = s.isDue(0.0) && s.takeSequence() == 0;
        ok = ok && !s.isDue(0.01) && s.isDue(1.0 / 60.0) && s.takeSequence() == 1;
        ok = ok && !s.isDue(0.02) && s.isDue(2.0 / 60.0) && s.takeSequence() == 2;
        return ok && s.rate() == 60.0;
    });
    run("sampler: one frame when a step skips intervals", [](const fs::path&) {
        fsx::FrameSampler s(60.0);
        return s.isDue(0.0) && s.isDue(10.0) && !s.isDue(10.0);
    });
    run("sampler: rejects an invalid rate and time", [](const fs::path&) {
        fsx::FrameSampler s;
        return throws<std::invalid_argument>([] { fsx::FrameSampler bad(0.0); }) &&
               throws<std::invalid_argument>([] {
                   fsx::FrameSampler bad(std::numeric_limits<double>::infinity());
               }) &&
               throws<std::invalid_argument>([&] { (void)s.isDue(-1.0); }) &&
               throws<std::invalid_argument>([&] { (void)s.isDue(std::nan("")); });
    });
    run("writes a self-describing version-3 frame atomically", [](const fs::path& dir) {
        const fsx::PositionBounds b{{-2.5, 3.5, -5.0}, {1.25, 4.75, 6.125}};
        const std::vector<std::uint16_t> q{65535, 0, 0, 0, 65535, 65535};
        const std::array<std::uint8_t, 2> types{1, 4};
        fsx::FrameSpool spool(dir, "test-session");
        const fs::path frame = spool.writeQuantizedPositions(
            7, 42, 0.5, 2, b, from(q), 20,
            [&](std::size_t off, std::size_t n, std::uint8_t* out) {
                std::copy_n(types.data() + off, n, out);
            });
        bool ok = frame.parent_path() == dir && frame.extension() == ".nbsnap" &&
                  frame.filename() ==
                      "run-test-session-frame-00000000000000000007.nbsnap" &&
                  !fs::exists(frame.string() + ".part") &&
                  fs::file_size(frame) == 128 + 12 + 2;
        const fsx::FrameHeader h = fsx::readFrameHeader(frame);
        ok = ok && h.version == 3 && h.headerBytes == 128 && h.sequence == 7 &&
             h.simulationStep == 42 && h.simulationTime == 0.5 &&
             h.sourceParticleCount == 20 && h.particleCount == 2 && h.components == 3 &&
             h.scalarBytes == 2 && h.payloadBytes == 14 && h.particleTypeBytes == 2 &&
             h.bounds.minimum == b.minimum && h.bounds.maximum == b.maximum;
        const std::string bytes = read_all(frame);
        std::array<std::uint16_t, 6> payload{};
        std::memcpy(payload.data(), bytes.data() + 128, 12);
        ok = ok && std::equal(payload.begin(), payload.end(), q.begin());
        ok = ok && fsx::readFramePositions(frame) ==
                       std::vector<float>{1.25F, 3.5F, -5.0F, -2.5F, 4.75F, 6.125F};
        return ok && fsx::readFrameParticleTypes(frame) == std::vector<std::uint8_t>{1, 4};
    });
    run("writes a typed float frame with the total step count", [](const fs::path& dir) {
        const std::array<float, 6> p{1, 2, 3, 4, 5, 6};
        const std::array<std::uint8_t, 2> types{1, 2};
        fsx::FrameSpool spool(dir, "typed-session");
        const fs::path frame = spool.writeTypedPositions(
            3, 7, 1.25, 40, 2,
            [&](std::size_t off, std::size_t n, float* out) {
                std::copy_n(p.data() + 3 * off, 3 * n, out);
            },
            [&](std::size_t off, std::size_t n, std::uint8_t* out) {
                std::copy_n(types.data() + off, n, out);
            },
            20);
        const fsx::FrameHeader h = fsx::readFrameHeader(frame);
        bool ok = h.version == 2 && h.headerBytes == 80 && h.sequence == 3 &&
                  h.simulationStep == 7 && h.totalSteps == 40 && h.particleCount == 2 &&
                  h.payloadBytes == sizeof(p) + 2;
        const std::string bytes = read_all(frame);
        ok = ok && std::memcmp(bytes.data() + 80, p.data(), sizeof(p)) == 0 &&
             static_cast<std::uint8_t>(bytes[80 + 24]) == 1 &&
             static_cast<std::uint8_t>(bytes[80 + 25]) == 2;
        ok = ok && fsx::readFramePositions(frame) == std::vector<float>(p.begin(), p.end());
        return ok && fsx::readFrameParticleTypes(frame) == std::vector<std::uint8_t>{1, 2};
    });
    run("removes the partial file when capture fails", [](const fs::path& dir) {
        fsx::FrameSpool spool(dir, "failed-session");
        const fsx::PositionBounds b{{0, 0, 0}, {1, 1, 1}};
        const bool ok = throws<std::runtime_error>(
            [&] {
                (void)spool.writeQuantizedPositions(
                    0, 0, 0.0, 2, b, [](std::size_t, std::size_t, std::uint16_t*) {
                        throw std::runtime_error("copy failed");
                    });
            },
            "copy failed");
        return ok && fs::is_empty(dir);
    });
    run("refuses to overwrite an unconsumed frame", [](const fs::path& dir) {
        fsx::FrameSpool spool(dir, "durable-session");
        const fsx::PositionBounds b{{0, 0, 0}, {0, 0, 0}};
        const std::vector<std::uint16_t> q{0, 0, 0};
        (void)spool.writeQuantizedPositions(0, 0, 0.0, 1, b, from(q));
        return throws<std::runtime_error>(
            [&] { (void)spool.writeQuantizedPositions(0, 1, 0.1, 1, b, from(q)); },
            "Refusing to overwrite queued frame: ");
    });
    run("decodes collapsed quantization axes", [](const fs::path& dir) {
        fsx::FrameSpool spool(dir, "collapsed-axis");
        const fsx::PositionBounds b{{8.0, -1.0, 5.0}, {8.0, 3.0, 5.0}};
        const std::vector<std::uint16_t> q{65535, 32768, 12345};
        const fs::path frame = spool.writeQuantizedPositions(0, 0, 0.0, 1, b, from(q));
        const std::vector<float> p = fsx::readFramePositions(frame);
        const fsx::FrameHeader h = fsx::readFrameHeader(frame);
        return h.version == 2 && h.headerBytes == 120 && p.size() == 3 && p[0] == 8.0F &&
               std::abs(p[1] - 1.0F) <= 4.0F / 65535.0F && p[2] == 5.0F &&
               fsx::readFrameParticleTypes(frame) == std::vector<std::uint8_t>{0};
    });
    run("reads a legacy float frame", [](const fs::path& dir) {
        fs::create_directories(dir);
        const fs::path frame = dir / "legacy.nbsnap";
        write_v1(frame);
        const fsx::FrameHeader h = fsx::readFrameHeader(frame);
        return h.version == 1 && h.headerBytes == 72 && h.scalarBytes == 4 &&
               h.sequence == 3 && h.simulationStep == 9 && h.simulationTime == 0.25 &&
               fsx::readFramePositions(frame) == std::vector<float>{1.5F, -2.0F, 4.25F} &&
               fsx::readFrameParticleTypes(frame) == std::vector<std::uint8_t>{0};
    });
    run("publishes the completion marker atomically", [](const fs::path& dir) {
        fsx::FrameSpool spool(dir, "completed-session");
        const fs::path marker = spool.markComplete();
        return marker.filename() == "run-completed-session.complete" &&
               fs::is_regular_file(marker) && !fs::exists(marker.string() + ".part") &&
               read_all(marker) == "completed-session\n" && spool.markComplete() == marker;
    });
    run("rejects invalid quantization bounds", [](const fs::path& dir) {
        fsx::FrameSpool spool(dir, "invalid-bounds");
        const std::vector<std::uint16_t> q{0, 0, 0};
        const auto inf = std::numeric_limits<double>::infinity();
        return throws<std::invalid_argument>(
                   [&] {
                       (void)spool.writeQuantizedPositions(
                           0, 0, 0.0, 1, fsx::PositionBounds{{1, 0, 0}, {-1, 0, 0}}, from(q));
                   },
                   "Position frame bounds must be finite and ordered") &&
               throws<std::invalid_argument>([&] {
                   (void)spool.writeQuantizedPositions(
                       0, 0, 0.0, 1, fsx::PositionBounds{{0, 0, 0}, {inf, 0, 0}}, from(q));
               }) &&
               throws<std::invalid_argument>([&] {
                   (void)spool.writeQuantizedPositions(
                       0, 0, 0.0, 1, fsx::PositionBounds{{-1e308, 0, 0}, {1e308, 0, 0}},
                       from(q));
               });
    });
    // Cases beyond the reference's unit tests: the reader's validation and
    // the spool's argument checks.
    run("writes a version-1 frame without types", [](const fs::path& dir) {
        fsx::FrameSpool spool(dir, "v1");
        const std::array<float, 3> p{0.5F, -1.0F, 2.0F};
        const fs::path frame = spool.writePositions(
            5, 6, 0.75, 1,
            [&](std::size_t, std::size_t, float* out) { std::copy_n(p.data(), 3, out); }, 4);
        const fsx::FrameHeader h = fsx::readFrameHeader(frame);
        return h.version == 1 && h.headerBytes == 72 && h.sourceParticleCount == 4 &&
               h.payloadBytes == 12 && fs::file_size(frame) == 84 &&
               fsx::readFramePositions(frame) == std::vector<float>(p.begin(), p.end());
    });
    run("rejects a bad magic, a corrupt header and a size mismatch", [](const fs::path& dir) {
        fsx::FrameSpool spool(dir, "corrupt");
        const fsx::PositionBounds b{{0, 0, 0}, {1, 1, 1}};
        const std::vector<std::uint16_t> q{1, 2, 3};
        const fs::path frame = spool.writeQuantizedPositions(0, 0, 0.0, 1, b, from(q));
        const std::string good = read_all(frame);
        const fs::path other = dir / "other.nbsnap";
        std::string s = good;
        s[0] = 'X';
        write_all(other, s);
        bool ok = throws<std::runtime_error>([&] { (void)fsx::readFrameHeader(other); },
                                             "Invalid N-body frame magic: ");
        s = good;
        s[60] = 4; // components
        write_all(other, s);
        ok = ok && throws<std::runtime_error>([&] { (void)fsx::readFrameHeader(other); },
                                              "Unsupported or inconsistent N-body frame header: ");
        s = good;
        s[8] = 3; // version 3 with a 120-byte header
        write_all(other, s);
        ok = ok && throws<std::runtime_error>([&] { (void)fsx::readFrameHeader(other); },
                                              "Unsupported or inconsistent N-body frame header: ");
        s = good;
        std::memcpy(&s[72], "\x00\x00\x00\x00\x00\x00\xf0\x7f", 8); // min x = +inf
        write_all(other, s);
        ok = ok && throws<std::runtime_error>([&] { (void)fsx::readFrameHeader(other); },
                                              "Unsupported or inconsistent N-body frame header: ");
        write_all(other, good + std::string(1, '\0'));
        ok = ok && throws<std::runtime_error>([&] { (void)fsx::readFrameHeader(other); },
                                              "N-body frame size does not match its header: ");
        write_all(other, good.substr(0, 40));
        return ok && throws<std::runtime_error>([&] { (void)fsx::readFrameHeader(other); });
    });
    run("rejects an invalid type in a frame and from a type reader", [](const fs::path& dir) {
        fsx::FrameSpool spool(dir, "types");
        const fsx::PositionBounds b{{0, 0, 0}, {1, 1, 1}};
        const std::vector<std::uint16_t> q{1, 2, 3};
        bool ok = throws<std::runtime_error>(
            [&] {
                (void)spool.writeQuantizedPositions(
                    0, 0, 0.0, 1, b, from(q), 0,
                    [](std::size_t, std::size_t, std::uint8_t* t) { t[0] = 5; });
            },
            "Particle type reader returned an unsupported value");
        ok = ok && fs::is_empty(dir);
        const fs::path frame = spool.writeQuantizedPositions(
            1, 0, 0.0, 1, b, from(q), 0,
            [](std::size_t, std::size_t, std::uint8_t* t) { t[0] = 3; });
        std::string s = read_all(frame);
        s.back() = 9;
        write_all(frame, s);
        return ok && throws<std::runtime_error>(
                         [&] { (void)fsx::readFrameParticleTypes(frame); },
                         "N-body frame contains an invalid particle type: ");
    });
    run("validates session ids and write arguments", [](const fs::path& dir) {
        const fsx::PositionBounds b{{0, 0, 0}, {1, 1, 1}};
        const std::vector<std::uint16_t> q{1, 2, 3, 4, 5, 6};
        fsx::FrameSpool spool(dir, "args");
        fsx::FrameSpool generated(dir);
        const std::string id = generated.sessionId();
        const bool id_ok = !id.empty() && id.find('-') != std::string::npos &&
                           std::all_of(id.begin(), id.end(), [](char c) {
                               return (c >= '0' && c <= '9') || c == '-';
                           });
        return id_ok &&
               throws<std::invalid_argument>([&] { fsx::FrameSpool s(dir, "a/b"); }) &&
               throws<std::invalid_argument>([&] { fsx::FrameSpool s(dir, "a b"); }) &&
               throws<std::invalid_argument>([&] { fsx::FrameSpool s{fs::path{}}; }) &&
               throws<std::invalid_argument>(
                   [&] { (void)spool.writeQuantizedPositions(0, 0, -1.0, 2, b, from(q)); },
                   "Frame simulation time is invalid") &&
               throws<std::invalid_argument>(
                   [&] { (void)spool.writeQuantizedPositions(0, 0, 0.0, 0, b, from(q)); },
                   "Cannot write an empty position frame") &&
               throws<std::invalid_argument>(
                   [&] { (void)spool.writeQuantizedPositions(0, 0, 0.0, 2, b, {}); },
                   "Position frame reader is not callable") &&
               throws<std::invalid_argument>(
                   [&] { (void)spool.writeQuantizedPositions(0, 0, 0.0, 2, b, from(q), 1); },
                   "Frame source particle count cannot be smaller than its sample") &&
               throws<std::invalid_argument>(
                   [&] {
                       (void)spool.writeTypedPositions(
                           0, 5, 0.0, 4, 1, [](std::size_t, std::size_t, float*) {},
                           [](std::size_t, std::size_t, std::uint8_t*) {});
                   },
                   "Frame total step count cannot be smaller than its current step") &&
               fs::is_empty(dir);
    });
    run("writes multi-chunk payloads in order", [](const fs::path& dir) {
        const std::size_t n = fsx::kFrameChunkParticles + 3;
        std::vector<std::uint16_t> q(3 * n);
        for (std::size_t i = 0; i < q.size(); ++i)
            q[i] = static_cast<std::uint16_t>(i * 2654435761u >> 16);
        std::vector<std::uint8_t> t(n);
        for (std::size_t i = 0; i < n; ++i)
            t[i] = static_cast<std::uint8_t>(i % 5);
        fsx::FrameSpool spool(dir, "chunks");
        const fsx::PositionBounds b{{0, -1, 2}, {65535, 65534, 65537}};
        const fs::path frame = spool.writeQuantizedPositions(
            0, 0, 0.0, n, b, from(q), 0, [&](std::size_t off, std::size_t m, std::uint8_t* o) {
                std::copy_n(t.data() + off, m, o);
            });
        const std::vector<float> p = fsx::readFramePositions(frame);
        bool ok = p.size() == 3 * n && fsx::readFrameParticleTypes(frame) == t;
        // lo + range * q / 65535 with range 65535 is exact for every q.
        for (std::size_t i = 0; ok && i < 3 * n; ++i)
            ok = p[i] == static_cast<float>(b.minimum[i % 3] + q[i]);
        return ok;
    });

    std::printf("%d failed\n", failed);
    return failed == 0 ? 0 : 1;
}
