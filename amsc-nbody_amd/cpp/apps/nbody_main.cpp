// nbody -- command-line driver of the simulation (the reference's
// apps/simulation/main.cpp:44-333): load a planar binary particle file, run
// N leapfrog steps and print the timing summary that the benchmark scripts
// parse.  Exit status 0 on success, 2 for argument errors (with the usage on
// stderr), 1 for any other failure.
//
// --stream-dir DIR writes sampled, quantized .nbsnap frames of the positions
// at --sample-rate wall-clock Hz into a durable spool, as the reference does
// (the initial state, then after steps when a frame is due, and the final
// state), and marks the run complete.  The stream server (--cluster and its
// network options) is not part of this build: those options are parsed and
// validated, and --cluster is reported as an argument error.  Additions:
// --precision fp64|mixed, --key-bits 30|63 and --gravitational-constant G
// (defaults: fp64, 30, the SI constant).
#include <cerrno>
#include <charconv>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <memory>
#include <stdexcept>
#include <string>
#include <string_view>
#include <utility>

#include "particle.hpp"
#include "simulation.hpp"
#include "streaming/device_capture.hpp"
#include "streaming/frame.hpp"

namespace {

struct Options {
    std::string input;
    std::size_t steps = 0;
    SimulationConfig config{0.0, 0.5, 1.0e-6};
    bool has_time_step = false;
    bool has_steps = false;
    bool help = false;
    bool cluster = false;
    bool streaming = false;
    double sample_rate = 60.0;
    std::size_t maximum_stream_particles = 100000;
    std::string stream_dir = "nbody-stream";
    std::string stream_bind = "127.0.0.1";
};

void usage(std::ostream& out, const char* program)
{
    out << "Usage:\n"
        << "  " << program << " --input FILE --time-step DT --steps N [OPTIONS]\n\n"
        << "Required arguments:\n"
        << "  -i, --input FILE       Binary particle input file\n"
        << "  -t, --time-step DT     Positive shared simulation timestep\n"
        << "  -n, --steps N          Positive number of timesteps\n\n"
        << "Force options:\n"
        << "      --theta VALUE      Barnes-Hut opening angle (default: 0.5)\n"
        << "      --softening VALUE  Non-negative softening length (default: 1e-6)\n"
        << "      --precision P      fp64 or mixed (FP32 traversal; default: fp64)\n"
        << "      --key-bits B       Morton key width, 30 or 63 (default: 30)\n"
        << "      --gravitational-constant G  (default: 6.6743e-11)\n\n"
        << "Graphical stream options:\n"
        << "      --stream-dir DIR   Durable frame queue; also enables output\n"
        << "                         (default: nbody-stream)\n"
        << "      --sample-rate HZ   Wall-clock frame rate (default: 60)\n"
        << "      --stream-max-particles N  Display sample size; 0 sends all "
           "(default: 100000)\n"
        << "      --cluster, --stream-bind ADDR, --stream-port PORT, --stream-server FILE\n"
        << "                         Stream server options (parsed; the server is not\n"
        << "                         available in this build)\n\n"
        << "Other options:\n"
        << "  -h, --help             Show this help and exit\n";
}

std::string_view value_of(int argc, char** argv, int& i, std::string_view option)
{
    if (i + 1 >= argc)
        throw std::invalid_argument("Missing value for " + std::string(option));
    return argv[++i];
}

double parse_double(std::string_view text, std::string_view option)
{
    const std::string s(text);
    char* end = nullptr;
    errno = 0;
    const double v = std::strtod(s.c_str(), &end);
    if (end == s.c_str() || *end != '\0' || errno == ERANGE || !std::isfinite(v))
        throw std::invalid_argument("Invalid numeric value for " + std::string(option) +
                                    ": " + s);
    return v;
}

template <class T>
T parse_unsigned(std::string_view text, std::string_view option, const char* what,
                 bool positive)
{
    T v{};
    const auto r = std::from_chars(text.data(), text.data() + text.size(), v);
    if (r.ec != std::errc{} || r.ptr != text.data() + text.size() || (positive && v == 0))
        throw std::invalid_argument(std::string("Invalid ") + what + " for " +
                                    std::string(option) + ": " + std::string(text));
    return v;
}

Options parse(int argc, char** argv)
{
    Options o;
    for (int i = 1; i < argc; ++i) {
        const std::string_view a(argv[i]);
        if (a == "-h" || a == "--help") {
            o.help = true;
            return o;
        }
        if (a == "-i" || a == "--input") {
            o.input = value_of(argc, argv, i, a);
        } else if (a == "-t" || a == "--time-step") {
            o.config.timeStep = parse_double(value_of(argc, argv, i, a), a);
            o.has_time_step = true;
        } else if (a == "-n" || a == "--steps") {
            o.steps = parse_unsigned<std::size_t>(value_of(argc, argv, i, a), a,
                                                  "positive integer", true);
            o.has_steps = true;
        } else if (a == "--theta") {
            o.config.theta = parse_double(value_of(argc, argv, i, a), a);
        } else if (a == "--softening") {
            o.config.softening = parse_double(value_of(argc, argv, i, a), a);
        } else if (a == "--precision") {
            const std::string_view p = value_of(argc, argv, i, a);
            if (p == "fp64")
                o.config.precision = NBH_PRECISION_FP64;
            else if (p == "mixed")
                o.config.precision = NBH_PRECISION_MIXED;
            else
                throw std::invalid_argument("--precision must be fp64 or mixed");
        } else if (a == "--key-bits") {
            const auto b = parse_unsigned<unsigned>(value_of(argc, argv, i, a), a,
                                                    "key width", true);
            if (b != 30 && b != 63)
                throw std::invalid_argument("--key-bits must be 30 or 63");
            o.config.bitsPerAxis = b == 30 ? 10 : 21;
        } else if (a == "--gravitational-constant") {
            o.config.gravitationalConstant = parse_double(value_of(argc, argv, i, a), a);
        } else if (a == "--cluster") {
            o.cluster = true;
            o.streaming = true;
        } else if (a == "--stream-dir") {
            o.stream_dir = value_of(argc, argv, i, a);
            o.streaming = true;
        } else if (a == "--sample-rate") {
            o.sample_rate = parse_double(value_of(argc, argv, i, a), a);
        } else if (a == "--stream-max-particles") {
            o.maximum_stream_particles = parse_unsigned<std::size_t>(
                value_of(argc, argv, i, a), a, "non-negative integer", false);
        } else if (a == "--stream-bind") {
            o.stream_bind = value_of(argc, argv, i, a);
        } else if (a == "--stream-port") {
            const auto port = parse_unsigned<unsigned>(value_of(argc, argv, i, a), a,
                                                       "TCP port", true);
            if (port > 65535)
                throw std::invalid_argument("Invalid TCP port for " + std::string(a) +
                                            ": " + std::to_string(port));
        } else if (a == "--stream-server") {
            (void)value_of(argc, argv, i, a);
        } else {
            throw std::invalid_argument("Unknown option: " + std::string(a));
        }
    }
    if (o.input.empty())
        throw std::invalid_argument("--input is required");
    if (!o.has_time_step)
        throw std::invalid_argument("--time-step is required");
    if (!o.has_steps)
        throw std::invalid_argument("--steps is required");
    if (o.config.timeStep <= 0.0)
        throw std::invalid_argument("--time-step must be positive");
    if (o.config.theta < 0.0)
        throw std::invalid_argument("--theta must be non-negative");
    if (o.config.softening < 0.0)
        throw std::invalid_argument("--softening must be non-negative");
    if (!(o.config.gravitationalConstant > 0.0))
        throw std::invalid_argument("--gravitational-constant must be positive");
    if (o.sample_rate <= 0.0)
        throw std::invalid_argument("--sample-rate must be positive");
    if (o.stream_dir.empty())
        throw std::invalid_argument("--stream-dir must not be empty");
    if (o.stream_bind.empty())
        throw std::invalid_argument("--stream-bind must not be empty");
    if (o.cluster)
        throw std::invalid_argument(
            "--cluster: the graphical stream server is not available in this build "
            "(--stream-dir writes frames)");
    return o;
}

} // namespace

int main(int argc, char** argv)
{
    try {
        const Options o = parse(argc, argv);
        if (o.help) {
            usage(std::cout, argv[0]);
            return 0;
        }
        std::unique_ptr<nbody::streaming::FrameSpool> spool;
        std::unique_ptr<nbody::streaming::FrameSampler> sampler;
        std::size_t frames_written = 0;
        if (o.streaming) {
            spool = std::make_unique<nbody::streaming::FrameSpool>(o.stream_dir);
            sampler = std::make_unique<nbody::streaming::FrameSampler>(o.sample_rate);
        }

        std::ifstream input(o.input, std::ios::binary);
        if (!input.is_open())
            throw std::runtime_error("Could not open particle input file: " + o.input);
        Particles particles(input);
        Simulation simulation(std::move(particles), o.config);
        const DeviceParticlesView view = simulation.particles();
        const std::size_t count = view.count;
        std::unique_ptr<nbody::streaming::DeviceFrameWriter> writer;
        if (spool)
            writer = std::make_unique<nbody::streaming::DeviceFrameWriter>(
                ConstDeviceParticlesView{view.count, view.mass, view.x, view.y, view.z,
                                         view.vx, view.vy, view.vz, view.type},
                o.maximum_stream_particles);
        std::cout << "Loaded " << count << " particles from " << o.input << '\n';

        using clock = std::chrono::steady_clock;
        const auto t0 = clock::now();
        simulation.initialize();
        const auto t1 = clock::now();
        std::size_t last_captured = std::numeric_limits<std::size_t>::max();
        const auto capture = [&](bool force) {
            if (!sampler)
                return;
            const std::chrono::duration<double> since = clock::now() - t1;
            if (!force && !sampler->isDue(since.count()))
                return;
            writer->write(*spool, sampler->takeSequence(), simulation.stepNumber(),
                          simulation.time());
            ++frames_written;
            last_captured = simulation.stepNumber();
        };
        capture(false);
        for (std::size_t s = 0; s < o.steps; ++s) {
            simulation.step();
            capture(false);
        }
        if (sampler && last_captured != simulation.stepNumber())
            capture(true);
        if (spool)
            (void)spool->markComplete();
        const auto t2 = clock::now();
        const std::chrono::duration<double> init = t1 - t0;
        const std::chrono::duration<double> evolution = t2 - t1;
        const std::chrono::duration<double> total = t2 - t0;
        std::cout << std::setprecision(17) << "Simulation complete\n"
                  << "  steps: " << simulation.stepNumber() << '\n'
                  << "  simulated time: " << simulation.time() << '\n'
                  << "  initialization wall time: " << init.count() << " seconds\n"
                  << "  evolution wall time: " << evolution.count() << " seconds\n"
                  << "  average step time: "
                  << 1000.0 * evolution.count() / static_cast<double>(o.steps) << " ms\n"
                  << "  total wall time: " << total.count() << " seconds\n";
        if (spool)
            std::cout << "  frames written: " << frames_written << '\n'
                      << "  particles per frame: " << writer->sampleParticleCount() << " of "
                      << count << '\n'
                      << "  frame spool: " << spool->directory() << '\n';
        return 0;
    } catch (const std::invalid_argument& e) {
        std::cerr << "Argument error: " << e.what() << "\n\n";
        usage(std::cerr, argv[0]);
        return 2;
    } catch (const std::exception& e) {
        std::cerr << "Simulation failed: " << e.what() << '\n';
        return 1;
    }
}
