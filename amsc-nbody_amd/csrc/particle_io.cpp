// particle_io.cpp -- the planar binary particle file reader behind the C ABI
// (host code; the reader itself is cpp/nbody/particle_file.hpp, the same
// one the C++ Particles class uses).
#include "nbody_amd.h"

#include <cstring>
#include <fstream>
#include <stdexcept>
#include <string>

#include "particle_file.hpp"

namespace nbh {
void set_last_error(int code, const std::string& message);
}

namespace {

int fail_with(int status, const std::string& message)
{
    nbh::set_last_error(0, message);
    return status;
}

template <class F>
int guarded_io(F&& f)
{
    try {
        f();
        nbh::set_last_error(0, "");
        return NBH_OK;
    } catch (const std::invalid_argument& e) {
        return fail_with(NBH_INVALID_ARGUMENT, e.what());
    } catch (const std::exception& e) {
        return fail_with(NBH_RUNTIME_ERROR, e.what());
    }
}

std::ifstream open_particle_file(const char* path)
{
    if (path == nullptr)
        throw std::invalid_argument("particle file path must not be null");
    std::ifstream in(path, std::ios::binary);
    if (!in.is_open())
        throw std::runtime_error(std::string("Could not open particle input file: ") + path);
    return in;
}

} // namespace

extern "C" int nbh_particle_file_info(const char* path, size_t* count, int* has_types)
{
    return guarded_io([&] {
        if (count == nullptr)
            throw std::invalid_argument("count must not be null");
        std::ifstream in = open_particle_file(path);
        const nbody_amd::ParticleFileLayout layout = nbody_amd::detect_particle_file_layout(in);
        *count = layout.count;
        if (has_types)
            *has_types = layout.has_types ? 1 : 0;
    });
}

extern "C" int nbh_particle_file_read(const char* path, size_t capacity, double* h_mass,
                                      double* h_x, double* h_y, double* h_z, double* h_vx,
                                      double* h_vy, double* h_vz, uint8_t* h_type,
                                      size_t* count)
{
    return guarded_io([&] {
        if (!h_mass || !h_x || !h_y || !h_z || !h_vx || !h_vy || !h_vz || !count)
            throw std::invalid_argument("particle arrays must not be null");
        std::ifstream in = open_particle_file(path);
        const nbody_amd::ParticleFileData d = nbody_amd::read_particle_file(in);
        if (d.size() > capacity)
            throw std::invalid_argument("particle file holds more particles than capacity");
        const size_t bytes = d.size() * sizeof(double);
        std::memcpy(h_mass, d.mass.data(), bytes);
        std::memcpy(h_x, d.x.data(), bytes);
        std::memcpy(h_y, d.y.data(), bytes);
        std::memcpy(h_z, d.z.data(), bytes);
        std::memcpy(h_vx, d.vx.data(), bytes);
        std::memcpy(h_vy, d.vy.data(), bytes);
        std::memcpy(h_vz, d.vz.data(), bytes);
        if (h_type)
            std::memcpy(h_type, d.type.data(), d.size());
        *count = d.size();
    });
}
