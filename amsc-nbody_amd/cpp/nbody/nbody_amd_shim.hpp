// nbody_amd_shim.hpp -- glue between the drop-in C++ headers and the C ABI:
// the status -> exception mapping and the process-wide default context.
#pragma once

#include <hip/hip_runtime_api.h>

#include <stdexcept>
#include <string>

#include "nbody_amd.h"

namespace nbody_amd {

// Rethrows a C ABI status as the exception class the reference throws for
// the same condition (std::invalid_argument, std::logic_error,
// std::runtime_error, std::domain_error, std::overflow_error).
inline void check(int status)
{
    if (status == NBH_OK)
        return;
    const std::string message = nbh_last_error_message();
    switch (status) {
    case NBH_INVALID_ARGUMENT:
        throw std::invalid_argument(message);
    case NBH_LOGIC_ERROR:
        throw std::logic_error(message);
    case NBH_DOMAIN_ERROR:
        throw std::domain_error(message);
    case NBH_OVERFLOW_ERROR:
        throw std::overflow_error(message);
    default:
        throw std::runtime_error(message);
    }
}

inline void check_hip(hipError_t error, const char* what)
{
    if (error != hipSuccess)
        throw std::runtime_error(std::string(what) + ": " + hipGetErrorString(error));
}

// The context every free function runs on: the null stream, as the
// reference's synchronous wrappers use the default stream.  Lives until the
// process ends.
inline nbh_context* default_context()
{
    static nbh_context* ctx = [] {
        nbh_context* c = nullptr;
        check(nbh_context_create(&c, nullptr));
        return c;
    }();
    return ctx;
}

// Gravitational constant used by the reference's Simulation
// (src/particle.hpp:16).
inline constexpr double kSiGravitationalConstant = 6.67430e-11;

} // namespace nbody_amd
