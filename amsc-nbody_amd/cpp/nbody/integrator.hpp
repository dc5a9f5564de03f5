// integrator.hpp -- leapfrog half kick + drift and closing half kick
// (src/integrator.hpp:5-7).
#pragma once

#include <cstddef>

#include "nbody_amd_shim.hpp"

inline void leapfrogKickDrift(double* d_positionX, double* d_positionY,
                              double* d_positionZ, double* d_velocityX,
                              double* d_velocityY, double* d_velocityZ,
                              const double* d_accelerationX,
                              const double* d_accelerationY,
                              const double* d_accelerationZ, std::size_t particleCount,
                              double dt)
{
    nbody_amd::check(nbh_leapfrog_kick_drift(
        nbody_amd::default_context(), d_positionX, d_positionY, d_positionZ, d_velocityX,
        d_velocityY, d_velocityZ, d_accelerationX, d_accelerationY, d_accelerationZ,
        particleCount, dt));
}

inline void leapfrogKick(double* d_velocityX, double* d_velocityY, double* d_velocityZ,
                         const double* d_accelerationX, const double* d_accelerationY,
                         const double* d_accelerationZ, std::size_t particleCount,
                         double dt)
{
    nbody_amd::check(nbh_leapfrog_kick(nbody_amd::default_context(), d_velocityX,
                                       d_velocityY, d_velocityZ, d_accelerationX,
                                       d_accelerationY, d_accelerationZ, particleCount, dt));
}
