// particle_type.hpp -- the per-particle type codes of the particle file
// format (src/particle_type.hpp:7-18).
#pragma once

#include <cstdint>

enum class ParticleType : std::uint8_t {
    Unknown = 0,
    Star = 1,
    Planet = 2,
    Moon = 3,
    Asteroid = 4,
};

[[nodiscard]] inline constexpr bool isValidParticleType(std::uint8_t value) noexcept
{
    return value <= static_cast<std::uint8_t>(ParticleType::Asteroid);
}
