// streaming/device_capture.hpp -- DeviceFrameWriter
// (src/streaming/device_capture.hpp:10-36, device_capture.cu:155-330) over
// the C ABI.  Samples at most `maximumParticles` particles (0 = all) of a
// device particle view and writes them as version-3 quantized typed frames.
// Sample j of S reads particle j * N / S.  Types are sampled and validated
// once, at construction.  Each write() runs the bounds and quantize kernels
// on the device (nbh_frame_capture), copies 6 S bytes to the host and
// publishes the frame through the spool.
#pragma once

#include <algorithm>
#include <cstdint>
#include <cstring>
#include <filesystem>
#include <stdexcept>
#include <vector>

#include "../device_buffer.hpp"
#include "../nbody_amd_shim.hpp"
#include "../particle.hpp"
#include "frame.hpp"

namespace nbody::streaming {

class DeviceFrameWriter {
public:
    DeviceFrameWriter(ConstDeviceParticlesView particles, std::size_t maximumParticles)
        : particles_(particles),
          sample_(maximumParticles == 0 ? particles.count
                                        : std::min(particles.count, maximumParticles))
    {
        if (particles_.count == 0 || particles_.x == nullptr || particles_.y == nullptr ||
            particles_.z == nullptr || particles_.type == nullptr)
            throw std::invalid_argument(
                "Cannot capture an empty or incomplete device particle view");
        positions_.allocate(sample_ * kPositionComponents);
        hostPositions_.resize(sample_ * kPositionComponents);
        DeviceBuffer<std::uint8_t> types(sample_);
        nbody_amd::check(nbh_frame_sample_types(nbody_amd::default_context(), particles_.type,
                                                particles_.count, sample_, types.data()));
        sampledTypes_.resize(sample_);
        nbody_amd::check_hip(hipMemcpy(sampledTypes_.data(), types.data(), sample_,
                                       hipMemcpyDeviceToHost),
                             "Could not copy sampled particle types");
    }

    DeviceFrameWriter(const DeviceFrameWriter&) = delete;
    DeviceFrameWriter& operator=(const DeviceFrameWriter&) = delete;

    [[nodiscard]] std::size_t sourceParticleCount() const noexcept { return particles_.count; }
    [[nodiscard]] std::size_t sampleParticleCount() const noexcept { return sample_; }

    std::filesystem::path write(const FrameSpool& spool, std::uint64_t sequence,
                                std::uint64_t simulationStep, double simulationTime) const
    {
        double b[6];
        nbody_amd::check(nbh_frame_capture(nbody_amd::default_context(), particles_.x,
                                           particles_.y, particles_.z, particles_.count,
                                           sample_, positions_.data(), b));
        nbody_amd::check_hip(hipMemcpy(hostPositions_.data(), positions_.data(),
                                       hostPositions_.size() * sizeof(std::uint16_t),
                                       hipMemcpyDeviceToHost),
                             "Could not copy quantized samples to the frame writer");
        PositionBounds bounds;
        for (std::size_t c = 0; c < kPositionComponents; ++c) {
            bounds.minimum[c] = b[c];
            bounds.maximum[c] = b[kPositionComponents + c];
        }
        const QuantizedPositionChunkReader reader =
            [this](std::size_t offset, std::size_t count, std::uint16_t* xyz) {
                std::memcpy(xyz, hostPositions_.data() + offset * kPositionComponents,
                            count * kPositionComponents * sizeof(std::uint16_t));
            };
        const ParticleTypeChunkReader typeReader =
            [this](std::size_t offset, std::size_t count, std::uint8_t* types) {
                std::copy_n(sampledTypes_.data() + offset, count, types);
            };
        return spool.writeQuantizedPositions(sequence, simulationStep, simulationTime, sample_,
                                             bounds, reader, particles_.count, typeReader);
    }

private:
    ConstDeviceParticlesView particles_{};
    std::size_t sample_ = 0;
    mutable DeviceBuffer<std::uint16_t> positions_;
    mutable std::vector<std::uint16_t> hostPositions_;
    std::vector<std::uint8_t> sampledTypes_;
};

} // namespace nbody::streaming
