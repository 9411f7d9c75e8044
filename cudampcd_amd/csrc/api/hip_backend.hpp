// HipBackend: the MI355X backend behind the mpcd::Backend interface.
//
// A thin C++ wrapper over the C-ABI of libmpcd_hip.so (include/mpcd_hip.h):
// every call forwards to one mpcd_hip_* entry point and turns a non-zero
// status into std::runtime_error(mpcd_hip_last_error()), the error convention
// of the reference backends (cuda_backend/gpu_error_check.hpp:10-39).
#pragma once

#include <cstdint>
#include <vector>

#include "backend.hpp"
#include "mpcd_hip.h"

namespace mpcd::hip {

class HipBackend : public Backend {
   public:
    explicit HipBackend(SimulationParameters const& params);
    ~HipBackend() override;
    HipBackend(HipBackend const&) = delete;
    HipBackend& operator=(HipBackend const&) = delete;

    void step(int n_steps) override;
    void stepAndAccumulateSample(int n_steps) override;
    void getSampleMean(std::vector<float>& mean_density, std::vector<float>& mean_velocity) override;
    size_t getNParticles() override;
    void getParticlePositions(std::vector<float>& positions) override;
    void getParticleVelocities(std::vector<float>& velocities) override;
    void setParticlePositions(std::vector<float> const& positions) override;
    void setParticleVelocities(std::vector<float> const& velocities) override;
    void setParticles(std::vector<float> const& positions, std::vector<float> const& velocities) override;

    // beyond the reference interface
    void getSampleMeanF64(std::vector<double>& mean_density, std::vector<double>& mean_velocity);
    uint64_t seed() const { return seed_; }
    mpcd_hip_t* handle() const { return h_; }

   private:
    void check(int status) const;
    mpcd_hip_t* h_ = nullptr;
    uint64_t seed_ = 0;
    size_t n_cells_ = 0;
};

}  // namespace mpcd::hip
