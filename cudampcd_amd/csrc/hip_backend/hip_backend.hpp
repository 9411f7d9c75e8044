// HipBackend: the MI355X backend behind the mpcd::Backend interface.
//
// A thin C++ wrapper over the C-ABI of libmpcd_hip.so (include/mpcd_hip.h):
// every call forwards to one mpcd_hip_* entry point and turns a non-zero
// status into std::runtime_error(mpcd_hip_last_error()), the error convention
// of the reference backends (cuda_backend/gpu_error_check.hpp:10-39).
#pragma once

#include <cstdint>
#include <vector>

#include <mpcd/api/simulation_parameters.hpp>

#include "common/backend.hpp"
#include "mpcd_hip.h"

namespace mpcd::hip {

// the C-ABI parameter block of one system; `seed` keys its random stream
mpcd_params_t to_c_params(SimulationParameters const& params, uint64_t seed);
// a fresh seed from std::random_device (the reference seeds from /dev/urandom, random.hpp:39-46)
uint64_t fresh_seed();

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

// An ensemble of independent systems advanced together on one device
// (mpcd_hip_create_ensemble): every kernel of a step covers all systems in one
// launch.  No reference counterpart; system i reproduces bit for bit a
// HipBackend created from the same parameters and seed.
class HipEnsemble {
   public:
    explicit HipEnsemble(std::vector<SimulationParameters> const& params);
    ~HipEnsemble();
    HipEnsemble(HipEnsemble const&) = delete;
    HipEnsemble& operator=(HipEnsemble const&) = delete;

    void step(int n_steps);
    void stepAndAccumulateSample(int n_steps);
    // FP32 / FP64 mean fields of system i (the current average is kept; resetSampling ends it)
    void getSampleMean(uint32_t i, std::vector<float>& density, std::vector<float>& velocity);
    void getSampleMeanF64(uint32_t i, std::vector<double>& density, std::vector<double>& velocity);
    void resetSampling();
    void getParticles(uint32_t i, std::vector<float>* positions, std::vector<float>* velocities);
    uint32_t nSystems() const { return static_cast<uint32_t>(params_.size()); }
    size_t nParticles(uint32_t i) const;
    SimulationParameters const& parameters(uint32_t i) const;
    uint64_t seed(uint32_t i) const;
    void useGraph(bool enable);
    // tries the compiled tile shapes on real steps and keeps the fastest
    // (results do not depend on the tile shape); returns cells per tile of
    // system 0, LDS capacity, milliseconds per step
    void autotune(int steps_per_trial, uint32_t& cells_per_tile, uint32_t& lds_capacity, double& ms_per_step);
    mpcd_hip_t* handle() const { return h_; }

   private:
    void check(int status) const;
    mpcd_hip_t* h_ = nullptr;
    std::vector<SimulationParameters> params_;
    std::vector<uint64_t> seeds_;
};

}  // namespace mpcd::hip
