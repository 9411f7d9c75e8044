// SimulationHandle: owner of one backend, the bridge the bindings talk to.
// Same interface as mpcd::api::SimulationHandle
// (reference include/mpcd/api/simulation_handle.hpp:14-45,
// src/libmpcd/api/simulator_handle.cpp:9-48).
#pragma once

#include <memory>
#include <string>
#include <vector>

#include <mpcd/api/simulation_parameters.hpp>

namespace mpcd {
struct Backend;

namespace api {

class SimulationHandle {
   public:
    SimulationHandle(SimulationParameters const&, std::string backend_type);
    ~SimulationHandle();

    SimulationHandle(SimulationHandle&&) noexcept;
    SimulationHandle& operator=(SimulationHandle&&) noexcept;
    SimulationHandle(const SimulationHandle&) = delete;
    SimulationHandle& operator=(const SimulationHandle&) = delete;

    SimulationParameters const& getParameters() const;

    void step(int n_steps);
    void stepAndAccumulateSample(int n_steps);
    void getSampleMean(std::vector<float>& mean_density, std::vector<float>& mean_velocity);

    void getParticlePositions(std::vector<float>& positions) const;
    void getParticleVelocities(std::vector<float>& velocities) const;
    void setParticlePositions(std::vector<float> const& positions);
    void setParticleVelocities(std::vector<float> const& velocities);
    void setParticles(std::vector<float> const& positions, std::vector<float> const& velocities);

    Backend& backend() { return *backend_; }

   private:
    std::unique_ptr<Backend> backend_;
};

}  // namespace api
}  // namespace mpcd
