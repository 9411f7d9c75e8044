#include <mpcd/api/simulation_handle.hpp>

#include <stdexcept>

#include "common/backend.hpp"
#include "hip_backend/hip_backend.hpp"

namespace mpcd {

// backend_factory.hpp:15-27 of the reference, with the "hip" backend
std::unique_ptr<Backend> create_backend(SimulationParameters const& params, std::string const& backend_type) {
    if (backend_type == "hip") return std::make_unique<hip::HipBackend>(params);
    throw std::runtime_error("Unsupported backend type");
}

namespace api {

SimulationHandle::SimulationHandle(SimulationParameters const& params, std::string backend_type)
    : backend_(create_backend(params, backend_type)) {}
SimulationHandle::~SimulationHandle() = default;
SimulationHandle::SimulationHandle(SimulationHandle&&) noexcept = default;
SimulationHandle& SimulationHandle::operator=(SimulationHandle&&) noexcept = default;

SimulationParameters const& SimulationHandle::getParameters() const { return backend_->getParameters(); }
void SimulationHandle::step(int n_steps) { backend_->step(n_steps); }
void SimulationHandle::stepAndAccumulateSample(int n_steps) { backend_->stepAndAccumulateSample(n_steps); }
void SimulationHandle::getSampleMean(std::vector<float>& d, std::vector<float>& v) { backend_->getSampleMean(d, v); }
void SimulationHandle::getParticlePositions(std::vector<float>& p) const { backend_->getParticlePositions(p); }
void SimulationHandle::getParticleVelocities(std::vector<float>& v) const { backend_->getParticleVelocities(v); }
void SimulationHandle::setParticlePositions(std::vector<float> const& p) { backend_->setParticlePositions(p); }
void SimulationHandle::setParticleVelocities(std::vector<float> const& v) { backend_->setParticleVelocities(v); }
void SimulationHandle::setParticles(std::vector<float> const& p, std::vector<float> const& v) {
    backend_->setParticles(p, v);
}

}  // namespace api
}  // namespace mpcd
