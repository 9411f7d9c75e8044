#include "hip_backend.hpp"

#include <random>
#include <stdexcept>
#include <string>

namespace mpcd::hip {

namespace {
uint64_t fresh_seed() {
    std::random_device rd;
    return (static_cast<uint64_t>(rd()) << 32) ^ rd();
}
}  // namespace

HipBackend::HipBackend(SimulationParameters const& params) : Backend(params) {
    mpcd_params_t p{};
    p.experiment = params.experiment == ExperimentType::standard ? 0 : 1;
    p.algorithm = params.algorithm == MPCDAlgorithm::srd ? 0 : 1;
    p.collision_kernel = static_cast<int>(params.collision_kernel);
    for (int i = 0; i < 3; ++i) {
        p.periodicity[i] = params.periodicity[i];
        p.volume_size[i] = params.volume_size[i];
    }
    p.N = params.N;
    p.n = params.n;
    p.device_id = params.device_id;
    p.delta_t = params.delta_t;
    p.temperature = params.temperature;
    p.thermal_velocity = params.thermal_velocity;
    p.drag = params.drag;
    p.rotation_angle = params.rotation_angle;
    seed_ = params.seed_set ? params.seed : fresh_seed();
    p.seed = seed_;
    p.system_id = params.system_id;
    if (mpcd_hip_create(&p, &h_) != 0) throw std::runtime_error(mpcd_hip_last_error(nullptr));
    uint64_t c = 0;
    mpcd_hip_n_cells(h_, &c);
    n_cells_ = c;
}

HipBackend::~HipBackend() { mpcd_hip_destroy(h_); }

void HipBackend::check(int status) const {
    if (status != 0) throw std::runtime_error(mpcd_hip_last_error(h_));
}

void HipBackend::step(int n_steps) { check(mpcd_hip_step(h_, n_steps)); }

void HipBackend::stepAndAccumulateSample(int n_steps) { check(mpcd_hip_step_and_sample(h_, n_steps)); }

void HipBackend::getSampleMean(std::vector<float>& mean_density, std::vector<float>& mean_velocity) {
    mean_density.resize(n_cells_);
    mean_velocity.resize(3 * n_cells_);
    check(mpcd_hip_get_sample_mean(h_, mean_density.data(), mean_velocity.data(), n_cells_));
}

void HipBackend::getSampleMeanF64(std::vector<double>& mean_density, std::vector<double>& mean_velocity) {
    mean_density.resize(n_cells_);
    mean_velocity.resize(3 * n_cells_);
    check(mpcd_hip_get_sample_mean_f64(h_, mean_density.data(), mean_velocity.data(), n_cells_));
}

size_t HipBackend::getNParticles() {
    uint64_t n = 0;
    check(mpcd_hip_n_particles(h_, &n));
    return n;
}

void HipBackend::getParticlePositions(std::vector<float>& positions) {
    size_t const n = getNParticles();
    positions.resize(3 * n);
    check(mpcd_hip_get_positions(h_, positions.data(), n));
}

void HipBackend::getParticleVelocities(std::vector<float>& velocities) {
    size_t const n = getNParticles();
    velocities.resize(3 * n);
    check(mpcd_hip_get_velocities(h_, velocities.data(), n));
}

static size_t triples(std::vector<float> const& v) {
    if (v.size() % 3 != 0) throw std::runtime_error("particle array length must be a multiple of 3");
    return v.size() / 3;
}

void HipBackend::setParticlePositions(std::vector<float> const& positions) {
    check(mpcd_hip_set_positions(h_, positions.data(), triples(positions)));
}

void HipBackend::setParticleVelocities(std::vector<float> const& velocities) {
    check(mpcd_hip_set_velocities(h_, velocities.data(), triples(velocities)));
}

void HipBackend::setParticles(std::vector<float> const& positions, std::vector<float> const& velocities) {
    if (positions.size() != velocities.size()) throw std::runtime_error("positions and velocities differ in length");
    check(mpcd_hip_set_particles(h_, positions.data(), velocities.data(), triples(positions)));
}

}  // namespace mpcd::hip
