#include "hip_backend.hpp"

#include <random>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>

namespace mpcd::hip {

uint64_t fresh_seed() {
    std::random_device rd;
    return (static_cast<uint64_t>(rd()) << 32) ^ rd();
}

// The reference's SimulationParameters (include/mpcd/api/simulation_parameters.hpp:23-55)
// has no rotation_angle, seed / seed_set or system_id: it fixes the angle to 120
// degrees (simulation_functions.cpp:118-119), seeds from /dev/urandom
// (random.hpp:39-46) and runs one system.  These accessors read the fields
// where the parameter struct has them and take the reference's choices where
// it does not, so this file compiles unchanged against either struct.
namespace {
template <typename P, typename = void>
struct has_rotation_angle : std::false_type {};
template <typename P>
struct has_rotation_angle<P, std::void_t<decltype(std::declval<P const&>().rotation_angle)>> : std::true_type {};
template <typename P, typename = void>
struct has_seed : std::false_type {};
template <typename P>
struct has_seed<P, std::void_t<decltype(std::declval<P const&>().seed), decltype(std::declval<P const&>().seed_set)>>
    : std::true_type {};
template <typename P, typename = void>
struct has_system_id : std::false_type {};
template <typename P>
struct has_system_id<P, std::void_t<decltype(std::declval<P const&>().system_id)>> : std::true_type {};

template <typename P>
float rotation_angle_of(P const& p) {
    if constexpr (has_rotation_angle<P>::value) return p.rotation_angle;
    else return 120.0f;
}
template <typename P>
uint64_t seed_of(P const& p) {
    if constexpr (has_seed<P>::value) return p.seed_set ? static_cast<uint64_t>(p.seed) : fresh_seed();
    else return fresh_seed();
}
template <typename P>
uint32_t system_id_of(P const& p) {
    if constexpr (has_system_id<P>::value) return static_cast<uint32_t>(p.system_id);
    else return 0u;
}
}  // namespace

mpcd_params_t to_c_params(SimulationParameters const& params, uint64_t seed) {
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
    p.rotation_angle = rotation_angle_of(params);
    p.seed = seed;
    p.system_id = system_id_of(params);
    return p;
}

HipBackend::HipBackend(SimulationParameters const& params) : Backend(params) {
    seed_ = seed_of(params);
    mpcd_params_t const p = to_c_params(params, seed_);
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

HipEnsemble::HipEnsemble(std::vector<SimulationParameters> const& params) : params_(params) {
    if (params.empty()) throw std::runtime_error("an ensemble needs at least one system");
    std::vector<mpcd_params_t> c;
    for (auto const& p : params) {
        seeds_.push_back(seed_of(p));
        c.push_back(to_c_params(p, seeds_.back()));
    }
    if (mpcd_hip_create_ensemble(c.data(), static_cast<uint32_t>(c.size()), &h_) != 0)
        throw std::runtime_error(mpcd_hip_last_error(nullptr));
}

HipEnsemble::~HipEnsemble() { mpcd_hip_destroy(h_); }

void HipEnsemble::check(int status) const {
    if (status != 0) throw std::runtime_error(mpcd_hip_last_error(h_));
}

void HipEnsemble::step(int n_steps) { check(mpcd_hip_step(h_, n_steps)); }

void HipEnsemble::stepAndAccumulateSample(int n_steps) { check(mpcd_hip_step_and_sample(h_, n_steps)); }

static size_t system_cells(mpcd_hip_t* h, uint32_t i) {
    uint64_t c = 0;
    if (mpcd_hip_system_info(h, i, nullptr, nullptr, nullptr, &c) != 0) throw std::runtime_error("system index out of range");
    return c;
}

void HipEnsemble::getSampleMean(uint32_t i, std::vector<float>& density, std::vector<float>& velocity) {
    size_t const c = system_cells(h_, i);
    density.resize(c);
    velocity.resize(3 * c);
    check(mpcd_hip_system_sample_mean(h_, i, density.data(), velocity.data(), c));
}

void HipEnsemble::getSampleMeanF64(uint32_t i, std::vector<double>& density, std::vector<double>& velocity) {
    size_t const c = system_cells(h_, i);
    density.resize(c);
    velocity.resize(3 * c);
    check(mpcd_hip_system_sample_mean_f64(h_, i, density.data(), velocity.data(), c));
}

void HipEnsemble::resetSampling() { check(mpcd_hip_reset_sampling(h_)); }

size_t HipEnsemble::nParticles(uint32_t i) const {
    uint64_t n = 0;
    if (mpcd_hip_system_info(h_, i, nullptr, &n, nullptr, nullptr) != 0) throw std::runtime_error("system index out of range");
    return n;
}

SimulationParameters const& HipEnsemble::parameters(uint32_t i) const {
    if (i >= params_.size()) throw std::runtime_error("system index out of range");
    return params_[i];
}

uint64_t HipEnsemble::seed(uint32_t i) const {
    if (i >= seeds_.size()) throw std::runtime_error("system index out of range");
    return seeds_[i];
}

void HipEnsemble::getParticles(uint32_t i, std::vector<float>* positions, std::vector<float>* velocities) {
    size_t const n = nParticles(i);
    if (positions) positions->resize(3 * n);
    if (velocities) velocities->resize(3 * n);
    check(mpcd_hip_system_get_particles(h_, i, positions ? positions->data() : nullptr,
                                        velocities ? velocities->data() : nullptr, n));
}

void HipEnsemble::useGraph(bool enable) { check(mpcd_hip_use_graph(h_, enable ? 1 : 0)); }

void HipEnsemble::autotune(int steps_per_trial, uint32_t& cells_per_tile, uint32_t& lds_capacity, double& ms_per_step) {
    check(mpcd_hip_autotune(h_, steps_per_trial, &cells_per_tile, &lds_capacity, &ms_per_step));
}

}  // namespace mpcd::hip
