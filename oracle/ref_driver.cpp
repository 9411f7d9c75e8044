// Driver for the reference CPU backend, compiled in place from the reference
// sources (test infrastructure: builds oracle/_ref/libref.so).
//
// It constructs the reference's own mpcd::cpu::CPUBackend (cpu_backend.cpp:8-21),
// reseeds its Xoshiro128Plus stream (random.hpp:66-74) so runs are reproducible,
// and exposes step / stepAndAccumulateSample / getSampleMean plus the
// translation half-step and the cell table, through a C interface for ctypes.
// Nothing here re-implements reference behaviour: every state change is made
// by the reference's code.

#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iostream>
#include <memory>
#include <string>
#include <vector>

#include "cpu_backend/cpu_backend.hpp"
#include "cpu_backend/simulation_functions.hpp"

// The backend keeps its stream and tables private; the driver reaches them
// through member pointers formed in explicit template instantiations, where
// the language does not apply access checks.
namespace {
template <typename Tag, typename Tag::type M>
struct Expose {
    friend typename Tag::type member(Tag) { return M; }
};
#define EXPOSE(NAME, TYPE, MEMBER)                                   \
    struct NAME {                                                    \
        using type = TYPE;                                           \
        friend type member(NAME);                                    \
    };                                                               \
    template struct Expose<NAME, &MEMBER>;
using B = mpcd::cpu::CPUBackend;
EXPOSE(TagRandom, mpcd::Xoshiro128Plus B::*, B::random)
EXPOSE(TagParticles, std::vector<mpcd::Particle> B::*, B::particles)
EXPOSE(TagCells, mpcd::cpu::VolumeContainer<mpcd::MPCCell> B::*, B::mpc_cells)
EXPOSE(TagCounter, std::vector<uint32_t> B::*, B::uniform_counter)
EXPOSE(TagList, std::vector<uint32_t> B::*, B::uniform_list)
EXPOSE(TagShift, mpcd::Vector B::*, B::grid_shift)
EXPOSE(TagTranslate, void (B::*)(), B::translationStep)
EXPOSE(TagCollide, void (B::*)(), B::collisionStep)
#undef EXPOSE
}  // namespace

extern "C" {

struct ref_params {
    float volume[3];       // cells per axis, wall layers included
    int periodic[3];
    unsigned n;
    unsigned long long N;
    float delta_t;
    float drag;
    float thermal_velocity;
    int algorithm;         // 0 srd, 1 extended
    unsigned long long seed_a, seed_b;
    int experiment;        // 0 standard, 1 channel
};

void* ref_create(const ref_params* q) {
    mpcd::SimulationParameters p;
    p.experiment = q->experiment == 1 ? mpcd::ExperimentType::channel : mpcd::ExperimentType::standard;
    p.algorithm = q->algorithm == 0 ? mpcd::MPCDAlgorithm::srd : mpcd::MPCDAlgorithm::extended;
    for (int i = 0; i < 3; ++i) {
        p.volume_size[i] = q->volume[i];
        p.periodicity[i] = q->periodic[i];
    }
    p.n = q->n;
    p.N = static_cast<unsigned>(q->N);
    p.delta_t = q->delta_t;
    p.drag = q->drag;
    p.thermal_velocity = q->thermal_velocity;
    p.temperature = q->thermal_velocity * q->thermal_velocity;
    auto* b = new mpcd::cpu::CPUBackend(p);
    // replace the /dev/urandom-seeded start with a seeded one
    (b->*member(TagRandom{})).seed(q->seed_a, q->seed_b);
    mpcd::cpu::distribute_particles(b->*member(TagParticles{}), b->*member(TagCells{}), p, b->*member(TagRandom{}));
    return b;
}

void ref_destroy(void* h) { delete static_cast<mpcd::cpu::CPUBackend*>(h); }

void ref_step(void* h, int n) { static_cast<mpcd::cpu::CPUBackend*>(h)->step(n); }
void ref_step_and_sample(void* h, int n) { static_cast<mpcd::cpu::CPUBackend*>(h)->stepAndAccumulateSample(n); }

void ref_get_sample_mean(void* h, float* dens, float* vel) {
    std::vector<float> d, v;
    static_cast<mpcd::cpu::CPUBackend*>(h)->getSampleMean(d, v);
    std::memcpy(dens, d.data(), d.size() * sizeof(float));
    std::memcpy(vel, v.data(), v.size() * sizeof(float));
}

void ref_get_particles(void* h, float* pos, float* vel) {
    std::vector<float> a, b;
    auto* be = static_cast<mpcd::cpu::CPUBackend*>(h);
    be->getParticlePositions(a);
    be->getParticleVelocities(b);
    if (pos) std::memcpy(pos, a.data(), a.size() * sizeof(float));
    if (vel) std::memcpy(vel, b.data(), b.size() * sizeof(float));
}

// the reference's set* are TODO (cpu_backend.cpp:146-156); write the particle array directly
void ref_set_particles(void* h, const float* pos, const float* vel, unsigned long long n) {
    auto* be = static_cast<mpcd::cpu::CPUBackend*>(h);
    (be->*member(TagParticles{})).resize(n);
    for (size_t i = 0; i < n; ++i) {
        (be->*member(TagParticles{}))[i].position = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
        (be->*member(TagParticles{}))[i].velocity = {vel[3 * i], vel[3 * i + 1], vel[3 * i + 2]};
        (be->*member(TagParticles{}))[i].cell_idx = (be->*member(TagCells{})).getIndex((be->*member(TagParticles{}))[i].position);
    }
}

void ref_reseed(void* h, unsigned long long a, unsigned long long b) {
    (static_cast<mpcd::cpu::CPUBackend*>(h)->*member(TagRandom{})).seed(a, b);
}

// translation half-step only (cpu_backend.cpp:23-33)
void ref_translate(void* h, float* shift_out) {
    auto* be = static_cast<mpcd::cpu::CPUBackend*>(h);
    (be->*member(TagTranslate{}))();
    if (shift_out) {
        shift_out[0] = (be->*member(TagShift{})).x;
        shift_out[1] = (be->*member(TagShift{})).y;
        shift_out[2] = (be->*member(TagShift{})).z;
    }
}
void ref_collide(void* h) { (static_cast<mpcd::cpu::CPUBackend*>(h)->*member(TagCollide{}))(); }

unsigned long long ref_get_cells(void* h, unsigned* cell_of_particle, unsigned* counter, unsigned* flat_list) {
    auto* be = static_cast<mpcd::cpu::CPUBackend*>(h);
    size_t C = (be->*member(TagCells{})).size();
    size_t cap = (be->*member(TagList{})).size() / C;
    for (size_t i = 0; i < (be->*member(TagParticles{})).size(); ++i) cell_of_particle[i] = (be->*member(TagParticles{}))[i].cell_idx;
    std::memcpy(counter, (be->*member(TagCounter{})).data(), C * sizeof(unsigned));
    size_t k = 0;
    for (size_t c = 0; c < C; ++c) {
        size_t m = std::min<size_t>(cap, (be->*member(TagCounter{}))[c]);
        for (size_t i = 0; i < m; ++i) flat_list[k++] = (be->*member(TagList{}))[c + i * C];
    }
    return k;
}

unsigned long long ref_n_particles(void* h) { return static_cast<mpcd::cpu::CPUBackend*>(h)->getNParticles(); }

}  // extern "C"
