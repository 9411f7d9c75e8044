/* C-ABI of the MI355X (gfx950) MPCD backend: libmpcd_hip.so.
 *
 * Every entry point returns 0 on success and a non-zero status on failure;
 * mpcd_hip_last_error() then holds the message.  No entry point takes or
 * returns C++ or torch types: plain pointers, sizes and a POD parameter block.
 * Host arrays are caller-allocated.  Particle arrays are interleaved xyz
 * (3 floats per particle), cell arrays follow the reference cell index
 * c = x + X*(y + Y*z) (volume_container.hpp:69-75).
 *
 * Each entry point names the reference interface it replaces.  The reference
 * reaches its backends through the abstract mpcd::Backend
 * (src/libmpcd/common/backend.hpp:11-31) selected by create_backend
 * (src/libmpcd/common/backend_factory.hpp:15-27); HipBackend (C++) wraps this
 * C-ABI behind that same interface.
 */
#ifndef MPCD_HIP_H
#define MPCD_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Mirrors mpcd::SimulationParameters (include/mpcd/api/simulation_parameters.hpp:23-55),
 * numeric fields only, plus the options the reference fixes in code. */
typedef struct mpcd_params_t {
    int experiment;          /* 0 standard, 1 channel: cylinder about x (simulation_functions.cpp:11-24, 75-82) */
    int algorithm;           /* 0 srd, 1 extended */
    int collision_kernel;    /* accepted for compatibility; this backend always runs its fused kernels */
    int periodicity[3];
    float volume_size[3];    /* cells per axis, wall layers included (bindings.cpp:108-123) */
    uint64_t N;              /* particles */
    uint32_t n;              /* particles per cell */
    uint32_t device_id;
    float delta_t;
    float temperature;
    float thermal_velocity;
    float drag;
    float rotation_angle;    /* SRD rotation angle in degrees; the reference fixes 120 (simulation_functions.cpp:118-119) */
    uint64_t seed;           /* Philox key of the counter-based random stream */
    uint32_t system_id;      /* ensemble member id folded into the random stream */
    uint32_t reserved;
} mpcd_params_t;

typedef struct mpcd_hip_s mpcd_hip_t;

/* Message of the last failure on this handle (or of the last failed create when h is NULL). */
const char* mpcd_hip_last_error(const mpcd_hip_t* h);

/* Replaces CudaBackend::CudaBackend -> BackendContext::BackendContext
 * (cuda_backend.cu:7-8, backend_context.cu:17-127) and CPUBackend::CPUBackend
 * (cpu_backend.cpp:8-21): allocates device state and distributes particles
 * uniformly in the fluid region with Maxwell-Boltzmann velocities
 * (simulation_functions.cpp:7-33), seeded from params->seed.  Returns nonzero,
 * with the reason in mpcd_hip_last_error(NULL), for parameters the step does
 * not support: n or N zero, extents not whole positive numbers or above 2^21,
 * walls on x or y, fewer than 3 layers between z walls, 2^32 or more
 * particles or cells, an unknown algorithm or experiment, no device. */
int mpcd_hip_create(const mpcd_params_t* params, mpcd_hip_t** out);
/* Replaces the backend destructor (cuda_backend_impl.cu:19-21). */
int mpcd_hip_destroy(mpcd_hip_t* h);

/* Replaces Backend::step (backend.hpp:12; cuda_backend_impl.cu:111-131,
 * cpu_backend.cpp:65-70): n_steps of grid shift + streaming + cell sort +
 * collision.  Returns after the steps finished on the device. */
int mpcd_hip_step(mpcd_hip_t* h, int n_steps);
/* Replaces Backend::stepAndAccumulateSample (backend.hpp:14;
 * cuda_backend_impl.cu:182-204, cpu_backend.cpp:72-109). Accumulates in FP64. */
int mpcd_hip_step_and_sample(mpcd_hip_t* h, int n_steps);
/* One SRD step of a single-system handle with the caller's draws in place of
 * its own: the parity hook that injects another implementation's random
 * numbers (no reference counterpart; the reference draws inside its step,
 * simulation_functions.cpp:108-202), so that a step can be compared with that
 * implementation's particle by particle.  shift[3]: the grid shift; axes[3 C]:
 * the rotation axis of every cell (x, y, z; reference cell order c = x +
 * X (y + Y z); read for the cells that collide); with walls in z, for the
 * cells w = (upper ? X Y : 0) + x + X y of this shift's two ghost layers
 * (simulation_functions.cpp:137-138: the lower at z = (shift_z > 0 ? 1 : 0),
 * the upper at z = Z - (shift_z > 0 ? 1 : 2)), ghost_counts[2 X Y] (each at
 * most n) and ghost_velocities[2 X Y n 3] (velocities with the thermal
 * velocity applied; the first count of a cell's n are used), NULL without
 * walls.  n_cells must be C, n_wall_cells 2 X Y with walls and 0 without,
 * n_ghost_values the length of ghost_velocities (2 X Y n 3, or 0).
 * Host memory.  The step index advances by one; later steps draw their own. */
int mpcd_hip_step_injected(mpcd_hip_t* h, const float* shift, const float* axes, size_t n_cells,
                           const uint32_t* ghost_counts, const float* ghost_velocities, size_t n_wall_cells,
                           size_t n_ghost_values);
/* The same for one extended-rule step (simulation_functions.cpp:207-406):
 * shift[3]; cell_draws[3 C] per cell (reference order): the direction index
 * 0-12 and sign 0/1 of simulation_functions.cpp:272-297 and the bits of the
 * gate uniform; thermal[C 4 n]: the thermal Gaussians of a cell's members
 * (table particles in slot order, then ghosts; read for the cells that
 * collide); with walls, ghost_counts[2 X Y] and ghosts[2 X Y n 6] (position
 * x, y, z and velocity with the thermal velocity applied, n slots per
 * wall-layer cell w as above), NULL without walls; n_thermal (C 4 n) and
 * n_ghost_values (2 X Y n 6, or 0) are the two arrays' lengths.  The direction and gate
 * draws of the collision and its thermal draws come from these arrays
 * (copied to the device for the step), the rest of the step as usual. */
int mpcd_hip_step_injected_extended(mpcd_hip_t* h, const float* shift, const uint32_t* cell_draws,
                                    const float* thermal, size_t n_cells, size_t n_thermal,
                                    const uint32_t* ghost_counts, const float* ghosts, size_t n_wall_cells,
                                    size_t n_ghost_values);
/* Replaces Backend::getSampleMean (backend.hpp:15; cpu_backend.cpp:111-122):
 * density[n_cells], velocity[3*n_cells], scaled by 1/sample_counter as the
 * reference does, and ends the current average. */
int mpcd_hip_get_sample_mean(mpcd_hip_t* h, float* density, float* velocity, size_t n_cells);
/* Same fields in FP64 (no reference counterpart; the FP64 accumulators themselves). */
int mpcd_hip_get_sample_mean_f64(mpcd_hip_t* h, double* density, double* velocity, size_t n_cells);

/* Replaces Backend::getNParticles (backend.hpp:17). */
int mpcd_hip_n_particles(const mpcd_hip_t* h, uint64_t* out);
/* Number of cells X*Y*Z (volume_container size). */
int mpcd_hip_n_cells(const mpcd_hip_t* h, uint64_t* out);
/* Replace Backend::getParticlePositions / getParticleVelocities (backend.hpp:18-19). */
int mpcd_hip_get_positions(mpcd_hip_t* h, float* out, size_t n_particles);
int mpcd_hip_get_velocities(mpcd_hip_t* h, float* out, size_t n_particles);
/* Replace Backend::setParticlePositions / setParticleVelocities / setParticles
 * (backend.hpp:20-22).  n_particles must equal the particle count. */
int mpcd_hip_set_positions(mpcd_hip_t* h, const float* pos, size_t n_particles);
int mpcd_hip_set_velocities(mpcd_hip_t* h, const float* vel, size_t n_particles);
int mpcd_hip_set_particles(mpcd_hip_t* h, const float* pos, const float* vel, size_t n_particles);

/* ---- parity and measurement hooks (no reference counterpart) ---- */

/* Step index that keys the counter-based random stream of the next step. */
int mpcd_hip_get_step_index(const mpcd_hip_t* h, uint64_t* out);
int mpcd_hip_set_step_index(mpcd_hip_t* h, uint64_t step);
/* The grid shift the given step draws (the reference draws it on the host, cpu_backend.cpp:23-26). */
int mpcd_hip_grid_shift(const mpcd_hip_t* h, uint64_t step, float* out3);
/* Streaming half-step of the next step without collision, as translate_particles
 * (simulation_functions.cpp:38-103) leaves it: shifted cell index per particle,
 * per-cell counts (uniform_counter), and the stable cell-ordered permutation
 * (uniform_list flattened cell by cell, all entries of each cell).  Leaves the
 * state unchanged; the next step reuses the sort. */
int mpcd_hip_peek_sort(mpcd_hip_t* h, uint32_t* cell_of_particle, uint32_t* counts, uint32_t* permutation);
/* When enabled, each step records for every output slot the input slot it came
 * from; mpcd_hip_last_permutation returns the record of the last step. */
int mpcd_hip_record_permutation(mpcd_hip_t* h, int enable);
int mpcd_hip_last_permutation(mpcd_hip_t* h, uint32_t* out, size_t n_particles);
/* Record HIP events around every collide launch of later step() calls. */
int mpcd_hip_enable_timing(mpcd_hip_t* h, int enable);
/* Device-side timing of the last step() call: milliseconds of the collision
 * kernels and of the whole step sequence, from HIP events on the handle's stream. */
int mpcd_hip_last_timing(const mpcd_hip_t* h, double* collide_ms, double* total_ms, int* launches);
/* Kernel launch shape: cells per tile (at most 24 for the SRD capacity 384,
 * 32 for SRD 448, 48 for extended 512, else 64) and LDS particle capacity of one
 * batch (rounded up to 256,
 * 384, 448, 512, 640, 1152 or 2048; 256 runs one-wave workgroups, 384 to 640
 * two-wave workgroups, larger ones four waves; 256 to 448 are SRD only, and
 * the extended rule takes 1152 for 640; 0 = keep).  Cells holding more than
 * one batch are processed in several batches.  Results do not depend on it. */
int mpcd_hip_set_tiling(mpcd_hip_t* h, uint32_t cells_per_tile, uint32_t lds_capacity);
int mpcd_hip_get_tiling(const mpcd_hip_t* h, uint32_t* cells_per_tile, uint32_t* lds_capacity);
/* The tile shape of one system (mpcd_hip_get_tiling gives system 0's).  Systems
 * of an ensemble may run different shapes: each system keeps its own, and
 * the systems sharing an LDS capacity take one collide launch. */
int mpcd_hip_system_tiling(const mpcd_hip_t* h, uint32_t system, uint32_t* cells_per_tile, uint32_t* lds_capacity);
/* ---- pathwise adjoint (no reference counterpart) ----
 *
 * Loss of `steps` SRD steps of a single-system handle from its current state,
 * with drag g, rotation angle (degrees) and an x body force f per reference
 * cell (v_x += g dt + f dt for the particles that collide; f = NULL: none):
 *   L = mean over the fluid z layers of (<vbar_x>(z) - target(z))^2,
 * <vbar_x>(z) = mean over the steps of the mean x velocity of the particles
 * whose stored position lies in layer z after the step.  target holds one
 * value per fluid layer (n_layers = Z - 2 with walls, Z without), host memory.
 * The gradient is pathwise: cells, cell tables, wall bounces, rotation axes
 * and ghosts of the trajectory held fixed (the step is then linear in the
 * velocities); the trajectory is recomputed from checkpoints taken every
 * checkpoint_every steps.  force and grad_force hold C floats in reference
 * cell order, in host memory (a process may hold a second HIP runtime, e.g.
 * torch's, whose device allocations this library cannot address safely);
 * NULL: no force / no force gradient.  The handle's state,
 * step index and parameters are unchanged afterwards (the same call repeats
 * the same trajectory: common random numbers).  profile (n_layers doubles,
 * host) may be NULL.  The layers' x-velocity sums are exact fixed-point sums
 * (multiples of 2^-40) while every |v_x| < 128: a run that meets a larger
 * velocity, or one that is not a number, fails with an error. */
int mpcd_hip_profile_loss(mpcd_hip_t* h, int steps, float drag, float rotation_angle, const float* force,
                          const double* target, size_t n_layers, double* loss, double* profile);
int mpcd_hip_profile_loss_grad(mpcd_hip_t* h, int steps, int checkpoint_every, float drag, float rotation_angle,
                               const float* force, const double* target, size_t n_layers, double* loss,
                               double* profile, double* grad_drag, double* grad_angle, float* grad_force);

/* ---- ensembles (no reference counterpart: the reference runs one system per
 * process, cpu_backend.cpp / cuda_backend_impl.cu) ---- */

/* n_systems independent systems advanced together: each kernel of a step
 * covers every system in one launch (per-system parameter table, block ->
 * system maps).  System s owns particle slots [first_particle, first_particle +
 * n_particles) of the flattened particle arrays the entry points above read and
 * write, and cells [first_cell, first_cell + n_cells) of the flattened cell
 * arrays (get_sample_mean, peek_sort: cell index first_cell + x + X (y + Y z)).
 * All systems use the same algorithm and device; n, extents, walls, drag,
 * temperature, rotation angle and seed are per system.  System s reproduces,
 * bit for bit, a handle created from params[s] alone.  mpcd_hip_create is
 * mpcd_hip_create_ensemble with one system. */
int mpcd_hip_create_ensemble(const mpcd_params_t* params, uint32_t n_systems, mpcd_hip_t** out);
int mpcd_hip_n_systems(const mpcd_hip_t* h, uint32_t* out);
int mpcd_hip_system_info(const mpcd_hip_t* h, uint32_t system, uint64_t* first_particle, uint64_t* n_particles,
                         uint64_t* first_cell, uint64_t* n_cells);
/* Positions and/or velocities (either may be NULL) of one system, interleaved xyz. */
int mpcd_hip_system_get_particles(mpcd_hip_t* h, uint32_t system, float* pos, float* vel, size_t n_particles);
/* Mean fields of one system's cells, scaled as mpcd_hip_get_sample_mean; the
 * current average continues (mpcd_hip_reset_sampling ends it for all systems). */
int mpcd_hip_system_sample_mean(mpcd_hip_t* h, uint32_t system, float* density, float* velocity, size_t n_cells);
int mpcd_hip_system_sample_mean_f64(mpcd_hip_t* h, uint32_t system, double* density, double* velocity, size_t n_cells);
int mpcd_hip_reset_sampling(mpcd_hip_t* h);
/* Sampling state, for checkpoints (the reference's backup writer is a stub,
 * cuda_backend_impl.cu:77-79, cpu_backend.cpp:61-63): the FP64 accumulators
 * (4 per cell in the order of the mean-field arrays, cell first_cell + x +
 * X (y + Y z): density and velocity sums of the per-step means, n_cells = all
 * systems' cells), the samples after the first, and whether an average is
 * running.  Setting it resumes the average exactly where it was saved. */
int mpcd_hip_get_sampling_state(mpcd_hip_t* h, double* acc, size_t n_cells, uint64_t* sample_counter, int* active);
int mpcd_hip_set_sampling_state(mpcd_hip_t* h, const double* acc, size_t n_cells, uint64_t sample_counter, int active);
/* Replay each step from a HIP graph captured on first use (default on; steps
 * with timing, permutation recording or the first sample of an average run
 * eagerly).  The step state (index, grid shifts) lives in device memory and is
 * advanced by the last kernel of the step, so one graph serves every step. */
int mpcd_hip_use_graph(mpcd_hip_t* h, int enable);
/* Autotuner: for each density class (the systems of one n, the most particles
 * first) tries each compiled LDS capacity that holds a whole collision table
 * with several cells-per-tile limits, timing that class's own collide
 * launches with HIP events over steps_per_trial real steps (after one untimed
 * step; a trial counts its fastest step), times the three fastest shapes once
 * more, and keeps the fastest for the class.  Returns system 0's shape and the
 * time per step with the chosen shapes.  The systems advance by these steps;
 * results do not depend on the tile shapes. */
int mpcd_hip_autotune(mpcd_hip_t* h, int steps_per_trial, uint32_t* cells_per_tile, uint32_t* lds_capacity,
                      double* ms_per_step);
/* LDS capacities the autotuner tries for `algorithm` (0 srd, 1 extended) when
 * the densest system has nmax particles per cell: the compiled capacities the
 * rule can run whose batch holds a whole collision table (4 nmax).  Writes up
 * to max_caps of them and their number to *n_caps; needs no device. */
int mpcd_hip_tune_capacities(int algorithm, uint32_t nmax, uint32_t* caps, uint32_t max_caps, uint32_t* n_caps);

/* ---- allocation-free step (no reference counterpart) ----
 *
 * For callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc): n_steps of
 * the step mpcd_hip_step runs, for one system described by *params, on a
 * particle state the caller holds in device memory, with the same results
 * bit for bit as a handle created from the same parameters, holding the same
 * state at the same step index.  It calls no hipMalloc or hipFree: the step's
 * sort and per-cell arrays live in a static device array of the library
 * (14 MiB per device, held until the process ends).
 *
 * state_in / state_out: device buffers of n_particles records (x, y, z, vx,
 * vy, vz as 6 floats, 24 bytes per particle), of the HIP runtime this library
 * uses; the result is written to state_out in cell order (as a handle keeps
 * it); state_in is overwritten when n_steps > 1.  step_index keys the random
 * stream (mpcd_hip_get_step_index).  Limits: at most 2^20 particles and 2^17
 * cells, and with walls 2 X Y at most 2^20 / 3 (SRD) or n * 2 X Y at most
 * 2^20 / 6 (extended); larger systems are rejected with an error before any
 * device work.  Runs on the
 * null stream and returns after the device finished; calls from one process
 * are serialised, each call leaves the array holding its own data and clears
 * what the next needs cleared.  Errors: nonzero return, message in
 * mpcd_hip_last_error(NULL). */
int mpcd_hip_step_noalloc(const mpcd_params_t* params, uint64_t step_index, int n_steps, float* state_in,
                          float* state_out, size_t n_particles);

/* Copies `bytes` between two device buffers `reps` times with each of four
 * copy kernels (4 or 8 float4 loads in flight per lane, plain or nontemporal
 * stores) and returns the mean time per copy of the fastest: the practical
 * HBM reference.  No counterpart in the reference (measurement only). */
int mpcd_hip_bench_copy(size_t bytes, int reps, double* ms_per_copy);
/* Device math of the random stream, for known-answer tests: Philox4x32-10 of
 * n counters (4 words each), and the float transforms. */
int mpcd_hip_selftest_philox(uint32_t k0, uint32_t k1, const uint32_t* ctr, uint32_t* out, size_t n);
int mpcd_hip_selftest_math(const float* in, float* out, size_t n, int which);

#ifdef __cplusplus
}
#endif
#endif /* MPCD_HIP_H */
