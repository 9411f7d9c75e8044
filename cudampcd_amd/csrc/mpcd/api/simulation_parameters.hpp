// Simulation parameters of the MPCD API.
//
// Same fields, defaults and meaning as mpcd::SimulationParameters
// (reference include/mpcd/api/simulation_parameters.hpp:7-55), with every
// field initialised (the reference leaves N and the CLI counters undefined
// until a setter runs), plus the options the reference fixes in code:
//   rotation_angle  SRD rotation angle in degrees (fixed to 120 at
//                   simulation_functions.cpp:118-119)
//   seed            key of the counter-based random stream (the reference
//                   seeds from /dev/urandom, random.hpp:39-46); drawn from
//                   std::random_device when seed_set is false
//   system_id       ensemble member folded into the random stream
#pragma once

#include <array>
#include <cstdint>
#include <string>

namespace mpcd {

enum class ExperimentType {
    standard,  // 3D periodic domain, optional walls in z
    channel,   // cylindrical channel about the x axis (simulation_functions.cpp:11-24, 75-82)
};

enum class MPCDAlgorithm {
    srd,       // stochastic rotation dynamics
    extended,  // extended MPC with non-ideal equation of state
};

enum class CollisionKernel {
    trivial,
    sorting,
    optimized,
};

struct SimulationParameters {
    ExperimentType experiment = ExperimentType::standard;
    MPCDAlgorithm algorithm = MPCDAlgorithm::srd;
    CollisionKernel collision_kernel = CollisionKernel::optimized;

    std::array<int, 3> periodicity = {1, 1, 1};
    std::array<float, 3> volume_size = {10.0f, 10.0f, 10.0f};

    unsigned int N = 0,       // number of particles
        n = 10,               // number of particles per cell
        device_id = 0;

    float delta_t = 0.01f,    // SRD time step
        temperature = 1.0f,
        thermal_velocity = 1.0f,
        drag = 0.0f;          // body force driving Poiseuille flow

    // CLI-only fields of the reference, kept for API compatibility
    unsigned int equilibration_steps = 0, steps = 0, sample_every = 1, average_samples = 1;
    bool sample_fluid = false;
    bool write_backup = false, read_backup = false;
    std::string output_directory, load_directory;

    // options the reference fixes in code
    float rotation_angle = 120.0f;
    uint64_t seed = 0;
    bool seed_set = false;
    uint32_t system_id = 0;
};

}  // namespace mpcd
