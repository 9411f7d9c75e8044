// Python module cudampcd_amd._pympcd: the drop-in surface of the reference's
// pybind11 module _pympcd (src/bindings/bindings.cpp:49-178).
//
// Simulation(params, backend) with step / step_and_sample / get_mean_fields /
// get_particle_positions / get_particle_velocities, and Params with the same
// properties and setter side effects (wall layers added by volume_size and
// periodicity, N recomputed, thermal_velocity = sqrt(temperature)).
// Additions: Params.rotation_angle / seed / system_id, and Simulation
// set_particle_* (the reference binds none), get_mean_fields_f64, seed,
// n_particles and the raw C-ABI handle for the parity hooks; Ensemble, a set
// of independent systems advanced together (no reference counterpart).
//
// Deliberate fix: get_mean_fields uses strides (4, 4X, 4XY) for shape (X, Y, Z);
// the reference uses (4, 4Y, 4XY) (bindings.cpp:214-216), which equals it only
// when X == Y and reads outside the buffer when Y > X.

#include <pybind11/numpy.h>
#include <pybind11/pybind11.h>
#include <pybind11/stl.h>

#include <cmath>
#include <cstdint>
#include <vector>

#include <mpcd/api/simulation_handle.hpp>
#include <mpcd/api/simulation_parameters.hpp>

#include "common/backend.hpp"
#include "hip_backend/hip_backend.hpp"

namespace py = pybind11;

using mpcd::CollisionKernel;
using mpcd::ExperimentType;
using mpcd::MPCDAlgorithm;
using mpcd::SimulationParameters;
using mpcd::api::SimulationHandle;

namespace {

template <typename T>
py::array_t<T> to_numpy(std::vector<T>&& vec, std::vector<ssize_t> shape, std::vector<ssize_t> strides) {
    auto* heap = new std::vector<T>(std::move(vec));
    py::capsule owner(heap, [](void* p) { delete static_cast<std::vector<T>*>(p); });
    return py::array_t<T>(shape, strides, heap->data(), owner);
}

// N as the reference setters compute it: n times the fluid extents, in float
// arithmetic (bindings.cpp:113-116)
void recompute_fluid_N(SimulationParameters& p) {
    p.N = p.n * (p.volume_size[0] - 2 * (1 - p.periodicity[0])) * (p.volume_size[1] - 2 * (1 - p.periodicity[1])) *
          (p.volume_size[2] - 2 * (1 - p.periodicity[2]));
}

std::vector<float> flat3(py::array_t<float, py::array::c_style | py::array::forcecast> a) {
    if (a.size() % 3 != 0) throw std::runtime_error("expected an array of 3-vectors");
    return std::vector<float>(a.data(), a.data() + a.size());
}

mpcd::hip::HipBackend& hip_of(SimulationHandle& s) {
    auto* b = dynamic_cast<mpcd::hip::HipBackend*>(&s.backend());
    if (!b) throw std::runtime_error("not a hip backend");
    return *b;
}

template <typename T>
py::tuple mean_fields_t(SimulationParameters const& p, std::vector<T>&& density, std::vector<T>&& velocity) {
    ssize_t const x = p.volume_size[0], y = p.volume_size[1], z = p.volume_size[2];
    ssize_t const sf = static_cast<ssize_t>(sizeof(T));
    auto d = to_numpy(std::move(density), {x, y, z}, {sf, x * sf, x * y * sf});
    auto v = to_numpy(std::move(velocity), {x, y, z, 3}, {3 * sf, 3 * x * sf, 3 * x * y * sf, sf});
    return py::make_tuple(d, v);
}

py::tuple mean_fields(SimulationParameters const& p, std::vector<float>&& density, std::vector<float>&& velocity) {
    ssize_t const x = p.volume_size[0], y = p.volume_size[1], z = p.volume_size[2];
    ssize_t const sf = static_cast<ssize_t>(sizeof(float));
    std::vector<ssize_t> shape = {x, y, z}, strides = {sf, x * sf, x * y * sf};
    auto d = to_numpy(std::move(density), shape, strides);
    for (auto& s : strides) s *= 3;
    strides.push_back(sf);
    shape.push_back(3);
    auto v = to_numpy(std::move(velocity), shape, strides);
    return py::make_tuple(d, v);
}

}  // namespace

PYBIND11_MODULE(_pympcd, m) {
    m.doc() = "MPCD simulation on AMD Instinct GPUs (hip backend); drop-in for pympcd.";

    py::class_<SimulationHandle>(m, "Simulation", "Handle class for managing simulation state and interaction with backend")
        .def(py::init<SimulationParameters const&, std::string>(), py::arg("params"), py::arg("backend"))
        .def("step", &SimulationHandle::step, "Perform n_steps simulation step.", py::call_guard<py::gil_scoped_release>())
        .def("step_and_sample", &SimulationHandle::stepAndAccumulateSample,
             "Perform n_steps simulation step and accumulate statistics.", py::call_guard<py::gil_scoped_release>())
        .def("get_mean_fields",
             [](SimulationHandle& self) {
                 std::vector<float> density, velocity;
                 self.getSampleMean(density, velocity);
                 return mean_fields(self.getParameters(), std::move(density), std::move(velocity));
             },
             "Return the mean density and velocity fields (array of voxels) of the fluid after accumulating statistics.")
        .def("get_mean_fields_f64",
             [](SimulationHandle& self) {
                 std::vector<double> d, v;
                 hip_of(self).getSampleMeanF64(d, v);
                 auto const& p = self.getParameters();
                 ssize_t const x = p.volume_size[0], y = p.volume_size[1], z = p.volume_size[2], sd = 8;
                 auto dn = to_numpy(std::move(d), {x, y, z}, {sd, x * sd, x * y * sd});
                 auto vn = to_numpy(std::move(v), {x, y, z, 3}, {3 * sd, 3 * x * sd, 3 * x * y * sd, sd});
                 return py::make_tuple(dn, vn);
             },
             "FP64 mean fields of the current average (does not end it).")
        .def("get_particle_positions",
             [](SimulationHandle const& self) {
                 std::vector<float> pos;
                 self.getParticlePositions(pos);
                 ssize_t const n = pos.size() / 3;
                 return to_numpy(std::move(pos), {n, 3}, {12, 4});
             },
             "Returns the positions of the particles as array of vectors.")
        .def("get_particle_velocities",
             [](SimulationHandle const& self) {
                 std::vector<float> vel;
                 self.getParticleVelocities(vel);
                 ssize_t const n = vel.size() / 3;
                 return to_numpy(std::move(vel), {n, 3}, {12, 4});
             },
             "Returns the velocities of the particles as array of vectors.")
        .def("set_particle_positions",
             [](SimulationHandle& self, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 self.setParticlePositions(flat3(a));
             })
        .def("set_particle_velocities",
             [](SimulationHandle& self, py::array_t<float, py::array::c_style | py::array::forcecast> a) {
                 self.setParticleVelocities(flat3(a));
             })
        .def("set_particles",
             [](SimulationHandle& self, py::array_t<float, py::array::c_style | py::array::forcecast> p,
                py::array_t<float, py::array::c_style | py::array::forcecast> v) { self.setParticles(flat3(p), flat3(v)); })
        .def("get_parameters", [](SimulationHandle const& self) { return self.getParameters(); },
             "A copy of the parameters the simulation was created with.")
        .def_property_readonly("n_particles", [](SimulationHandle& self) { return hip_of(self).getNParticles(); })
        .def_property_readonly("seed", [](SimulationHandle& self) { return hip_of(self).seed(); })
        .def("_handle", [](SimulationHandle& self) { return reinterpret_cast<uintptr_t>(hip_of(self).handle()); },
             "Address of the mpcd_hip_t handle, for calls into the C-ABI.");

    using mpcd::hip::HipEnsemble;
    py::class_<HipEnsemble>(m, "Ensemble",
                            "Independent systems (one Params each) advanced together on one device; system i "
                            "reproduces Simulation(params[i], 'hip') with the same seed bit for bit")
        .def(py::init<std::vector<SimulationParameters> const&>(), py::arg("params"))
        .def("step", &HipEnsemble::step, py::call_guard<py::gil_scoped_release>())
        .def("step_and_sample", &HipEnsemble::stepAndAccumulateSample, py::call_guard<py::gil_scoped_release>())
        .def_property_readonly("n_systems", &HipEnsemble::nSystems)
        .def("n_particles", &HipEnsemble::nParticles)
        .def("seed", &HipEnsemble::seed)
        .def("get_mean_fields",
             [](HipEnsemble& self, uint32_t i) {
                 std::vector<float> d, v;
                 self.getSampleMean(i, d, v);
                 return mean_fields_t(self.parameters(i), std::move(d), std::move(v));
             },
             "Mean density (X, Y, Z) and velocity (X, Y, Z, 3) of system i; the average continues until reset_sampling().")
        .def("get_mean_fields_f64",
             [](HipEnsemble& self, uint32_t i) {
                 std::vector<double> d, v;
                 self.getSampleMeanF64(i, d, v);
                 return mean_fields_t(self.parameters(i), std::move(d), std::move(v));
             })
        .def("reset_sampling", &HipEnsemble::resetSampling)
        .def("get_particle_positions",
             [](HipEnsemble& self, uint32_t i) {
                 std::vector<float> pos;
                 self.getParticles(i, &pos, nullptr);
                 ssize_t const n = pos.size() / 3;
                 return to_numpy(std::move(pos), {n, 3}, {12, 4});
             })
        .def("get_particle_velocities",
             [](HipEnsemble& self, uint32_t i) {
                 std::vector<float> vel;
                 self.getParticles(i, nullptr, &vel);
                 ssize_t const n = vel.size() / 3;
                 return to_numpy(std::move(vel), {n, 3}, {12, 4});
             })
        .def("use_graph", &HipEnsemble::useGraph, py::arg("enable") = true)
        .def("autotune",
             [](HipEnsemble& self, int steps_per_trial) {
                 uint32_t tc = 0, cap = 0;
                 double ms = 0;
                 {
                     py::gil_scoped_release release;
                     self.autotune(steps_per_trial, tc, cap, ms);
                 }
                 return py::make_tuple(tc, cap, ms);
             },
             py::arg("steps_per_trial") = 3,
             "Times the compiled tile shapes on real steps (the systems advance) and keeps the fastest; "
             "returns (cells per tile of system 0, LDS capacity, ms per step).")
        .def("_handle", [](HipEnsemble& self) { return reinterpret_cast<uintptr_t>(self.handle()); });

    py::class_<SimulationParameters>(m, "Params", "Class representing simulation parameters")
        .def(py::init<>())
        .def_readwrite("device_id", &SimulationParameters::device_id)
        .def_readwrite("delta_t", &SimulationParameters::delta_t)
        .def_readwrite("N", &SimulationParameters::N)
        .def_readwrite("drag", &SimulationParameters::drag)
        .def_readwrite("equilibration_steps", &SimulationParameters::equilibration_steps)
        .def_readwrite("steps", &SimulationParameters::steps)
        .def_readwrite("sample_rate", &SimulationParameters::sample_every)
        .def_readwrite("average_samples", &SimulationParameters::average_samples)
        .def_property(
            "temperature", [](const SimulationParameters& p) { return p.temperature; },
            [](SimulationParameters& p, double temperature) {
                p.temperature = temperature;
                p.thermal_velocity = std::sqrt(temperature);
            })
        .def_property(
            "volume_size",
            [](const SimulationParameters& p) { return py::make_tuple(p.volume_size[0], p.volume_size[1], p.volume_size[2]); },
            [](SimulationParameters& p, py::tuple t) {
                if (t.size() != 3) throw std::runtime_error("Expected 3 elements");
                for (int i = 0; i < 3; ++i) {
                    p.volume_size[i] = t[i].cast<float>();
                    if (p.periodicity[i] == 0) p.volume_size[i] += 2;  // wall layer
                }
                recompute_fluid_N(p);
            })
        .def_property(
            "n", [](const SimulationParameters& p) { return p.n; },
            [](SimulationParameters& p, int n) {
                p.n = n;
                p.N = p.n * p.volume_size[0] * p.volume_size[1] * p.volume_size[2];
            })
        .def_property(
            "periodicity",
            [](const SimulationParameters& p) { return py::make_tuple(p.periodicity[0], p.periodicity[1], p.periodicity[2]); },
            [](SimulationParameters& p, py::tuple t) {
                if (t.size() != 3) throw std::runtime_error("Expected 3 elements");
                for (int i = 0; i < 3; ++i) {
                    p.periodicity[i] = t[i].cast<int>();
                    if (p.periodicity[i] == 0 and p.volume_size[i] != 0) p.volume_size[i] += 2;  // wall layer
                }
                recompute_fluid_N(p);
            })
        .def_property(
            "experiment",
            [](const SimulationParameters& p) { return p.experiment == ExperimentType::standard ? "standard" : "channel"; },
            [](SimulationParameters& p, const std::string& e) {
                p.experiment = e == "standard" ? ExperimentType::standard : ExperimentType::channel;
            })
        .def_property(
            "algorithm", [](const SimulationParameters& p) { return p.algorithm == MPCDAlgorithm::srd ? "srd" : "extended"; },
            [](SimulationParameters& p, const std::string& a) {
                p.algorithm = a == "srd" ? MPCDAlgorithm::srd : MPCDAlgorithm::extended;
            })
        .def_property(
            "collision_kernel",
            [](const SimulationParameters& p) -> std::string {
                switch (p.collision_kernel) {
                    case CollisionKernel::trivial: return "trivial";
                    case CollisionKernel::sorting: return "sorting";
                    default: return "optimized";
                }
            },
            [](SimulationParameters& p, const std::string& k) {
                if (k == "trivial") p.collision_kernel = CollisionKernel::trivial;
                else if (k == "sorting") p.collision_kernel = CollisionKernel::sorting;
                else p.collision_kernel = CollisionKernel::optimized;
            })
        // Not a property of the reference's Params: the temperature setter derives
        // it as sqrt(double T) while the parameter file reader takes sqrtf of the
        // float T (simulation_parameters.cpp:182), which can differ in the last
        // bit; checkpoints and the file driver set it directly.
        .def_readwrite("thermal_velocity", &SimulationParameters::thermal_velocity)
        .def_readwrite("rotation_angle", &SimulationParameters::rotation_angle)
        .def_property(
            "seed", [](const SimulationParameters& p) -> py::object { if (p.seed_set) return py::int_(p.seed); return py::none(); },
            [](SimulationParameters& p, py::object s) {
                if (s.is_none()) {
                    p.seed_set = false;
                    p.seed = 0;
                } else {
                    p.seed = s.cast<uint64_t>();
                    p.seed_set = true;
                }
            })
        .def_readwrite("system_id", &SimulationParameters::system_id);

    m.attr("__version__") = "1.0.0";
}
