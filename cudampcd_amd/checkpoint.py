"""Checkpoint and restart of a Simulation (the reference's writeBackupFile is a
stub: cuda_backend_impl.cu:77-79, cpu_backend.cpp:61-63).

A checkpoint holds the particles in their current order, the step index that
keys the counter-based random stream, the seed and parameters, and the FP64
sampling accumulators, so a restored simulation continues bit for bit where
the saved one would have (tested): same draws, same cell order, same average.
Files are numpy .npz archives without pickled objects.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np

from . import capi

_SCALARS = ("n", "temperature", "thermal_velocity", "drag", "delta_t", "rotation_angle", "system_id", "N", "device_id")


def save(sim, path) -> Path:
    """Write the simulation's state to ``path`` (.npz)."""
    h = capi.Handle(sim)
    p = sim.get_parameters()
    acc, count, active = h.sampling_state()
    data = {
        "positions": sim.get_particle_positions(),
        "velocities": sim.get_particle_velocities(),
        "step_index": np.uint64(h.step_index),
        "sampling_acc": acc,
        "sampling_count": np.uint64(count),
        "sampling_active": np.bool_(active),
    }
    data.update(param_record(p, sim.seed))
    path = Path(path)
    np.savez(path, **data)
    return path if path.suffix == ".npz" else path.with_name(path.name + ".npz")


def param_record(p, seed) -> dict:
    """The arrays a checkpoint keeps of Params `p` (and the stream's seed)."""
    data = {
        "seed": np.uint64(seed),
        "volume_size": np.asarray(p.volume_size, np.float32),
        "periodicity": np.asarray(p.periodicity, np.int32),
        "algorithm": np.asarray(p.algorithm),
        "experiment": np.asarray(p.experiment),
    }
    for k in _SCALARS:
        data[k] = np.asarray(getattr(p, k))
    return data


def params_of(z):
    """The Params a checkpoint archive was saved with."""
    import cudampcd_amd as pympcd

    p = pympcd.Params()
    p.n = int(z["n"])
    p.temperature = float(z["temperature"])
    # the thermal velocity as it was: recomputing it from the stored float32
    # temperature can change its last bit
    if "thermal_velocity" in z:
        p.thermal_velocity = float(z["thermal_velocity"])
    perio = tuple(int(v) for v in z["periodicity"])
    stored = [float(v) for v in z["volume_size"]]
    p.periodicity = perio
    p.volume_size = tuple(stored[a] - (0 if perio[a] else 2) for a in range(3))   # the setter adds the wall layers
    p.drag = float(z["drag"])
    p.delta_t = float(z["delta_t"])
    p.rotation_angle = float(z["rotation_angle"])
    p.system_id = int(z["system_id"])
    p.device_id = int(z["device_id"])
    p.algorithm = str(z["algorithm"])
    p.experiment = str(z["experiment"])
    p.seed = int(z["seed"])
    if tuple(float(v) for v in p.volume_size) != tuple(stored) or int(p.N) != int(z["N"]):
        raise ValueError("checkpoint parameters do not reproduce the saved grid")
    return p


def restore(path, backend: str = "hip"):
    """A new Simulation in the saved state."""
    import cudampcd_amd as pympcd

    z = np.load(path, allow_pickle=False)
    p = params_of(z)
    sim = pympcd.Simulation(p, backend)
    sim.set_particles(z["positions"], z["velocities"])
    h = capi.Handle(sim)
    h.step_index = int(z["step_index"])
    h.set_sampling_state(z["sampling_acc"], int(z["sampling_count"]), bool(z["sampling_active"]))
    return sim
