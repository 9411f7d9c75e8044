"""Checkpoint and restart: a restored simulation continues bit for bit, also
in the middle of an average; snapshots of the mean fields round-trip."""
from __future__ import annotations

import numpy as np
import pytest

from cudampcd_amd import checkpoint, io
from cudampcd_amd.capi import Handle
from helpers import make_params

pytestmark = pytest.mark.gpu


# T = 0.9: the thermal velocity (wall ghosts, extended thermal draws) must be
# restored as saved, not recomputed from the float32 temperature
@pytest.mark.parametrize("algorithm,temperature", [("srd", 1.0), ("extended", 1.0), ("srd", 0.9), ("extended", 0.9)])
def test_restart_continues_bitwise(pympcd, tmp_path, algorithm, temperature):
    p = make_params(pympcd, (10, 8, 12), (1, 1, 0), n=8, algorithm=algorithm, seed=321, drag=0.002,
                    temperature=temperature)
    a = pympcd.Simulation(p, "hip")
    a.step(12)
    a.step_and_sample(5)
    f = checkpoint.save(a, tmp_path / "ck.npz")
    b = checkpoint.restore(f)
    assert Handle(b).step_index == Handle(a).step_index
    for s in (a, b):
        s.step(7)
        s.step_and_sample(4)
    assert np.array_equal(a.get_particle_positions(), b.get_particle_positions())
    assert np.array_equal(a.get_particle_velocities(), b.get_particle_velocities())
    da, va = a.get_mean_fields_f64()
    db, vb = b.get_mean_fields_f64()
    assert np.array_equal(da, db) and np.array_equal(va, vb)
    d32a, v32a = a.get_mean_fields()
    d32b, v32b = b.get_mean_fields()
    assert np.array_equal(d32a, d32b) and np.array_equal(v32a, v32b)


def test_restore_without_sampling_and_snapshot(pympcd, tmp_path):
    p = make_params(pympcd, (6, 6, 10), (1, 1, 1), n=5, seed=9)
    a = pympcd.Simulation(p, "hip")
    a.step(3)
    b = checkpoint.restore(checkpoint.save(a, tmp_path / "c0.npz"))
    a.step(4)
    b.step(4)
    assert np.array_equal(a.get_particle_velocities(), b.get_particle_velocities())
    a.step_and_sample(6)
    d, v = a.get_mean_fields()
    io.write_fluid_snapshot(tmp_path / "out", 13, d, v)
    d2, v2 = io.read_fluid_snapshot(tmp_path / "out", 13)
    assert np.array_equal(d, d2) and np.array_equal(v, v2)
    # 6 samples summed and divided by 5, the reference's S - 1 (cpu_backend.cpp:111-122)
    assert abs(float(d2.sum()) - 6 / 5 * a.n_particles) < 1e-3 * a.n_particles
