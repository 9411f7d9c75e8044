"""Physics checks of the HIP backend at sizes the oracle cannot reach.

* Poiseuille flow at the reference's README geometry (10x10x100 fluid cells,
  walls in z, n=10, T=1, g=1e-3, dt=0.02; README.md:52-102): viscosity within
  2% of the SRD formula (37.50) and of the reference's measurement (37.27).
  The flow is sampled once developed (slowest mode relaxes in about 13500
  steps from rest).
* Conservation with periodic boundaries and no drag: total momentum (SRD and
  extended) and kinetic energy (SRD) to FP32 round-off; drag adds N g dt.
"""
from __future__ import annotations

import numpy as np
import pytest

from cudampcd_amd.capi import Handle
from cudampcd_amd.physics import poiseuille_viscosity, srd_viscosity
from helpers import make_params

pytestmark = pytest.mark.gpu


def test_poiseuille_viscosity_within_2_percent(pympcd):
    n, g, dt = 10, 0.001, 0.02
    p = make_params(pympcd, (10, 10, 100), (1, 1, 0), n=n, seed=777, drag=g, delta_t=dt)
    sim = pympcd.Simulation(p, "hip")
    sim.step(1000)
    sim.step_and_sample(10000)
    _, vel_short = sim.get_mean_fields()
    eta_short = poiseuille_viscosity(vel_short, n, g, 100)
    sim.step(60000)
    sim.step_and_sample(40000)
    dens, vel = sim.get_mean_fields()
    eta = poiseuille_viscosity(vel, n, g, 100)
    theory = srd_viscosity(n, dt)
    print(f"eta developed={eta:.3f} (1000+10000 protocol: {eta_short:.3f}) theory={theory:.3f} reference=37.27")
    assert abs(eta - theory) / theory < 0.02
    assert abs(eta - 37.27) / 37.27 < 0.02
    assert abs(dens[:, :, 1:-1].mean() - n) < 0.05


@pytest.mark.parametrize("algorithm", ["srd", "extended"])
def test_momentum_energy_conservation(pympcd, algorithm):
    p = make_params(pympcd, (16, 12, 10), (1, 1, 1), n=10, algorithm=algorithm, seed=5, drag=0.0)
    sim = pympcd.Simulation(p, "hip")
    v0 = sim.get_particle_velocities().astype(np.float64)
    sim.step(50)
    v1 = sim.get_particle_velocities().astype(np.float64)
    P0, P1 = v0.sum(axis=0), v1.sum(axis=0)
    assert np.abs(P1 - P0).max() < 1e-5 * len(v0)
    assert not np.array_equal(np.sort(v0[:, 0]), np.sort(v1[:, 0]))
    if algorithm == "srd":
        E0, E1 = (v0 ** 2).sum(), (v1 ** 2).sum()
        assert abs(E1 - E0) / E0 < 1e-5


def test_drag_momentum_gain(pympcd):
    g, dt = 0.01, 0.02
    p = make_params(pympcd, (16, 12, 10), (1, 1, 1), n=10, seed=6, drag=g, delta_t=dt)
    sim = pympcd.Simulation(p, "hip")
    v0 = sim.get_particle_velocities().astype(np.float64)
    sim.step(10)
    v1 = sim.get_particle_velocities().astype(np.float64)
    gain = (v1[:, 0].sum() - v0[:, 0].sum()) / 10
    assert abs(gain - len(v0) * np.float32(g * dt)) / (len(v0) * g * dt) < 2e-3


def test_timing_hooks(pympcd):
    p = make_params(pympcd, (64, 64, 20), (1, 1, 0), n=20, seed=8)
    sim = pympcd.Simulation(p, "hip")
    h = Handle(sim)
    h.enable_timing(True)
    sim.step(5)
    collide_ms, total_ms, launches = h.last_timing()
    assert 0 < collide_ms <= total_ms
    assert launches >= 5
