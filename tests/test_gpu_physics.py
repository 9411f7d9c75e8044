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
from oracle_lib import Oracle

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


@pytest.mark.parametrize("algorithm", ["srd", "extended"])
def test_walled_cells_conserve_momentum(pympcd, algorithm):
    """Walls in z: every collision cell outside the two ghost layers conserves
    the momentum of its particles plus g dt each (simulation_functions.cpp:121,
    187, 384-391; the extended rule's thermal and angular terms sum to zero over
    the cell), a lone particle keeps its velocity (no drag, simulation_functions.cpp:169-189), and the ghost
    layers exchange momentum with the walls.  Device velocities after a step
    against the streamed velocities and shifted cells of the oracle's
    streaming half-step from the same state (bit-exact with the device's)."""
    g, dt, n = 0.002, 0.02, 10
    p = make_params(pympcd, (20, 16, 12), (1, 1, 0), n=n, algorithm=algorithm, seed=515, drag=g, delta_t=dt)
    sim = pympcd.Simulation(p, "hip")
    h = Handle(sim)
    sim.step(3)
    h.record_permutation(True)
    X, Y, Z = (int(v) for v in p.volume_size)
    cells = np.arange(X * Y * Z)
    checked = 0
    for _ in range(3):
        pos, vel = sim.get_particle_positions(), sim.get_particle_velocities()
        o = Oracle((X, Y, Z), (1, 1, 0), n=n, drag=g, delta_t=dt, algorithm=algorithm, rng="counter",
                   seed=(p.seed, 0), distribute=False)
        o.set_particles(pos, vel)
        o.set_step(h.step_index)
        shift = o.translate()
        cell, cnt, _ = o.cells()
        _, streamed = o.particles()
        sim.step(1)
        perm = h.last_permutation()
        after = np.empty_like(vel)
        after[perm] = sim.get_particle_velocities()       # back to the input order
        dP = np.zeros((len(cells), 3))
        np.add.at(dP, cell, after.astype(np.float64) - streamed)
        sign = int(shift[2] > 0)
        cz = cells // (X * Y)
        ghost_layer = (cz == sign) | (cz == Z - (1 if sign else 2))
        full = ~ghost_layer & (cnt >= 2) & (cnt <= 4 * n)
        lone = ~ghost_layer & (cnt == 1)
        want = np.zeros_like(dP)
        want[:, 0] = cnt * np.float64(np.float32(g * dt))
        assert np.abs(dP - want)[full].max() < 5e-5
        assert not dP[lone].any()
        assert np.abs(dP[ghost_layer & (cnt > 0)]).max() > 0.1
        checked += int(full.sum())
    # the fluid's shifted cells span Z - 1 layers, two of them ghost layers
    assert checked > 0.9 * 3 * X * Y * (Z - 3)


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
