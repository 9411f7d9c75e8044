"""Physics checks of the HIP backend at sizes the oracle cannot reach.

* Poiseuille flow at the reference's README geometry (10x10x100 fluid cells,
  walls in z, n=10, T=1, g=1e-3, dt=0.02; README.md:52-102): under the
  reference's sampling protocols (1000 + 10,000 and 1000 + 100,000 steps) the
  viscosity estimate matches the reference CPU backend's under the same
  protocols; sampled once developed (the slowest mode relaxes in about 13500
  steps from rest) it lies within 2% of the SRD formula (37.50) and of the
  reference's printed measurement (37.27).
* Conservation with periodic boundaries and no drag: total momentum (SRD and
  extended) and kinetic energy (SRD) to FP32 round-off; drag adds N g dt.
"""
from __future__ import annotations

import json
from pathlib import Path

import numpy as np
import pytest

from cudampcd_amd.capi import Handle
from cudampcd_amd.physics import poiseuille_viscosity, srd_viscosity
from helpers import make_params
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu


REFERENCE_C1 = json.loads((Path(__file__).parent / "golden" / "c1_reference_cpu.json").read_text())


def _reference(protocol):
    """The reference CPU backend's viscosity under a protocol: the mean over its seeds."""
    r = REFERENCE_C1[f"reference_{protocol}"]
    return r.get("eta_mean", r["eta"]), np.array(r.get("profile_x_mean", r["profile_x"]))


def test_poiseuille_viscosity_within_2_percent(pympcd):
    """Sampled from the start of the flow, as the reference's protocols do, the
    viscosity estimate carries the flow's development: the reference CPU
    backend itself reads about 113 after 1000 + 10,000 steps and 42 after
    1000 + 100,000 (the README's protocol, whose printed 37.27 it does not
    reproduce; tests/golden/c1_reference_cpu.json).  One run's estimate
    scatters by about 4% and 0.6% under these protocols, so eight seeds run as
    one ensemble and their mean is compared with the reference's mean under
    the same protocol; once the flow is developed the mean lies within 2% of
    theory (37.50) and of the README's figure."""
    n, g, dt, K = 10, 0.001, 0.02, 8
    theory = srd_viscosity(n, dt)

    def channels(seed0):
        ps = []
        for i in range(K):
            p = make_params(pympcd, (10, 10, 100), (1, 1, 0), n=n, seed=seed0 + i, drag=g, delta_t=dt)
            p.system_id = i
            ps.append(p)
        return pympcd.Ensemble(ps)

    def etas(ens):
        return np.array([poiseuille_viscosity(ens.get_mean_fields(i)[1], n, g, 100) for i in range(K)])

    a = channels(777)
    a.step(1000)
    a.step_and_sample(10000)
    eta_c1 = etas(a)
    a.reset_sampling()
    a.step(50000)
    a.step_and_sample(40000)
    eta_dev = etas(a)
    dens = np.mean([a.get_mean_fields(i)[0][:, :, 1:-1].mean() for i in range(K)])
    b = channels(1777)
    b.step(1000)
    b.step_and_sample(100000)
    eta_readme = etas(b)
    prof = np.mean([b.get_mean_fields(i)[1][..., 0].astype(np.float64).mean(axis=(0, 1)) for i in range(K)], axis=0)
    ref_c1, _ = _reference("c1")
    ref_readme, ref_prof = _reference("readme")
    print(f"eta over {K} seeds: 1000+10000 {eta_c1.mean():.2f} +- {eta_c1.std(ddof=1):.2f} (reference CPU {ref_c1:.2f}); "
          f"1000+100000 {eta_readme.mean():.3f} +- {eta_readme.std(ddof=1):.3f} (reference CPU {ref_readme:.3f}); "
          f"developed {eta_dev.mean():.3f} +- {eta_dev.std(ddof=1):.3f}; theory {theory:.3f}; README 37.27")
    assert abs(eta_c1.mean() - ref_c1) / ref_c1 < 0.05
    assert abs(eta_readme.mean() - ref_readme) / ref_readme < 0.02
    assert np.abs(prof[1:-1] - ref_prof[1:-1]).max() < 0.02 * ref_prof.max()
    assert abs(eta_dev.mean() - theory) / theory < 0.02
    assert abs(eta_dev.mean() - 37.27) / 37.27 < 0.02
    assert abs(dens - n) < 0.05


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
