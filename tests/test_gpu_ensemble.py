"""Ensembles (independent systems advanced together) and captured steps.

Every member of an ensemble must reproduce, bit for bit, a Simulation created
from the same parameters and seed (the counter-based random stream is keyed by
(seed, system id) and the cell-local order is stable, so neither the other
members nor the launch shape can change a result).  Steps replayed from a
captured HIP graph must equal eagerly launched steps, replay after replay,
with eager work in between; the autotuner may not change results.
"""
from __future__ import annotations

import numpy as np
import pytest

from cudampcd_amd.capi import Handle
from helpers import make_params

pytestmark = pytest.mark.gpu

MEMBERS = [
    # fluid extents, periodicity, n, drag, seed, system id, rotation angle
    ((8, 6, 10), (1, 1, 0), 10, 0.001, 11, 0, 120.0),
    ((5, 7, 6), (1, 1, 1), 5, 0.0, 12, 3, 120.0),
    ((12, 12, 20), (1, 1, 0), 20, 0.002, 13, 1, 90.0),
    ((4, 4, 3), (1, 1, 0), 1, 0.0005, 14, 7, 120.0),
    ((16, 8, 9), (1, 1, 0), 3, 0.0015, 15, 2, 130.0),
]


def _params(pympcd, m, algorithm):
    fluid, perio, n, drag, seed, sid, angle = m
    p = make_params(pympcd, fluid, perio, n=n, algorithm=algorithm, seed=seed, drag=drag, angle=angle)
    p.system_id = sid
    return p


@pytest.mark.parametrize("algorithm", ["srd", "extended"])
def test_members_match_separate_simulations_bitwise(pympcd, algorithm):
    params = [_params(pympcd, m, algorithm) for m in MEMBERS]
    ens = pympcd.Ensemble(params)
    sims = [pympcd.Simulation(p, "hip") for p in params]
    assert ens.n_systems == len(params)
    for k in range(len(params)):
        assert ens.n_particles(k) == sims[k].n_particles
        assert np.array_equal(ens.get_particle_positions(k), sims[k].get_particle_positions())
    ens.step(7)
    ens.step_and_sample(9)
    for s in sims:
        s.step(7)
        s.step_and_sample(9)
    for k, s in enumerate(sims):
        assert np.array_equal(ens.get_particle_positions(k), s.get_particle_positions()), k
        assert np.array_equal(ens.get_particle_velocities(k), s.get_particle_velocities()), k
        d, v = ens.get_mean_fields_f64(k)
        sd, sv = s.get_mean_fields_f64()
        assert np.array_equal(d, sd) and np.array_equal(v, sv), k
        d32, v32 = ens.get_mean_fields(k)
        s32, sv32 = s.get_mean_fields()
        assert d32.shape == s32.shape and np.array_equal(d32, s32) and np.array_equal(v32, sv32), k
    # the flattened C-ABI view places system k at its slot range
    h = Handle(ens)
    assert h.n_systems == len(params)
    first, count, cfirst, ccount = h.system_info(2)
    assert count == sims[2].n_particles and first == sum(s.n_particles for s in sims[:2])


def test_graph_replay_equals_eager_steps(pympcd):
    cfg = ((20, 16, 14), (1, 1, 0))
    a = pympcd.Simulation(make_params(pympcd, *cfg, n=10, seed=77), "hip")
    b = pympcd.Simulation(make_params(pympcd, *cfg, n=10, seed=77), "hip")
    Handle(a).use_graph(False)
    for rep in range(5):
        a.step(3)
        b.step(3)              # graph replays
        assert np.array_equal(a.get_particle_positions(), b.get_particle_positions()), rep
        assert np.array_equal(a.get_particle_velocities(), b.get_particle_velocities()), rep
        a.step_and_sample(2)   # first sample eager, second from the sampling graph
        b.step_and_sample(2)
        da, va = a.get_mean_fields_f64()
        db, vb = b.get_mean_fields_f64()
        assert np.array_equal(da, db) and np.array_equal(va, vb), rep
    assert Handle(a).step_index == Handle(b).step_index == 25


def test_autotune_keeps_results(pympcd):
    params = [_params(pympcd, m, "srd") for m in MEMBERS[:3]]
    tuned = pympcd.Ensemble(params)
    plain = pympcd.Ensemble(params)
    tc, cap, ms = tuned.autotune(2)
    # any compiled LDS capacity can win the timing (include/mpcd_hip.h)
    assert cap in (256, 384, 448, 512, 640, 1152, 2048) and 1 <= tc <= 64 and ms > 0
    steps = Handle(tuned).step_index
    plain.step(steps)
    tuned.step(3)
    plain.step(3)
    for k in range(len(params)):
        assert np.array_equal(tuned.get_particle_positions(k), plain.get_particle_positions(k))
        assert np.array_equal(tuned.get_particle_velocities(k), plain.get_particle_velocities(k))


def test_ensemble_member_set_and_reset(pympcd):
    params = [_params(pympcd, m, "srd") for m in MEMBERS[:2]]
    ens = pympcd.Ensemble(params)
    ens.step_and_sample(3)
    d1, _ = ens.get_mean_fields_f64(0)
    ens.step_and_sample(2)
    d2, _ = ens.get_mean_fields_f64(0)
    assert d1.sum() > 0 and d2.sum() > 0
    ens.reset_sampling()
    ens.step_and_sample(3)
    d3, v3 = ens.get_mean_fields(1)
    assert np.isfinite(v3).all()
    with pytest.raises(RuntimeError, match="out of range"):
        ens.get_particle_positions(5)


def _member_oracle(p, system_id):
    from helpers import stored_volume
    from oracle_lib import Oracle

    return Oracle(stored_volume(p), tuple(p.periodicity), n=p.n, drag=p.drag, delta_t=p.delta_t,
                  thermal_velocity=float(np.sqrt(p.temperature)), rotation_deg=p.rotation_angle,
                  algorithm=p.algorithm, rng="counter", seed=(p.seed, 0), system=system_id, distribute=False)


def _step_members_against_oracle(ens, params, members, steps=2):
    """One ensemble step at a time; each listed member's state after the step
    equals the oracle's step from that member's state before it, keyed by the
    member's (seed, system id), bit for bit.  The recorded permutation of the
    flattened arrays maps each member's output slots onto its own input slots."""
    h = Handle(ens)
    h.record_permutation(True)
    for _ in range(steps):
        before = {k: (ens.get_particle_positions(k), ens.get_particle_velocities(k)) for k in members}
        step = h.step_index
        ens.step(1)
        perm = h.last_permutation()
        for k in members:
            first, count, _, _ = h.system_info(k)
            sub = perm[first:first + count].astype(np.int64) - first
            assert sub.min() >= 0 and sub.max() < count, k
            o = _member_oracle(params[k], params[k].system_id)
            o.set_particles(*before[k])
            o.set_step(step)
            o.step(1)
            opos, ovel = o.particles()
            assert np.array_equal(ens.get_particle_positions(k), opos[sub]), k
            assert np.array_equal(ens.get_particle_velocities(k), ovel[sub]), k


@pytest.mark.parametrize("algorithm", ["srd", "extended"])
def test_members_match_the_oracle_bitwise(pympcd, algorithm):
    """Every member of a mixed ensemble (different grids, densities, walls,
    drags, rotation angles, seeds and system ids) against the oracle."""
    params = [_params(pympcd, m, algorithm) for m in MEMBERS]
    ens = pympcd.Ensemble(params)
    ens.step(3)
    _step_members_against_oracle(ens, params, range(len(params)))


def test_c4_channel_set(pympcd):
    """The ensemble workload itself: 64 seeded SRD channels, cross-sections
    32^2 / 64^2 / 128^2, wall gaps 20-200, n in {5, 10, 20}, drags 5e-4 to 2e-3
    (about 5e8 particles).  Three steps, then one member of each density
    stepped against the oracle, and three members of different density and
    cross-section against standalone Simulations of the same parameters."""
    from bench import c4_members, make_params as bench_params

    spec = c4_members(0x5EED0004)       # the seed bench.py --config c4 uses
    params = [bench_params(pympcd, m["fluid"], m["n"], "srd", m["drag"], m["seed"], m["system_id"]) for m in spec]
    ens = pympcd.Ensemble(params)
    assert ens.n_systems == 64
    ens.step(3)
    # the smallest member of each density for the oracle (a few 1e5 particles)
    by_n = {}
    for k, m in enumerate(spec):
        size = m["fluid"][0] * m["fluid"][1] * m["fluid"][2] * m["n"]
        if m["n"] not in by_n or size < by_n[m["n"]][1]:
            by_n[m["n"]] = (k, size)
    oracle_members = sorted(k for k, _ in by_n.values())
    assert len(oracle_members) == 3
    _step_members_against_oracle(ens, params, oracle_members, steps=1)
    # standalone runs of three members of different density and cross-section,
    # the smallest such ones
    size = [m["fluid"][0] * m["fluid"][1] * m["fluid"][2] * m["n"] for m in spec]
    picks = []
    for k in sorted(range(len(spec)), key=lambda j: size[j]):
        m = spec[k]
        if all(spec[j]["n"] != m["n"] and spec[j]["fluid"][0] != m["fluid"][0] for j in picks):
            picks.append(k)
    assert len(picks) == 3
    steps = Handle(ens).step_index
    for k in picks:
        s = pympcd.Simulation(params[k], "hip")
        s.step(steps)
        assert np.array_equal(ens.get_particle_positions(k), s.get_particle_positions()), k
        assert np.array_equal(ens.get_particle_velocities(k), s.get_particle_velocities()), k
        del s
