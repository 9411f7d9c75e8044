"""Parity of the HIP step with the CPU oracle (same state, same counter-based draws).

(a) cell index per particle, per-cell counts and the stable cell-ordered
    permutation: bit-exact against the oracle's translate_particles table
    (simulation_functions.cpp:84-94).
(b) per-particle positions and velocities after one step: expected bit-exact
    (same IEEE operations in the same order, no contraction; the oracle with
    the counter stream sums each cell in the GPU's lane-group order, which is
    pinned against the reference's serial order in test_oracle_pinning.py);
    the asserted bound is the norm-wise relative 1e-5 of the specification.
Also: device Philox known answers and float transforms bitwise against the
oracle, seeded initial state, sampling accumulators against the oracle's FP64
accumulation.
"""
from __future__ import annotations

import numpy as np
import pytest

from cudampcd_amd import capi
from cudampcd_amd.capi import Handle
from helpers import make_params, min_image, stored_volume, vel_error
from oracle_lib import Oracle, oracle_lib

pytestmark = pytest.mark.gpu

CONFIGS = [
    # fluid extents, periodicity, n, algorithm
    ((6, 5, 8), (1, 1, 0), 10, "srd"),
    ((6, 5, 9), (1, 1, 1), 10, "srd"),
    ((5, 4, 7), (1, 1, 0), 3, "srd"),
    ((4, 4, 6), (1, 1, 0), 1, "srd"),
    ((7, 5, 9), (1, 1, 0), 10, "srd"),
    ((32, 24, 18), (1, 1, 0), 20, "srd"),
    ((6, 5, 8), (1, 1, 0), 10, "extended"),
    ((6, 5, 9), (1, 1, 1), 10, "extended"),
    ((5, 4, 7), (1, 1, 0), 3, "extended"),
    ((4, 4, 6), (1, 1, 0), 1, "extended"),
    ((20, 16, 12), (1, 1, 0), 10, "extended"),
    # lane groups of four and eight lanes per cell in the per-cell sums
    # (n = 13 and 30; wall cells add n ghosts to the members)
    ((7, 6, 8), (1, 1, 0), 13, "extended"),
    ((8, 6, 8), (1, 1, 0), 30, "srd"),
    ((6, 5, 8), (1, 1, 0), 30, "extended"),
    # dense walled cells: up to 100 ghosts a wall cell (the SRD rule's one
    # velocity-sum draw scaled by sqrt(100)), 640- and 1152-particle tiles
    ((4, 4, 5), (1, 1, 0), 100, "srd"),
    ((4, 4, 5), (1, 1, 0), 100, "extended"),
    # longest axis above 128 cells: the internal key cuts it into 128-row
    # blocks (the last one shorter), so particles cross block edges, whose
    # keys fall outside the permutation's and histogram's windows
    ((6, 4, 300), (1, 1, 0), 5, "srd"),
    ((300, 5, 6), (1, 1, 1), 4, "srd"),
    ((5, 4, 260), (1, 1, 0), 5, "extended"),
    # the smallest grids: one periodic cell (every particle in it, the shift
    # wraps within one cell length) and walled channels one cell wide
    ((1, 1, 1), (1, 1, 1), 6, "srd"),
    ((2, 1, 3), (1, 1, 0), 4, "srd"),
    ((1, 1, 1), (1, 1, 1), 6, "extended"),
    ((1, 2, 3), (1, 1, 0), 4, "extended"),
]
IDS = [f"{a}-{'x'.join(map(str, f))}-p{''.join(map(str, p))}-n{n}" for f, p, n, a in CONFIGS]


def _sim(pympcd, cfg, seed=4242, drag=0.001):
    fluid, perio, n, algo = cfg
    p = make_params(pympcd, fluid, perio, n=n, algorithm=algo, seed=seed, drag=drag)
    sim = pympcd.Simulation(p, "hip")
    return p, sim, Handle(sim)


def _oracle(p, distribute=False):
    return Oracle(stored_volume(p), tuple(p.periodicity), n=p.n, drag=p.drag, delta_t=p.delta_t,
                  thermal_velocity=float(np.sqrt(p.temperature)), rotation_deg=p.rotation_angle,
                  algorithm=p.algorithm, rng="counter", seed=(p.seed, 0), distribute=distribute,
                  experiment=p.experiment)


def test_device_philox_known_answers():
    ctr = np.array([[0, 0, 0, 0], [0xffffffff] * 4, [0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344]], np.uint32)
    keys = [(0, 0), (0xffffffff, 0xffffffff), (0xa4093822, 0x299f31d0)]
    want = [(0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd),
            (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)]
    for (k0, k1), c, w in zip(keys, ctr, want):
        assert tuple(capi.selftest_philox(k0, k1, c[None])[0]) == w
    # a bulk of random counters against the oracle's Philox
    import ctypes
    rng = np.random.default_rng(0)
    many = rng.integers(0, 2 ** 32, size=(4096, 4), dtype=np.uint64).astype(np.uint32)
    dev = capi.selftest_philox(0x1234, 0x5678, many)
    lib = oracle_lib()
    U = ctypes.c_uint * 4
    for i in range(0, 4096, 97):
        out = U()
        lib.orc_philox(0x1234, 0x5678, U(*map(int, many[i])), out)
        assert tuple(out) == tuple(dev[i])


def test_device_transforms_bitwise():
    import ctypes
    lib = oracle_lib()
    rng = np.random.default_rng(1)
    u = rng.uniform(0, 1, 20000).astype(np.float32) + np.float32(1.17549435e-38)
    got = capi.selftest_math(u, 0)[:, 0]
    want = np.array([lib.orc_logf(float(x)) for x in u], np.float32)
    assert np.array_equal(got, want)
    t = rng.uniform(0, 2, 20000).astype(np.float32)
    got = capi.selftest_math(t, 1)
    s = ctypes.c_float()
    c = ctypes.c_float()
    want = []
    for x in t:
        lib.orc_sincospi(float(x), ctypes.byref(s), ctypes.byref(c))
        want.append((s.value, c.value))
    assert np.array_equal(got, np.array(want, np.float32))
    x = rng.uniform(-90, 90, 20000).astype(np.float32)
    got = capi.selftest_math(x, 2)[:, 0]
    want = np.array([lib.orc_expf(float(v)) for v in x], np.float32)
    assert np.array_equal(got, want)
    y = rng.uniform(0, 4, 20000).astype(np.float32)
    assert np.array_equal(capi.selftest_math(y, 3)[:, 0], np.sqrt(y))
    w = rng.uniform(-12, 12, 20000).astype(np.float32)
    got = capi.selftest_math(w, 4)
    L = np.float32(7.0)
    assert np.array_equal(got[:, 0], np.fmod(w, L))
    assert np.array_equal(got[:, 1], np.fmod(w + np.float32(1.5) * L, L) - L * np.float32(0.5))
    got = capi.selftest_math(y + np.float32(0.5), 5)
    assert np.array_equal(got[:, 0], np.float32(1) / (y + np.float32(0.5)))
    assert np.array_equal(got[:, 1], (y + np.float32(0.5)) / np.float32(3))


@pytest.mark.parametrize("cfg", CONFIGS[:6], ids=IDS[:6])
def test_initial_state_matches_oracle(pympcd, cfg):
    p, sim, h = _sim(pympcd, cfg)
    o = _oracle(p, distribute=True)
    pos, vel = o.particles()
    assert np.array_equal(sim.get_particle_positions(), pos)
    assert np.array_equal(sim.get_particle_velocities(), vel)


@pytest.mark.parametrize("cfg", CONFIGS, ids=IDS)
def test_cell_keys_counts_permutation_bitexact(pympcd, cfg):
    p, sim, h = _sim(pympcd, cfg)
    sim.step(3)                      # a sorted, collided state
    pos, vel = sim.get_particle_positions(), sim.get_particle_velocities()
    o = _oracle(p)
    o.set_particles(pos, vel)
    o.set_step(h.step_index)
    cell, cnt, perm = h.peek_sort()
    o.translate()
    ocell, ocnt, olist = o.cells()
    assert np.array_equal(cell, ocell)
    assert np.array_equal(cnt, ocnt)
    cap = 4 * p.n
    starts = np.concatenate([[0], np.cumsum(cnt.astype(np.int64))[:-1]]).astype(np.int64)
    kept = np.concatenate([perm[s:s + min(int(c), cap)] for s, c in zip(starts, cnt)])
    assert np.array_equal(kept, olist)
    # the permutation is a permutation
    assert np.array_equal(np.sort(perm), np.arange(len(perm), dtype=np.uint32))


@pytest.mark.parametrize("cfg", CONFIGS, ids=IDS)
def test_step_matches_oracle(pympcd, cfg):
    p, sim, h = _sim(pympcd, cfg)
    sim.step(2)
    L = np.array(stored_volume(p), np.float64)
    h.record_permutation(True)
    bitwise = 0
    total = 0
    for _ in range(4):
        pos, vel = sim.get_particle_positions(), sim.get_particle_velocities()
        o = _oracle(p)
        o.set_particles(pos, vel)
        o.set_step(h.step_index)
        sim.step(1)
        o.step(1)
        perm = h.last_permutation()
        gpos, gvel = sim.get_particle_positions(), sim.get_particle_velocities()
        opos, ovel = o.particles()
        opos, ovel = opos[perm], ovel[perm]
        assert vel_error(gvel, ovel) <= 1e-5
        dpos = np.abs(min_image(gpos.astype(np.float64) - opos, L)).max()
        assert dpos <= 1e-5 * L.max()
        bitwise += int((np.all(gpos == opos, axis=1) & np.all(gvel == ovel, axis=1)).sum())
        total += len(gpos)
    print(f"bitwise-identical particles: {bitwise}/{total}")
    assert bitwise == total


# the small grid fits one tile; the long ones span many tiles and several
# 128-row blocks of the cell key, so batches accumulate their samples in LDS
# windows and send cells outside them (block edges, periodic wrap) to memory
SAMPLING_CFGS = [((6, 5, 8), (1, 1, 0), 10, "srd"), ((6, 5, 8), (1, 1, 0), 10, "extended"),
                 ((6, 4, 300), (1, 1, 0), 5, "srd"), ((300, 5, 6), (1, 1, 1), 4, "srd"),
                 ((5, 4, 260), (1, 1, 0), 5, "extended"),
                 # four and eight lanes per cell in the per-cell sums
                 ((7, 6, 8), (1, 1, 0), 20, "srd"), ((6, 5, 8), (1, 1, 0), 30, "extended")]


@pytest.mark.parametrize("cfg", SAMPLING_CFGS, ids=lambda c: f"{c[3]}-{'x'.join(map(str, c[0]))}")
def test_sampling_matches_oracle(pympcd, cfg):
    p, sim, h = _sim(pympcd, cfg, drag=0.01)
    sim.step(2)
    o = _oracle(p)
    for _ in range(7):
        o.set_particles(sim.get_particle_positions(), sim.get_particle_velocities())
        o.set_step(h.step_index)
        sim.step_and_sample(1)
        o.step_and_sample(1)
    X, Y, Z = stored_volume(p)
    d64, v64 = sim.get_mean_fields_f64()
    od, ov = o.sample_mean_f64()
    od = od.reshape(Z, Y, X).transpose(2, 1, 0)
    ov = ov.reshape(Z, Y, X, 3).transpose(2, 1, 0, 3)
    assert np.array_equal(d64, od)
    assert np.abs(v64 - ov).max() < 1e-9
    d32, v32 = sim.get_mean_fields()
    fd, fv = o.sample_mean()
    fd = fd.reshape(Z, Y, X).transpose(2, 1, 0)
    fv = fv.reshape(Z, Y, X, 3).transpose(2, 1, 0, 3)
    assert np.abs(d32 - fd).max() <= 1e-5 * max(1.0, np.abs(fd).max())
    assert np.abs(v32 - fv).max() <= 1e-5
    # get_mean_fields ended the average: the next sample starts a new one
    sim.step_and_sample(2)
    d, v = sim.get_mean_fields()
    assert np.isfinite(v).all()


def test_determinism_and_tiling_independence(pympcd):
    cfg = ((24, 20, 14), (1, 1, 0), 10, "srd")
    states = []
    # (32, 256): every tile spans two or more batches; 384/448: the 32-cell
    # tiles of five-wave workgroups
    for tiling in [(0, 0), (0, 0), (7, 512), (64, 1024), (1, 256), (10, 256), (64, 2048), (32, 256),
                   (20, 448), (32, 448), (30, 384)]:
        p, sim, h = _sim(pympcd, cfg, seed=99)
        if tiling != (0, 0):
            h.set_tiling(*tiling)
        sim.step(8)
        states.append((sim.get_particle_positions(), sim.get_particle_velocities()))
    for s in states[1:]:
        assert np.array_equal(s[0], states[0][0]) and np.array_equal(s[1], states[0][1])


def test_extended_tiling_independence_and_ghost_bound(pympcd):
    """Extended rule, 5 cell layers (2 of them wall layers), n = 40: a 32-cell
    tile holds about 13 wall cells (520 ghosts), more than the 100-slot ghost
    region of a 512 tile (room for two wall cells), so batches are cut by the
    ghost bound as well as by the particle capacity.  Every tiling gives the same bits, and one step
    with the smallest tiles matches the oracle bitwise."""
    cfg = ((12, 10, 3), (1, 1, 0), 40, "extended")
    states = []
    for tiling in [(0, 0), (32, 512), (8, 512), (64, 1152), (32, 2048)]:
        p, sim, h = _sim(pympcd, cfg, seed=77)
        if tiling != (0, 0):
            h.set_tiling(*tiling)
        sim.step(6)
        states.append((sim.get_particle_positions(), sim.get_particle_velocities()))
    for st in states[1:]:
        assert np.array_equal(st[0], states[0][0]) and np.array_equal(st[1], states[0][1])
    p, sim, h = _sim(pympcd, cfg, seed=77)
    h.set_tiling(32, 512)
    sim.step(2)
    h.record_permutation(True)
    pos, vel = sim.get_particle_positions(), sim.get_particle_velocities()
    o = _oracle(p)
    o.set_particles(pos, vel)
    o.set_step(h.step_index)
    sim.step(1)
    o.step(1)
    perm = h.last_permutation()
    opos, ovel = o.particles()
    assert np.array_equal(sim.get_particle_positions(), opos[perm].astype(np.float32))
    assert np.array_equal(sim.get_particle_velocities(), ovel[perm].astype(np.float32))


def test_set_get_roundtrip_and_errors(pympcd):
    p, sim, h = _sim(pympcd, ((5, 5, 6), (1, 1, 0), 4, "srd"))
    rng = np.random.default_rng(3)
    pos = sim.get_particle_positions()
    vel = rng.normal(size=pos.shape).astype(np.float32)
    sim.set_particle_velocities(vel)
    assert np.array_equal(sim.get_particle_velocities(), vel)
    pos2 = (pos * np.float32(0.5)).astype(np.float32)
    sim.set_particles(pos2, vel)
    assert np.array_equal(sim.get_particle_positions(), pos2)
    with pytest.raises(RuntimeError, match="particles"):
        sim.set_particle_positions(pos[:-1])
    sim.step(2)
    assert np.isfinite(sim.get_particle_positions()).all()


@pytest.mark.parametrize("algorithm", ["srd", "extended"])
def test_cell_beyond_largest_tile_matches_oracle(pympcd, algorithm):
    """2500 particles at one point (one cell, more than the largest LDS tile,
    2048): the cell's source slots are sorted inside the collide kernel, its
    first 4n collide and the rest are moved through un-shifted, as the
    oracle's table does.  A plain step from the initial state, then two
    sampled steps from the resulting states (the big cell now in cell order,
    mostly not collided): states and the mean fields equal to the oracle's."""
    p, sim, h = _sim(pympcd, ((10, 10, 10), (1, 1, 0), 5, algorithm))
    assert sim.n_particles > 2500
    pos, vel = sim.get_particle_positions().copy(), sim.get_particle_velocities().copy()
    pos[:2500] = (0.3, -0.6, -3.6)      # next to the lower wall: ghosts join for some shifts
    vel[:2500] = (0.4, -0.1, 0.2)
    sim.set_particles(pos, vel)
    h.record_permutation(True)
    o = _oracle(p)
    for sampled in (False, True, True):
        o.set_particles(sim.get_particle_positions(), sim.get_particle_velocities())
        o.set_step(h.step_index)
        if sampled:
            sim.step_and_sample(1)
            o.step_and_sample(1)
        else:
            sim.step(1)
            o.step(1)
        assert h.tiling()[1] <= 2048
        perm = h.last_permutation()
        assert np.array_equal(np.sort(perm), np.arange(sim.n_particles))
        gpos, gvel = sim.get_particle_positions(), sim.get_particle_velocities()
        opos, ovel = o.particles()
        assert np.array_equal(gpos, opos[perm]) and np.array_equal(gvel, ovel[perm])
    X, Y, Z = stored_volume(p)
    d64, v64 = sim.get_mean_fields_f64()
    od, ov = o.sample_mean_f64()
    assert d64.max() >= 2000
    assert np.array_equal(d64, od.reshape(Z, Y, X).transpose(2, 1, 0))
    assert np.abs(v64 - ov.reshape(Z, Y, X, 3).transpose(2, 1, 0, 3)).max() < 1e-9


@pytest.mark.parametrize("algorithm", ["srd", "extended"])
def test_full_cell_raises_capacity_and_matches_oracle(pympcd, algorithm):
    """A cell fuller than the default LDS batch (600 particles at one point,
    one velocity, so they share a cell whatever the grid shift): the batch
    capacity is raised to fit it and the step matches the oracle, whose
    table takes the first 4n particles of the cell and un-shifts the rest."""
    p, sim, h = _sim(pympcd, ((6, 5, 8), (1, 1, 0), 10, algorithm))
    cap0 = h.tiling()[1]
    pos, vel = sim.get_particle_positions().copy(), sim.get_particle_velocities().copy()
    pos[:600] = (0.3, -0.6, -3.6)      # the fluid layer next to the lower wall: ghosts join for some shifts
    vel[:600] = (0.4, -0.1, 0.2)
    sim.set_particles(pos, vel)
    h.record_permutation(True)
    o = _oracle(p)
    o.set_particles(pos, vel)
    o.set_step(h.step_index)
    sim.step(1)
    o.step(1)
    assert h.tiling()[1] > cap0 and h.tiling()[1] >= 600
    perm = h.last_permutation()
    gpos, gvel = sim.get_particle_positions(), sim.get_particle_velocities()
    opos, ovel = o.particles()
    assert np.array_equal(gpos, opos[perm]) and np.array_equal(gvel, ovel[perm])


@pytest.mark.parametrize("algorithm", ["srd", "extended"])
def test_one_call_equals_single_steps(pympcd, algorithm):
    """Inside one step() call the intermediate states are kept streamed (the next
    gather skips re-streaming); the result must equal single-step calls bit for bit."""
    cfg = ((14, 12, 10), (1, 1, 0), 10, algorithm)
    _, a, _ = _sim(pympcd, cfg, seed=8)
    _, b, _ = _sim(pympcd, cfg, seed=8)
    a.step(9)
    a.step_and_sample(5)
    for _ in range(9):
        b.step(1)
    for _ in range(5):
        b.step_and_sample(1)
    assert np.array_equal(a.get_particle_positions(), b.get_particle_positions())
    assert np.array_equal(a.get_particle_velocities(), b.get_particle_velocities())
    da, va = a.get_mean_fields_f64()
    db, vb = b.get_mean_fields_f64()
    assert np.array_equal(da, db) and np.array_equal(va, vb)


CHANNELS = [
    # fluid extents, periodicity, n, algorithm: cylinder about x, radius (X-2)/2 at
    # initialisation and (Z-2)/2 in streaming (simulation_functions.cpp:11-24, 45)
    ((10, 8, 8), (1, 1, 1), 8, "srd"),
    ((8, 10, 10), (1, 1, 1), 8, "srd"),
    ((10, 8, 8), (1, 1, 1), 8, "extended"),
    ((12, 9, 7), (1, 1, 0), 6, "srd"),
    ((10, 8, 8), (1, 1, 1), 20, "extended"),   # four lanes per cell
]
CIDS = [f"{a}-{'x'.join(map(str, f))}-p{''.join(map(str, p))}" for f, p, n, a in CHANNELS]


@pytest.mark.parametrize("cfg", CHANNELS, ids=CIDS)
def test_channel_matches_oracle(pympcd, cfg):
    fluid, perio, n, algo = cfg
    p = make_params(pympcd, fluid, perio, n=n, algorithm=algo, seed=606, drag=0.002, experiment="channel")
    sim = pympcd.Simulation(p, "hip")
    h = Handle(sim)
    o = _oracle(p, distribute=True)
    pos, vel = o.particles()
    assert np.array_equal(sim.get_particle_positions(), pos)
    assert np.array_equal(sim.get_particle_velocities(), vel)
    sim.step(3)
    h.record_permutation(True)
    for _ in range(3):
        o = _oracle(p)
        o.set_particles(sim.get_particle_positions(), sim.get_particle_velocities())
        o.set_step(h.step_index)
        sim.step(1)
        o.step(1)
        perm = h.last_permutation()
        opos, ovel = o.particles()
        assert np.array_equal(sim.get_particle_positions(), opos[perm])
        assert np.array_equal(sim.get_particle_velocities(), ovel[perm])
