"""Pins the CPU oracle before it is trusted as the GPU's checker.

1. Against the reference CPU backend itself, compiled in place into
   oracle/_ref (cpu_backend.cpp + simulation_functions.cpp): with the
   reference's Xoshiro128+ stream and seed, the oracle reproduces the
   reference trajectory bit for bit (initial distribution, grid shifts, cell
   indices, counters, cell lists, post-collision state, sample means).
2. The counter-based stream: Philox4x32-10 known-answer vectors (Random123),
   accuracy of the float transforms against float64 numpy, and the
   statistical tests of the reference (tests/statistics/test_statistics.py:6-48).
3. Closed forms and invariants that hold for any draws: the reference's
   translate known answers (tests/unit/cpp/test_simulation_functions.cpp:4-43),
   the SRD rotation matrix about the drawn axis, momentum and energy
   conservation, the 3x3 inverse against numpy, the 13 extended directions.
4. The GPU's per-cell summation order (lane groups combined pairwise, the
   order the GPU tests compare with bit for bit) against the reference's
   serial order, pinned in 1.: the same step within 1e-5, momentum and SRD
   energy conserved (3. runs in that order).
"""
from __future__ import annotations

import math

import numpy as np
import pytest
import scipy.stats as stats

from oracle_lib import Oracle, Reference, oracle_lib, ref_available

needs_ref = pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference sources absent)")


# ---------------------------------------------------------------- 1. reference
@needs_ref
@pytest.mark.parametrize("algorithm", ["srd", "extended"])
@pytest.mark.parametrize("volume,periodic,n", [
    ((6, 5, 10), (1, 1, 0), 10),
    ((6, 5, 9), (1, 1, 1), 10),
    ((5, 4, 7), (1, 1, 0), 3),     # sparse: lone and empty cells
    ((4, 4, 6), (1, 1, 0), 1),     # cap 4n = 4: table overflow and dropped reals
])
def test_oracle_reproduces_reference_bitwise(algorithm, volume, periodic, n):
    kw = dict(n=n, drag=0.01, delta_t=0.02, algorithm=algorithm, seed=(0x1234567, 0x89abcdef))
    o = Oracle(volume, periodic, rng="reference", **kw)
    r = Reference(volume, periodic, **kw)
    p0, v0 = o.particles()
    q0, w0 = r.particles()
    assert np.array_equal(p0, q0) and np.array_equal(v0, w0)
    overflow_seen = False
    for _ in range(12):
        assert np.array_equal(o.translate(), r.translate())
        oc, ocnt, olist = o.cells()
        rc, rcnt, rlist = r.cells()
        assert np.array_equal(oc, rc)
        assert np.array_equal(ocnt, rcnt)
        assert np.array_equal(olist, rlist)
        overflow_seen |= bool((ocnt > 4 * n).any())
        o.collide()
        r.collide()
        p, v = o.particles()
        q, w = r.particles()
        assert np.array_equal(p, q)
        assert np.array_equal(v, w)
    if n == 1:
        assert overflow_seen


@needs_ref
@pytest.mark.parametrize("algorithm", ["srd", "extended"])
@pytest.mark.parametrize("volume,periodic", [
    ((10, 8, 8), (1, 1, 1)),    # initial radius from x (4) > streaming radius from z (3)
    ((8, 10, 10), (1, 1, 1)),   # initial radius 3 < streaming radius 4
    ((10, 8, 8), (1, 1, 0)),    # with z-wall layers (and their ghosts) as well
])
def test_oracle_channel_reproduces_reference_bitwise(algorithm, volume, periodic):
    """Channel geometry (simulation_functions.cpp:11-24, 45, 75-82): rejection
    sampling inside a cylinder about x, bounce back off its wall."""
    kw = dict(n=6, drag=0.01, delta_t=0.05, algorithm=algorithm, seed=(77, 88), experiment="channel")
    o = Oracle(volume, periodic, rng="reference", **kw)
    r = Reference(volume, periodic, **kw)
    p0, v0 = o.particles()
    q0, w0 = r.particles()
    assert np.array_equal(p0, q0) and np.array_equal(v0, w0)
    r2_init = np.float32((volume[0] - 2) ** 2 * 0.25)
    assert (p0[:, 2] * p0[:, 2] + p0[:, 1] * p0[:, 1] <= r2_init).all()
    for _ in range(10):
        assert np.array_equal(o.translate(), r.translate())
        assert np.array_equal(o.cells()[2], r.cells()[2])
        o.collide()
        r.collide()
        p, v = o.particles()
        q, w = r.particles()
        assert np.array_equal(p, q) and np.array_equal(v, w)


def test_counter_channel_initial_state_inside_cylinder():
    o = Oracle((12, 9, 9), (1, 1, 1), n=5, rng="counter", seed=(3, 0), experiment="channel")
    pos, _ = o.particles()
    r2 = np.float32((12 - 2) ** 2 * 0.25)
    assert (pos[:, 2] * pos[:, 2] + pos[:, 1] * pos[:, 1] <= r2).all()
    s = Oracle((12, 9, 9), (1, 1, 1), n=5, rng="counter", seed=(3, 0))
    spos, _ = s.particles()
    # particles whose first draw already lies inside keep it (attempt 0 = the standard draw)
    inside = spos[:, 2] * spos[:, 2] + spos[:, 1] * spos[:, 1] <= r2
    assert np.array_equal(pos[inside], spos[inside]) and (~inside).any()


@needs_ref
@pytest.mark.parametrize("algorithm", ["srd", "extended"])
def test_oracle_sampling_reproduces_reference(algorithm):
    kw = dict(n=10, drag=0.01, delta_t=0.02, algorithm=algorithm, seed=(5, 7))
    o = Oracle((6, 5, 10), (1, 1, 0), rng="reference", **kw)
    r = Reference((6, 5, 10), (1, 1, 0), **kw)
    o.step(3)
    r.step(3)
    for k in (1, 6, 3):      # the first sample only initialises (S-1 normalisation)
        o.step_and_sample(k)
        r.step_and_sample(k)
        d1, v1 = o.sample_mean()
        d2, v2 = r.sample_mean()
        assert np.array_equal(d1, d2, equal_nan=True)
        assert np.array_equal(v1, v2, equal_nan=True)


# ------------------------------------------------------- 2. counter stream
def test_philox_known_answers():
    lib = oracle_lib()
    U = __import__("ctypes").c_uint * 4
    kat = [
        ((0, 0), (0, 0, 0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff, 0xffffffff), (0xffffffff,) * 4, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0xa4093822, 0x299f31d0), (0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for key, ctr, want in kat:
        out = U()
        lib.orc_philox(key[0], key[1], U(*ctr), out)
        assert tuple(out) == want


def _ulp(x):
    x = np.abs(np.asarray(x, np.float32))
    return np.spacing(x).astype(np.float64)


def test_log_transform_accuracy():
    lib = oracle_lib()
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.uniform(0, 1, 4000), np.float32(2.0) ** rng.uniform(-126, 0, 2000),
                         np.array([1.0, 0.5, 2 ** -126, 1.41421356, 0.70710677])]).astype(np.float32)
    got = np.array([lib.orc_logf(float(x)) for x in xs], np.float64)
    want = np.log(xs.astype(np.float64))
    err = np.abs(got - want) / np.maximum(_ulp(want), 1e-45)
    assert err.max() <= 4.0, err.max()


def test_sincospi_accuracy():
    import ctypes
    lib = oracle_lib()
    ts = np.concatenate([np.linspace(0, 2, 5001), np.random.default_rng(2).uniform(0, 2, 3000)]).astype(np.float32)
    s = ctypes.c_float()
    c = ctypes.c_float()
    got = []
    for t in ts:
        lib.orc_sincospi(float(t), ctypes.byref(s), ctypes.byref(c))
        got.append((s.value, c.value))
    got = np.array(got)
    t64 = ts.astype(np.float64)
    assert np.abs(got[:, 0] - np.sin(np.pi * t64)).max() < 3e-7
    assert np.abs(got[:, 1] - np.cos(np.pi * t64)).max() < 3e-7


def test_exp_transform_accuracy_and_specials():
    lib = oracle_lib()
    xs = np.concatenate([np.linspace(-87, 88, 20001), np.random.default_rng(3).uniform(-5, 5, 5000)]).astype(np.float32)
    got = np.array([lib.orc_expf(float(x)) for x in xs], np.float64)
    want = np.exp(xs.astype(np.float64))
    rel = np.abs(got - want) / want
    assert rel.max() < 5e-7, rel.max()
    assert math.isnan(lib.orc_expf(float("nan")))
    assert lib.orc_expf(float("-inf")) == 0.0
    assert lib.orc_expf(float("inf")) == float("inf")
    assert lib.orc_expf(-200.0) == 0.0


def _counter_samples(purpose, n):
    o = Oracle((4, 4, 4), (1, 1, 1), n=1, rng="counter", seed=(12345, 0), distribute=False)
    per = 4096
    out = [o.draws(purpose, step, cell, per) for step in range(2) for cell in range((n // per) // 2 + 1)]
    return np.concatenate(out)[:n].astype(np.float64)


def test_counter_uniform_statistics():
    """Mirrors tests/statistics/test_statistics.py:6-28 of the reference."""
    samples = _counter_samples(2, 100000)
    assert 0.49 < samples.mean() < 0.51
    assert 0.08 < samples.var() < 0.09
    assert stats.kstest(samples, "uniform").pvalue > 0.01
    hist, _ = np.histogram(samples, bins=50, range=(0, 1))
    assert stats.chisquare(hist, len(samples) / 50).pvalue > 0.01


def test_counter_gaussian_statistics():
    """Mirrors tests/statistics/test_statistics.py:31-48 of the reference."""
    samples = _counter_samples(7, 100000)
    assert -0.01 < samples.mean() < 0.01
    assert 0.99 < samples.std() < 1.01
    subset = np.random.default_rng(4).choice(samples, 5000, replace=False)
    assert stats.shapiro(subset).pvalue > 0.01
    ad = stats.anderson(samples, "norm")
    assert ad.statistic < ad.critical_values[2]


def test_reference_xoshiro_statistics():
    """The reference-mode stream (pinned bitwise above) passes the same tests."""
    import ctypes
    lib = oracle_lib()
    out = np.empty(100000, np.float32)
    lib.orc_xoshiro_uniform(12345, 67890, len(out), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert 0.49 < out.mean() < 0.51 and stats.kstest(out.astype(np.float64), "uniform").pvalue > 0.01
    lib.orc_xoshiro_gaussian(12345, 67890, len(out), out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
    assert -0.01 < out.mean() < 0.01 and 0.99 < out.std() < 1.01


def test_counter_axes_are_uniform_unit_vectors():
    """Uniform on the sphere, as the reference's Marsaglia draw
    (random.hpp:237-253): each Cartesian component of a uniform unit vector is
    uniform on [-1, 1] (Archimedes), and the azimuth about any axis is uniform
    and independent of that component; a dropped square root, a wrong scale of
    the azimuth or a swapped component fails one of these."""
    o = Oracle((4, 4, 4), (1, 1, 1), n=1, rng="counter", seed=(99, 0), distribute=False)
    ax = np.concatenate([o.draws(5, s, 0, 5000) for s in range(4)]).astype(np.float64)
    assert np.abs(np.linalg.norm(ax, axis=1) - 1).max() < 1e-6
    assert np.abs(ax.mean(axis=0)).max() < 0.02
    # isotropy: <a_i a_j> = delta_ij / 3
    assert np.abs(ax.T @ ax / len(ax) - np.eye(3) / 3).max() < 0.02
    for k in range(3):
        assert stats.kstest(ax[:, k], stats.uniform(-1, 2).cdf).pvalue > 0.01
        i, j = [a for a in range(3) if a != k]
        phi = np.arctan2(ax[:, j], ax[:, i])
        assert stats.kstest(phi, stats.uniform(-np.pi, 2 * np.pi).cdf).pvalue > 0.01
        # independence of the azimuth and the component (|component| split at 1/2)
        lo, hi = phi[np.abs(ax[:, k]) < 0.5], phi[np.abs(ax[:, k]) >= 0.5]
        assert stats.ks_2samp(lo, hi).pvalue > 0.01


@pytest.mark.parametrize("added", [1, 3, 10, 20])
def test_counter_srd_ghost_sum_matches_the_reference_construction(added):
    """The SRD rule uses a wall cell's ghosts only through the sum of their
    velocities (simulation_functions.cpp:168-171; getCorrection returns the
    mean, mpc_cell.hpp:82-84).  The counter stream draws that sum at once;
    the reference draws `added` Maxwell-Boltzmann velocities at the thermal
    velocity and sums them.  Both must be N(0, added T) per component,
    independent: each component is compared with the reference stream's
    construction (two-sample KS) and with the closed form (KS against the
    normal of variance added T); a missing square root, a missing thermal
    velocity or correlated components fails."""
    T = 0.7
    kw = dict(n=20, thermal_velocity=math.sqrt(T), distribute=False)
    oc = Oracle((4, 4, 4), (1, 1, 0), rng="counter", seed=(5, 6), **kw)
    orf = Oracle((4, 4, 4), (1, 1, 0), rng="reference", seed=(5, 6), **kw)
    cnt = 6000
    c = np.concatenate([oc.srd_ghost_sums(s, 0, cnt // 3, added) for s in range(3)]).astype(np.float64)
    r = orf.srd_ghost_sums(0, 0, cnt, added).astype(np.float64)
    sd = math.sqrt(added * T)
    for k in range(3):
        assert stats.ks_2samp(c[:, k], r[:, k]).pvalue > 0.01
        assert stats.kstest(c[:, k], stats.norm(0, sd).cdf).pvalue > 0.01
    cov = np.cov(c.T) / (added * T)
    assert np.abs(cov - np.eye(3)).max() < 0.08


# ------------------------------------------------- 3. closed forms / invariants
def test_translate_known_answers():
    """tests/unit/cpp/test_simulation_functions.cpp:4-43 of the reference."""
    o = Oracle((10, 10, 10), (1, 1, 1), n=10, N=10, delta_t=1.0, rng="counter", distribute=False)
    o.force_shift([0, 0, 0])
    rng = np.random.default_rng(5)
    for _ in range(10):
        v = rng.normal(size=(10, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        o.set_particles(np.zeros((10, 3)), v)
        o.translate()
        pos, vel = o.particles()
        assert np.abs(pos - vel).max() <= 1e-6
    o.set_particles(np.full((10, 3), -5.0), np.zeros((10, 3)))
    o.translate()
    _, cnt, _ = o.cells()
    assert cnt[0] == 10


def _rotation(axis, alpha):
    """v' = v_par + cos a v_perp + sin a (v_perp x axis)  (simulation_functions.cpp:176-182)."""
    a = np.asarray(axis, np.float64)
    P = np.outer(a, a)
    K = np.array([[0, -a[2], a[1]], [a[2], 0, -a[0]], [-a[1], a[0], 0]])  # K v = a x v
    return P + math.cos(alpha) * (np.eye(3) - P) - math.sin(alpha) * K


def test_rotation_special_case():
    R = _rotation([0, 0, 1], math.radians(120))
    assert np.allclose(R @ [1, 0, 0], [math.cos(math.radians(120)), -math.sin(math.radians(120)), 0])


@pytest.mark.parametrize("angle", [120.0, 90.0, 47.0])
def test_srd_rotation_matches_closed_form(angle):
    """Per cell, v' - u = R(axis, alpha) (v - u) with the axis the stream draws."""
    vol = (5, 4, 6)
    o = Oracle(vol, (1, 1, 1), n=12, drag=0.0, rotation_deg=angle, rng="counter", seed=(77, 0))
    o.set_step(3)
    o.translate()
    cell, cnt, lst = o.cells()
    p0, v0 = o.particles()
    o.collide()
    p1, v1 = o.particles()
    axes = o.draws(5, 3, 0, o.C)
    checked = 0
    for c in range(o.C):
        idx = np.where(cell == c)[0]
        if len(idx) < 4 or len(idx) > 48:
            continue
        u = v0[idx].astype(np.float64).mean(axis=0)
        d0 = v0[idx] - u
        d1 = v1[idx] - u
        R = _rotation(axes[c], math.radians(angle))
        assert np.abs(d0 @ R.T - d1).max() < 2e-5
        checked += 1
    assert checked > o.C // 2


@pytest.mark.parametrize("algorithm", ["srd", "extended"])
def test_momentum_and_energy_conservation(algorithm):
    o = Oracle((6, 6, 6), (1, 1, 1), n=10, drag=0.0, algorithm=algorithm, rng="counter", seed=(3, 0))
    _, v = o.particles()
    P0 = v.astype(np.float64).sum(axis=0)
    E0 = (v.astype(np.float64) ** 2).sum()
    changed = 0
    for _ in range(20):
        _, vb = o.particles()
        o.step(1)
        _, v = o.particles()
        changed += int((np.abs(v - vb) > 1e-6).any(axis=1).sum())
    P1 = v.astype(np.float64).sum(axis=0)
    assert np.abs(P1 - P0).max() < 2e-2          # FP32 round-off over 2160 particles x 20 steps
    assert changed > 0
    if algorithm == "srd":
        E1 = (v.astype(np.float64) ** 2).sum()
        assert abs(E1 - E0) / E0 < 1e-4


def test_drag_adds_momentum():
    """Each particle of a colliding cell gains g dt in x (simulation_functions.cpp:121, 187)."""
    g, dt = 0.01, 0.02
    o = Oracle((6, 6, 6), (1, 1, 1), n=10, drag=g, delta_t=dt, rng="counter", seed=(8, 0))
    _, v0 = o.particles()
    o.step(1)
    _, v1 = o.particles()
    gain = (v1[:, 0].astype(np.float64) - v0[:, 0]).sum()
    assert abs(gain - o.N * np.float32(g * dt)) < 1e-3


def test_inverse_matches_numpy():
    import ctypes
    lib = oracle_lib()
    fp = ctypes.POINTER(ctypes.c_float)
    rng = np.random.default_rng(6)
    for _ in range(200):
        A = rng.normal(size=(3, 3))
        M = A @ A.T + 0.5 * np.eye(3)
        m6 = np.array([M[0, 0], M[1, 1], M[2, 2], M[0, 1], M[0, 2], M[1, 2]], np.float32)
        out = np.empty(6, np.float32)
        lib.orc_sym_inverse(m6.ctypes.data_as(fp), 1e-5, out.ctypes.data_as(fp))
        inv = np.linalg.inv(m6[[0, 3, 4, 3, 1, 5, 4, 5, 2]].reshape(3, 3).astype(np.float64))
        got = out[[0, 3, 4, 3, 1, 5, 4, 5, 2]].reshape(3, 3)
        assert np.abs(got - inv).max() / np.abs(inv).max() < 1e-4
    # collinear particles: det = 0, trace fallback 4 / tr^2 (mechanic.hpp:72-75)
    m6 = np.array([2.0, 2.0, 0.0, 0.0, 0.0, 0.0], np.float32)
    out = np.empty(6, np.float32)
    lib.orc_sym_inverse(m6.ctypes.data_as(fp), 1e-5, out.ctypes.data_as(fp))
    assert np.allclose(out, m6 * 4 / 16)
    lib.orc_sym_inverse(np.zeros(6, np.float32).ctypes.data_as(fp), 1e-5, out.ctypes.data_as(fp))
    assert not out.any()


def test_extended_directions():
    import ctypes
    lib = oracle_lib()
    out = np.empty(3, np.float32)
    dirs = []
    for sel in range(13):
        lib.orc_ext_direction(sel, out.ctypes.data_as(ctypes.POINTER(ctypes.c_float)))
        dirs.append(out.copy())
        phi_i, theta_i = sel % 4 + 1, sel // 4 + 1
        if theta_i % 2 and phi_i % 2:
            th = math.acos(1 / math.sqrt(3)) if theta_i == 1 else math.pi - math.acos(1 / math.sqrt(3))
        else:
            th = theta_i * math.pi / 4
        ph = phi_i * math.pi / 4
        want = [math.sin(th) * math.cos(ph), math.sin(th) * math.sin(ph), math.cos(th)]
        assert np.allclose(out, want, atol=1e-6)
    dirs = np.array(dirs)
    assert np.allclose(np.linalg.norm(dirs, axis=1), 1, atol=1e-6)


# ---- counter-mode draws that replace a rejection loop or a stateful draw ----

def _accepted_by_reference(z, shift, layer, sign):
    """Acceptance of the reference's ghost-z rejection loop
    (simulation_functions.cpp:155-161): it redraws while
    ((z < shift) xor layer) xor sign."""
    return ~(((z < np.float32(shift)) ^ bool(layer)) ^ bool(sign))


@pytest.mark.parametrize("shift", [0.03, 0.31, 0.5, 0.77, 0.98])
@pytest.mark.parametrize("layer,sign", [(0, 0), (0, 1), (1, 0), (1, 1)])
def test_counter_ghost_z_lies_in_the_accepted_half_cell_and_is_uniform(layer, sign, shift):
    """Counter mode maps one uniform onto the region the reference's rejection
    loop accepts ([shift, 1] when layer xor sign is 0, [0, shift) otherwise).
    Every draw must pass the reference's own acceptance predicate (a swapped
    layer/sign branch would put every draw in the rejected region), be uniform
    there, and follow the same distribution as the reference-mode stream, which
    runs the rejection loop itself (and is pinned bitwise to the compiled
    reference above).  x and y stay uniform on [0, 1]."""
    o = Oracle((4, 4, 4), (1, 1, 0), n=1, rng="counter", seed=(2024, 7), distribute=False)
    g = np.concatenate([o.ghost_positions(step, cell, 25, shift, layer, sign)
                        for step in range(4) for cell in range(100)]).astype(np.float64)
    z = g[:, 2].astype(np.float32)
    assert _accepted_by_reference(z, shift, layer, sign).all()
    lo, hi = (0.0, shift) if (layer ^ sign) else (shift, 1.0)
    assert stats.kstest(g[:, 2], "uniform", args=(lo, hi - lo)).pvalue > 1e-3
    for a in (0, 1):
        assert stats.kstest(g[:, a], "uniform").pvalue > 1e-3
    r = Oracle((4, 4, 4), (1, 1, 0), n=1, rng="reference", seed=(31337, 4242), distribute=False)
    gr = r.ghost_positions(0, 0, len(g), shift, layer, sign).astype(np.float64)
    assert _accepted_by_reference(gr[:, 2].astype(np.float32), shift, layer, sign).all()
    assert stats.ks_2samp(g[:, 2], gr[:, 2]).pvalue > 1e-3


def test_counter_extended_directions_uniform_over_26_and_independent():
    """Counter mode draws the extended rule's direction index as
    genUniformInt(0, 12) = (u32 / 2) % 13 and its sign as (u32 / 2) % 2 from two
    words of one Philox block (random.hpp:134-136, simulation_functions.cpp:276,
    293): the 26 signed directions are equally likely, index and sign are
    independent, the gate uniform (third word) is uniform and independent of
    the direction, and the joint table matches the reference-mode stream's."""
    o = Oracle((4, 4, 4), (1, 1, 1), n=1, rng="counter", seed=(99, 3), distribute=False)
    draws = [o.ext_axis_draws(step, 0, 20000) for step in range(13)]
    sel = np.concatenate([d[0] for d in draws])
    flip = np.concatenate([d[1] for d in draws])
    gate = np.concatenate([d[2] for d in draws]).astype(np.float64)
    assert sel.min() == 0 and sel.max() == 12 and set(np.unique(flip)) == {0, 1}
    table = np.zeros((13, 2))
    np.add.at(table, (sel, flip), 1)
    assert stats.chisquare(table.ravel()).pvalue > 1e-3
    assert stats.chi2_contingency(table).pvalue > 1e-3
    assert stats.kstest(gate, "uniform").pvalue > 1e-3
    # gate independent of the direction: its mean per direction index
    means = np.array([gate[sel == k].mean() for k in range(13)])
    assert np.abs(means - 0.5).max() < 5 * math.sqrt(1 / 12 / (len(gate) / 13))
    r = Oracle((4, 4, 4), (1, 1, 1), n=1, rng="reference", seed=(5, 6), distribute=False)
    rs, rf, _ = r.ext_axis_draws(0, 0, len(sel))
    rtable = np.zeros((13, 2))
    np.add.at(rtable, (rs, rf), 1)
    assert stats.chi2_contingency(np.stack([table.ravel(), rtable.ravel()])).pvalue > 1e-3


def test_oracle_poiseuille_viscosity_sanity():
    """Small developed channel on the oracle: a parabolic profile whose viscosity
    is within 8% of SRD theory.

    8x8x16 fluid, n=10, g=0.02, 2000 warm-up (about 6 relaxation times of the
    slowest mode) + 8000 sampled steps.  At this 16-cell gap the reference's
    algorithm (which the oracle reproduces bit for bit) reads about 6% below the
    bulk theory, from both the reference's max-velocity estimate and a
    slip-independent curvature fit; the 2% gate is applied on the GPU at the
    reference's 100-cell gap (README.md:80-102).
    """
    from cudampcd_amd.physics import srd_viscosity

    n, g, dt, L = 10, 0.02, 0.02, 16
    o = Oracle((8, 8, L + 2), (1, 1, 0), n=n, drag=g, delta_t=dt, rng="counter", seed=(2024, 0))
    o.step(2000)
    o.step_and_sample(8000)
    d, v = o.sample_mean_f64()
    vel = v.reshape(L + 2, 8, 8, 3).transpose(2, 1, 0, 3)
    prof = vel[..., 0].mean(axis=(0, 1))
    eta_max = L * L * n * g / (8 * prof.max())
    z = np.arange(L + 2) - (L + 2) / 2 + 0.5
    a2, a1, a0 = np.polyfit(z[2:L], prof[2:L], 2)
    eta_curv = -n * g / (2 * a2)
    theory = srd_viscosity(n, dt)
    walls = np.sort(np.roots([a2, a1, a0]).real)
    assert abs(eta_max - theory) / theory < 0.08, (eta_max, theory)
    assert abs(eta_curv - theory) / theory < 0.08, (eta_curv, theory)
    assert np.abs(np.abs(walls) - L / 2).max() < 0.5      # no-slip walls at +-L/2
    assert abs(d.reshape(L + 2, 8, 8)[2:L].mean() - n) < 0.1


# ---- frozen-assignment model (reference gradient of the pathwise adjoint) ----

def _frozen_setup(seed=31):
    volume, periodic, n = (6, 5, 10), (1, 1, 0), 8
    o = Oracle(volume, periodic, n=n, drag=0.002, delta_t=0.02, rotation_deg=120.0, rng="counter", seed=(seed, 0))
    o.step(3)
    zlo, zhi = 1, volume[2] - 1
    target = np.linspace(-0.02, 0.03, zhi - zlo)
    return o, target, zlo, zhi


def test_frozen_model_follows_the_real_trajectory():
    """At the oracle's own parameters the float64 shadow velocities are the real
    (bit-exact reference) velocities up to rounding: same loss and profile."""
    o, target, zlo, zhi = _frozen_setup()
    before = o.particles()
    lr, pr = o.frozen_profile_loss(12, 0.002, 120.0, target, zlo, zhi, use_real=True)
    ls, ps = o.frozen_profile_loss(12, 0.002, 120.0, target, zlo, zhi)
    assert np.abs(pr - ps).max() < 1e-6
    assert abs(lr - ls) <= 1e-5 * abs(lr)
    after = o.particles()
    assert np.array_equal(before[0], after[0]) and np.array_equal(before[1], after[1])
    # and the real trajectory equals stepping the oracle and averaging the layers itself
    prof = np.zeros(zhi - zlo)
    q = Oracle((6, 5, 10), (1, 1, 0), n=8, drag=0.002, delta_t=0.02, rng="counter", seed=(31, 0))
    q.step(3)
    for _ in range(12):
        q.step(1)
        pos, vel = q.particles()
        iz = np.floor(pos[:, 2] + np.float32(5.0)).astype(np.int64) % 10
        for z in range(zlo, zhi):
            prof[z - zlo] += vel[iz == z, 0].astype(np.float64).mean() / 12
    assert np.abs(prof - pr).max() < 1e-12
    assert abs(np.mean((prof - target) ** 2) - lr) < 1e-12


def test_frozen_model_is_affine_in_the_force_terms():
    """Shadow velocities are affine in g and f: the loss is an exact quadratic in g,
    and a uniform force field equals the same drag."""
    o, target, zlo, zhi = _frozen_setup()
    h = 5e-4
    L = [o.frozen_profile_loss(8, 0.002 + k * h, 120.0, target, zlo, zhi)[0] for k in range(-2, 3)]
    d2a = L[0] - 2 * L[1] + L[2]
    d2b = L[2] - 2 * L[3] + L[4]
    assert abs(d2a - d2b) < 1e-9 * max(1.0, abs(d2a)) + 1e-14
    assert d2a > 0
    C = o.C
    f = np.full(C, 7e-4)
    lf = o.frozen_profile_loss(8, 0.002, 120.0, target, zlo, zhi, force=f)[0]
    lg = o.frozen_profile_loss(8, 0.002 + 7e-4, 120.0, target, zlo, zhi)[0]
    assert lf == lg


def test_oracle_force_field_matches_drag():
    """The oracle's float force path: a uniform f with g' = g - f gives v_x += (g' + f) dt,
    identical to drag g when (g' dt + f dt) rounds to g dt; f = 0 is the plain rule bit for bit."""
    o1 = Oracle((6, 5, 10), (1, 1, 0), n=8, drag=0.002, delta_t=0.02, rng="counter", seed=(5, 0))
    o2 = Oracle((6, 5, 10), (1, 1, 0), n=8, drag=0.002, delta_t=0.02, rng="counter", seed=(5, 0))
    o2.set_force(np.zeros(o2.C, np.float32))
    o1.step(6)
    o2.step(6)
    assert np.array_equal(o1.particles()[1], o2.particles()[1])
    o2.set_force(np.full(o2.C, 0.001, np.float32))
    o1.step(4)
    o2.step(4)
    # force adds momentum n_collided * f dt per step in x
    dv = o2.particles()[1][:, 0].astype(np.float64).sum() - o1.particles()[1][:, 0].astype(np.float64).sum()
    assert dv > 0


def test_reference_fp32_kahan_accumulation_error_is_stated():
    """The reference accumulates the per-step mean fields in FP32 with Kahan
    compensation (cpu_backend.cpp:85-107); this build in FP64.  On the
    reference trajectory (the oracle in reference mode follows it bit for bit)
    of the README Poiseuille case, the two differ by at most 3.1e-8 in any
    cell's mean velocity after 100,000 samples and give the same viscosity to
    7 digits (tests/golden/c1_reference_cpu.json, written by
    tests/golden/make_c1_fixture.py); the oracle's FP32 path reproduces the
    reference's own mean exactly."""
    import json
    from pathlib import Path

    d = json.loads((Path(__file__).parent / "golden" / "c1_reference_cpu.json").read_text())
    for k in ("c1", "readme"):
        acc = d[f"accumulation_{k}"]
        assert acc["eta_fp32_kahan"] == d[f"reference_{k}"]["eta"]
        assert abs(acc["eta_fp64"] - acc["eta_fp32_kahan"]) < 1e-6 * acc["eta_fp64"]
        assert acc["max_abs_velocity_diff"] < 5e-8
        assert acc["max_abs_density_diff"] < 1e-6
    # the protocols' own viscosity: the developing flow biases the estimate up
    assert d["reference_c1"]["eta"] > 2.5 * d["theory_eta"]
    assert 1.05 * d["theory_eta"] < d["reference_readme"]["eta"] < 1.2 * d["theory_eta"]


# ------------------------------------------------------ 4. lane-group order
@pytest.mark.parametrize("algorithm", ["srd", "extended"])
@pytest.mark.parametrize("volume,periodic,n", [
    ((8, 7, 12), (1, 1, 0), 10),   # two lanes per cell
    ((6, 6, 8), (1, 1, 0), 20),    # four
    ((5, 5, 8), (1, 1, 0), 40),    # eight; extended cells of more than 64 members
    ((10, 9, 14), (1, 1, 0), 16),  # four, uneven member counts per lane
    ((6, 5, 10), (1, 1, 1), 3),    # one lane: the serial order itself
])
def test_lane_group_order_matches_serial_order(algorithm, volume, periodic, n):
    """Per-cell sums in the GPU's lane-group order against the reference's
    serial order, one step at a time from a common state (six steps, wall
    ghosts included): per-particle velocities within 1e-5 of the velocity
    scale and 1e-5 norm-wise, positions within FP32 rounding (SRD: bitwise,
    they do not depend on the sums), bitwise equal with one lane per cell, and
    not bitwise equal otherwise (the two orders do round differently)."""
    kw = dict(n=n, drag=0.01, delta_t=0.02, algorithm=algorithm, rng="counter", seed=(77, 0))
    serial = Oracle(volume, periodic, order="serial", **kw)
    lanes = Oracle(volume, periodic, order="lanes", **kw)
    differ = False
    for st in range(6):
        p, v = lanes.particles()
        serial.set_particles(p, v)
        serial.set_step(st)
        lanes.set_step(st)
        serial.step(1)
        lanes.step(1)
        ps, vs = serial.particles()
        pl, vl = lanes.particles()
        scale = np.abs(vs).max()
        assert np.abs(vl - vs).max() <= 1e-5 * scale
        assert np.linalg.norm(vl - vs) <= 1e-5 * np.linalg.norm(vs)
        if algorithm == "srd":
            assert np.array_equal(pl, ps)
        else:
            assert np.abs(pl - ps).max() <= 4 * np.finfo(np.float32).eps * max(volume)
        if n < 6:
            assert np.array_equal(vl, vs) and np.array_equal(pl, ps)
        differ |= not np.array_equal(vl, vs)
    assert differ == (n >= 6)
