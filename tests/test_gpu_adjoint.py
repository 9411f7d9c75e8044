"""Pathwise adjoint of the time-averaged velocity-profile loss (SRD).

Reference gradient: central finite differences of the oracle's frozen-
assignment model, which steps the oracle's bit-exact trajectory and carries
float64 velocities beside it with the perturbed drag, angle and force while
cells, tables, bounces, axes and ghosts stay those of the trajectory (the
model the pathwise gradient differentiates; pinned in test_oracle_pinning.py).
Also: the forward loss against the oracle's real trajectory with a force
field, independence of the checkpoint interval, central finite differences
of the full GPU model with common random numbers (cells allowed to change),
and the torch.autograd.Function.
"""
from __future__ import annotations

import numpy as np
import pytest

from cudampcd_amd import adjoint
from cudampcd_amd.capi import Handle
from helpers import make_params, stored_volume
from oracle_lib import Oracle

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

STEPS = 12


def _setup(pympcd, fluid=(6, 5, 8), n=8, drag=0.002, angle=120.0, seed=31, warm=3):
    p = make_params(pympcd, fluid, (1, 1, 0), n=n, seed=seed, drag=drag, angle=angle)
    sim = pympcd.Simulation(p, "hip")
    sim.step(warm)
    Z = stored_volume(p)[2]
    target = np.linspace(-0.02, 0.03, Z - 2)
    return p, sim, target


def _oracle_at(p, sim):
    o = Oracle(stored_volume(p), tuple(p.periodicity), n=p.n, drag=p.drag, delta_t=p.delta_t,
               thermal_velocity=float(np.sqrt(p.temperature)), rotation_deg=p.rotation_angle, algorithm="srd",
               rng="counter", seed=(p.seed, 0), distribute=False)
    o.set_particles(sim.get_particle_positions(), sim.get_particle_velocities())
    o.set_step(Handle(sim).step_index)
    return o


def _cells(p):
    X, Y, Z = stored_volume(p)
    return X * Y * Z


def test_forward_loss_matches_oracle_and_leaves_state(pympcd):
    p, sim, target = _setup(pympcd)
    Z = stored_volume(p)[2]
    pos0, vel0 = sim.get_particle_positions(), sim.get_particle_velocities()
    step0 = Handle(sim).step_index
    rng = np.random.default_rng(4)
    f_host = (rng.normal(size=_cells(p)) * 2e-3).astype(np.float32)
    o = _oracle_at(p, sim)
    o.set_force(f_host)
    # one step from the same state: the force path is bit-identical (the layer
    # sums differ only by the 2^-40 fixed point of the device accumulation)
    r1 = adjoint.profile_loss(sim, 1, p.drag, p.rotation_angle, target, force=f_host)
    l1, p1 = o.frozen_profile_loss(1, p.drag, p.rotation_angle, target, 1, Z - 1, use_real=True)
    assert np.abs(r1.profile - p1).max() < 1e-11
    # several steps: after the first step the device holds the particles in cell
    # order and the oracle in index order, so in-cell sums round differently
    r = adjoint.profile_loss(sim, STEPS, p.drag, p.rotation_angle, target, force=f_host)
    lo, po = o.frozen_profile_loss(STEPS, p.drag, p.rotation_angle, target, 1, Z - 1, use_real=True)
    assert np.abs(r.profile - po).max() < 1e-7
    assert abs(r.loss - lo) <= 1e-4 * abs(lo)
    # the call leaves the simulation where it was
    assert Handle(sim).step_index == step0
    assert np.array_equal(sim.get_particle_positions(), pos0)
    assert np.array_equal(sim.get_particle_velocities(), vel0)
    # and repeats itself exactly (common random numbers)
    r2 = adjoint.profile_loss(sim, STEPS, p.drag, p.rotation_angle, target, force=f_host)
    assert r2.loss == r.loss


# n = 8: the reverse step stages each block's 64 cells in LDS; n = 20: the
# blocks hold more particles than the staging buffer (768) and take the
# one-lane-per-cell path
@pytest.mark.parametrize("n", [8, 20], ids=["staged", "per-cell"])
def test_gradient_matches_frozen_model_finite_differences(pympcd, n):
    """Adjoint gradient vs central differences of the oracle's float64 frozen model."""
    p, sim, target = _setup(pympcd, n=n)
    Z = stored_volume(p)[2]
    C = _cells(p)
    r = adjoint.profile_loss_grad(sim, STEPS, 5, p.drag, p.rotation_angle, target)
    o = _oracle_at(p, sim)

    def L(g=p.drag, a=p.rotation_angle, f=None):
        return o.frozen_profile_loss(STEPS, g, a, target, 1, Z - 1, force=f)[0]

    assert abs(r.loss - L()) <= 1e-5 * abs(r.loss)
    hg = 1e-4
    fd_g = (L(g=p.drag + hg) - L(g=p.drag - hg)) / (2 * hg)     # exact: quadratic in g
    ha = 0.01
    fd_a = (L(a=p.rotation_angle + ha) - L(a=p.rotation_angle - ha)) / (2 * ha)
    print(f"dL/dg adjoint {r.grad_drag:.9e} fd {fd_g:.9e}; dL/dangle adjoint {r.grad_angle:.9e} fd {fd_a:.9e}")
    assert abs(r.grad_drag - fd_g) <= 1e-3 * abs(fd_g)
    assert abs(r.grad_angle - fd_a) <= 1e-3 * abs(fd_a)
    g = r.grad_force.astype(np.float64)
    rng = np.random.default_rng(8)
    d = rng.normal(size=C)
    hf = 1e-4
    fd_d = (L(f=hf * d) - L(f=-hf * d)) / (2 * hf)
    assert abs(g @ d - fd_d) <= 1e-3 * abs(fd_d)
    for c in np.argsort(-np.abs(g))[:3]:
        e = np.zeros(C)
        e[c] = 1.0
        fd_c = (L(f=hf * e) - L(f=-hf * e)) / (2 * hf)
        assert abs(g[c] - fd_c) <= 2e-3 * abs(fd_c) + 1e-9
    # the sum of the force gradient over cells is the drag gradient (same kick)
    assert abs(g.sum() - r.grad_drag) <= 1e-4 * abs(r.grad_drag)


def test_checkpoint_interval_does_not_change_the_gradient(pympcd):
    p, sim, target = _setup(pympcd, seed=77)
    out = []
    # one segment, two (both taped by the forward pass), several with a short last one
    for K in (1, 3, 5, 6, 7, STEPS, 1000):
        r = adjoint.profile_loss_grad(sim, STEPS, K, p.drag, p.rotation_angle, target)
        out.append((r.loss, r.grad_drag, r.grad_angle, r.grad_force))
    for o in out[1:]:
        assert o[0] == out[0][0] and o[1] == out[0][1] and o[2] == out[0][2]
        assert np.array_equal(o[3], out[0][3])


@pytest.mark.parametrize("seeds", [(5, 6, 7)])
def test_gradient_against_full_model_finite_differences(pympcd, seeds):
    """Central differences of the GPU forward itself, common random numbers.

    Under the perturbation particles near cell faces change cells, which the
    pathwise gradient leaves out by construction (the reassignment term).
    Near the loss minimum the gradient is small and that term dominates a
    single realisation (16x16x20 cells, n = 10, 20 steps, target between
    -0.02 and 0.03: adjoint against differences -74% to +38% over six
    seeds), so the target here lies away from the flow (0.05 in every
    layer): the gradient is then carried by the momentum input g dt, exact in
    the pathwise gradient, and the reassignment term is a correction (six
    seeds: -9% to +7%, about -5% on average).  The loss is quadratic in g
    (the profile is linear in it), so the central difference is exact at any
    step: a step of g itself averages over many reassignments, where 5% of g
    left differences scattered by up to 44%.  Checked: each seed within 15%
    with the same sign, the seed mean within 10%."""
    adj, fds = [], []
    for seed in seeds:
        p, sim, _ = _setup(pympcd, fluid=(16, 16, 20), n=10, seed=seed, warm=20)
        steps = 20
        Z = stored_volume(p)[2]
        target = np.full(Z - 2, 0.05)
        r = adjoint.profile_loss_grad(sim, steps, 7, p.drag, p.rotation_angle, target, want_force_grad=False)
        hg = p.drag
        lp = adjoint.profile_loss(sim, steps, p.drag + hg, p.rotation_angle, target).loss
        lm = adjoint.profile_loss(sim, steps, p.drag - hg, p.rotation_angle, target).loss
        fd_g = (lp - lm) / (2 * hg)
        print(f"seed {seed}: dL/dg adjoint {r.grad_drag:.6e} fd {fd_g:.6e} ({r.grad_drag / fd_g - 1:+.2e})")
        assert np.sign(r.grad_drag) == np.sign(fd_g)
        assert abs(r.grad_drag - fd_g) <= 0.15 * abs(fd_g)
        adj.append(r.grad_drag)
        fds.append(fd_g)
    assert abs(np.mean(adj) - np.mean(fds)) <= 0.10 * abs(np.mean(fds))


def test_torch_autograd_function(pympcd):
    p, sim, target = _setup(pympcd, seed=13)
    C = _cells(p)
    drag = torch.tensor(p.drag, dtype=torch.float64, requires_grad=True)
    angle = torch.tensor(p.rotation_angle, dtype=torch.float64, requires_grad=True)
    force = torch.zeros(C, dtype=torch.float32, requires_grad=True)
    loss = adjoint.ProfileLoss.apply(drag, angle, force, sim, target, STEPS, 4)
    (3.0 * loss).backward()
    r = adjoint.profile_loss_grad(sim, STEPS, 4, p.drag, p.rotation_angle, target, force=np.zeros(C, np.float32))
    assert loss.item() == r.loss
    assert drag.grad.item() == pytest.approx(3.0 * r.grad_drag, rel=1e-12)
    assert angle.grad.item() == pytest.approx(3.0 * r.grad_angle, rel=1e-12)
    assert np.allclose(force.grad.numpy(), 3.0 * r.grad_force, rtol=1e-6, atol=0)
    with torch.no_grad():
        l2 = adjoint.ProfileLoss.apply(drag, angle, force, sim, target, STEPS, 4)
    assert l2.item() == r.loss


def test_adjoint_rejects_what_it_does_not_model(pympcd):
    p = make_params(pympcd, (6, 5, 8), (1, 1, 0), n=4, algorithm="extended", seed=2)
    sim = pympcd.Simulation(p, "hip")
    with pytest.raises(RuntimeError, match="SRD"):
        adjoint.profile_loss(sim, 3, p.drag, p.rotation_angle, np.zeros(8))
    p, sim, target = _setup(pympcd)
    with pytest.raises(RuntimeError, match="layers"):
        adjoint.profile_loss(sim, 3, p.drag, p.rotation_angle, target[:-1])
    with pytest.raises(ValueError, match="cells"):
        adjoint.profile_loss(sim, 3, p.drag, p.rotation_angle, target, force=np.zeros(10, np.float32))


def test_layer_sums_reject_velocities_beyond_their_exact_range(pympcd):
    """The layers' x-velocity sums are fixed point (2^-40), exact while every
    |v_x| < 128: a particle beyond that makes the loss fail with a clear error,
    and the handle works again once the state is sane."""
    p, sim, target = _setup(pympcd)
    pos, vel = sim.get_particle_positions(), sim.get_particle_velocities()
    bad = vel.copy()
    bad[7, 0] = 200.0
    sim.set_particle_velocities(bad)
    with pytest.raises(RuntimeError, match="fixed point"):
        adjoint.profile_loss(sim, 3, p.drag, p.rotation_angle, target)
    sim.set_particle_positions(pos)
    sim.set_particle_velocities(vel)
    r = adjoint.profile_loss(sim, 3, p.drag, p.rotation_angle, target)
    assert np.isfinite(r.loss)
