"""The GPU's step against the reference CPU backend itself, with the
reference's own random numbers injected (mpcd_hip_step_injected and
mpcd_hip_step_injected_extended).

The reference (cpu_backend.cpp + simulation_functions.cpp, compiled in place
into oracle/_ref/libref.so) draws from its Xoshiro128+ stream inside its step.
The oracle in its reference stream reproduces that trajectory bit for bit
(test_oracle_pinning.py), so the draws it logs for a step are the reference's:
the grid shift, every colliding cell's rotation axis, and the ghost counts and
velocities of the wall layers; for the extended rule every colliding cell's
direction, sign and gate uniform, its members' thermal Gaussians and the wall
ghosts' positions.  From the reference's state, the GPU steps once
with those draws and must match the reference's result particle by particle:
velocities within 1e-5 (norm-wise relative; the GPU sums each cell in its
lane-group order, the reference serially, about 1e-7 apart), positions within
FP32 rounding, total momentum to FP32 round-off.
"""
from __future__ import annotations

import numpy as np
import pytest

from cudampcd_amd.capi import Handle
from helpers import make_params, min_image, stored_volume, vel_error
from oracle_lib import Oracle, Reference, ref_available

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not ref_available(), reason="oracle/_ref not built (reference sources absent)")]


@pytest.mark.parametrize("fluid,periodic,n", [
    ((8, 7, 10), (1, 1, 0), 10),     # walls: ghosts in both wall layers
    ((6, 5, 9), (1, 1, 1), 10),      # periodic
    ((7, 6, 8), (1, 1, 0), 20),      # four lanes per cell
    ((5, 4, 5), (1, 1, 0), 3),       # sparse: lone and empty cells
    ((4, 4, 4), (1, 1, 0), 1),       # table overflow (cap 4n = 4)
], ids=["walls-n10", "periodic-n10", "walls-n20", "sparse-n3", "overflow-n1"])
def test_step_with_reference_draws_matches_reference(pympcd, fluid, periodic, n):
    p = make_params(pympcd, fluid, periodic, n=n, algorithm="srd", seed=5, drag=0.01)
    volume = stored_volume(p)
    kw = dict(n=n, drag=p.drag, delta_t=p.delta_t, algorithm="srd", seed=(0x1234567, 0x89abcdef))
    ref = Reference(volume, periodic, **kw)
    orc = Oracle(volume, periodic, rng="reference", **kw)
    orc.log_draws(True)
    sim = pympcd.Simulation(p, "hip")
    h = Handle(sim)
    assert h.N == ref.N
    h.record_permutation(True)
    walls = periodic[2] == 0
    L = np.array(volume, np.float64)
    for _ in range(6):
        pos, vel = ref.particles()
        opos, ovel = orc.particles()
        assert np.array_equal(pos, opos) and np.array_equal(vel, ovel)
        sim.set_particles(pos, vel)
        orc.step(1)
        ref.step(1)
        shift, axes, gcnt, gvel = orc.last_draws()
        if walls:
            assert gcnt.sum() > 0
            h.step_injected(shift, axes, gcnt, gvel)
        else:
            h.step_injected(shift, axes)
        perm = h.last_permutation()
        rpos, rvel = ref.particles()
        rpos, rvel = rpos[perm], rvel[perm]
        gpos, gv = sim.get_particle_positions(), sim.get_particle_velocities()
        assert vel_error(gv, rvel) <= 1e-5
        assert np.linalg.norm(gv.astype(np.float64) - rvel) <= 1e-5 * np.linalg.norm(rvel.astype(np.float64))
        assert np.abs(min_image(gpos.astype(np.float64) - rpos, L)).max() <= 1e-5 * L.max()
        dp = np.abs(gv.astype(np.float64).sum(axis=0) - rvel.astype(np.float64).sum(axis=0)).max()
        assert dp <= 1e-5 * len(gv)


@pytest.mark.parametrize("fluid,periodic,n", [
    ((8, 7, 10), (1, 1, 0), 10),     # walls: ghosts with positions in both wall layers
    ((6, 5, 9), (1, 1, 1), 10),      # periodic
    ((6, 6, 8), (1, 1, 0), 13),      # four lanes per cell
    ((6, 5, 6), (1, 1, 0), 30),      # eight lanes per cell
    ((5, 4, 5), (1, 1, 0), 3),       # sparse
    ((4, 4, 4), (1, 1, 0), 1),       # ghosts fill the table: dropped table entries
], ids=["walls-n10", "periodic-n10", "walls-n13", "walls-n30", "sparse-n3", "overflow-n1"])
def test_extended_step_with_reference_draws_matches_reference(pympcd, fluid, periodic, n):
    """The extended rule with the reference's draws: its gate compares the same
    uniform with 1 - exp(.) (the reference's libm exp, the GPU's own series,
    equal to about an ulp), and which side of the cell a member is on follows
    the centre of mass (lane-group sums); a decision that came out differently
    would change a whole cell, so every cell is compared: all within 1e-5."""
    p = make_params(pympcd, fluid, periodic, n=n, algorithm="extended", seed=5, drag=0.01)
    volume = stored_volume(p)
    kw = dict(n=n, drag=p.drag, delta_t=p.delta_t, algorithm="extended", seed=(0x2468ace, 0x13579bdf))
    ref = Reference(volume, periodic, **kw)
    orc = Oracle(volume, periodic, rng="reference", **kw)
    orc.log_draws(True)
    sim = pympcd.Simulation(p, "hip")
    h = Handle(sim)
    assert h.N == ref.N
    h.record_permutation(True)
    walls = periodic[2] == 0
    L = np.array(volume, np.float64)
    collided = gated = 0
    for _ in range(6):
        pos, vel = ref.particles()
        sim.set_particles(pos, vel)
        orc.step(1)
        ref.step(1)
        shift, ext, gauss, gcnt, ghosts = orc.last_ext_draws()
        # the reference keeps a cell's sides in one 64-bit word: its behaviour
        # is defined for cells of at most 64 members (table and ghosts)
        assert (np.count_nonzero(gauss, axis=1) <= 64).all()
        collided += int(np.count_nonzero(gauss))
        gated += int(np.count_nonzero(ext[:, 2]))
        if walls:
            h.step_injected_extended(shift, ext, gauss, gcnt, ghosts)
        else:
            h.step_injected_extended(shift, ext, gauss)
        perm = h.last_permutation()
        rpos, rvel = ref.particles()
        rpos, rvel = rpos[perm], rvel[perm]
        gpos, gv = sim.get_particle_positions(), sim.get_particle_velocities()
        assert vel_error(gv, rvel) <= 1e-5
        assert np.linalg.norm(gv.astype(np.float64) - rvel) <= 1e-5 * np.linalg.norm(rvel.astype(np.float64))
        assert np.abs(min_image(gpos.astype(np.float64) - rpos, L)).max() <= 1e-5 * L.max()
        dp = np.abs(gv.astype(np.float64).sum(axis=0) - rvel.astype(np.float64).sum(axis=0)).max()
        assert dp <= 1e-5 * len(gv)
    assert gated > 0
    assert collided > 0 or n == 1   # one particle a cell: 1 - exp(.) is tiny, the gates stay shut
