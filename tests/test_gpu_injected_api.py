"""The injected-draws entry points (mpcd_hip_step_injected and
mpcd_hip_step_injected_extended) check their arguments before any device
work: arrays of the wrong length, draws out of range and the wrong collision
rule are rejected with a message, and the handle keeps stepping as before."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from cudampcd_amd.capi import Handle
from helpers import make_params

pytestmark = pytest.mark.gpu

_f32p = ctypes.POINTER(ctypes.c_float)
_u32p = ctypes.POINTER(ctypes.c_uint32)


def _p(a, t):
    return a.ctypes.data_as(t)


def _setup(pympcd, algorithm, n=4, fluid=(4, 3, 4)):
    p = make_params(pympcd, fluid, (1, 1, 0), n=n, algorithm=algorithm, seed=9, drag=0.001)
    sim = pympcd.Simulation(p, "hip")
    return p, sim, Handle(sim)


def _rejects(h, rc, words):
    assert rc != 0
    msg = h.L.mpcd_hip_last_error(h.h).decode()
    assert all(w in msg for w in words), msg


def _still_steps(pympcd, p, sim):
    """After the rejected calls the handle steps like a fresh one from the same state."""
    twin = pympcd.Simulation(p, "hip")
    twin.set_particles(sim.get_particle_positions(), sim.get_particle_velocities())
    Handle(twin).step_index = Handle(sim).step_index
    sim.step(2)
    twin.step(2)
    assert np.array_equal(sim.get_particle_velocities(), twin.get_particle_velocities())
    assert np.array_equal(sim.get_particle_positions(), twin.get_particle_positions())


def test_srd_injected_step_rejects_bad_arguments(pympcd):
    p, sim, h = _setup(pympcd, "srd")
    X, Y = 4, 3
    W = 2 * X * Y
    shift = np.zeros(3, np.float32)
    axes = np.tile(np.array([0, 0, 1], np.float32), (h.C, 1))
    gc = np.zeros(W, np.uint32)
    gv = np.zeros(W * p.n * 3, np.float32)
    L = h.L
    # ghost velocities of the wrong length
    _rejects(h, L.mpcd_hip_step_injected(h.h, _p(shift, _f32p), _p(axes, _f32p), h.C, _p(gc, _u32p), _p(gv, _f32p),
                                         W, gv.size - 3), ["ghost velocities"])
    # axes for another number of cells
    _rejects(h, L.mpcd_hip_step_injected(h.h, _p(shift, _f32p), _p(axes, _f32p), h.C - 1, _p(gc, _u32p),
                                         _p(gv, _f32p), W, gv.size), ["different number of cells"])
    # a system with walls without ghost arrays
    _rejects(h, L.mpcd_hip_step_injected(h.h, _p(shift, _f32p), _p(axes, _f32p), h.C, None, None, 0, 0),
             ["wall"])
    # a ghost count above n
    bad = gc.copy()
    bad[3] = p.n + 1
    _rejects(h, L.mpcd_hip_step_injected(h.h, _p(shift, _f32p), _p(axes, _f32p), h.C, _p(bad, _u32p), _p(gv, _f32p),
                                         W, gv.size), ["ghost count above n"])
    # the extended rule's entry point on an SRD handle
    ext = np.zeros((h.C, 3), np.uint32)
    th = np.zeros(h.C * 4 * p.n, np.float32)
    gg = np.zeros(W * p.n * 6, np.float32)
    _rejects(h, L.mpcd_hip_step_injected_extended(h.h, _p(shift, _f32p), _p(ext, _u32p), _p(th, _f32p), h.C,
                                                  th.size, _p(gc, _u32p), _p(gg, _f32p), W, gg.size),
             ["extended-rule"])
    _still_steps(pympcd, p, sim)


def test_extended_injected_step_rejects_bad_arguments(pympcd):
    p, sim, h = _setup(pympcd, "extended")
    X, Y = 4, 3
    W = 2 * X * Y
    shift = np.zeros(3, np.float32)
    ext = np.zeros((h.C, 3), np.uint32)
    th = np.zeros(h.C * 4 * p.n, np.float32)
    gc = np.zeros(W, np.uint32)
    gg = np.zeros(W * p.n * 6, np.float32)
    L = h.L

    def call(ext=ext, th=th, n_th=None, gc=gc, gg=gg, n_gg=None, n_cells=None):
        return L.mpcd_hip_step_injected_extended(h.h, _p(shift, _f32p), _p(ext, _u32p), _p(th, _f32p),
                                                 h.C if n_cells is None else n_cells,
                                                 th.size if n_th is None else n_th, _p(gc, _u32p), _p(gg, _f32p),
                                                 W, gg.size if n_gg is None else n_gg)

    _rejects(h, call(n_th=th.size - 4), ["thermal draws"])
    _rejects(h, call(n_gg=gg.size - 6), ["ghosts must hold"])
    _rejects(h, call(n_cells=h.C + 1), ["different number of cells"])
    bad = ext.copy()
    bad[5, 0] = 13
    _rejects(h, call(ext=bad), ["direction index"])
    bad = ext.copy()
    bad[2, 1] = 2
    _rejects(h, call(ext=bad), ["sign"])
    badc = gc.copy()
    badc[0] = p.n + 1
    _rejects(h, call(gc=badc), ["ghost count above n"])
    # the SRD entry point on an extended-rule handle
    axes = np.zeros((h.C, 3), np.float32)
    gv = np.zeros(W * p.n * 3, np.float32)
    _rejects(h, L.mpcd_hip_step_injected(h.h, _p(shift, _f32p), _p(axes, _f32p), h.C, _p(gc, _u32p), _p(gv, _f32p),
                                         W, gv.size), ["SRD"])
    _still_steps(pympcd, p, sim)
