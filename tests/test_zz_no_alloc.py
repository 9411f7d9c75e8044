"""The allocation-free step (mpcd_hip_step_noalloc, scratch in a static device
array) against the step of a handle (mpcd_hip_step) from the same state at
the same step index: bit-identical states, on two calls in a row, the second
on a smaller system (the array keeps the first call's contents), for an odd
and an even number of steps, with walls and periodic, both rules.  The device
buffers are the caller's, allocated with the HIP runtime the library uses.
"""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

from cudampcd_amd import capi
from cudampcd_amd.capi import Handle
from helpers import make_params

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


def _hip():
    capi.lib()   # loads libmpcd_hip.so and, with it, its HIP runtime
    hip = ctypes.CDLL("libamdhip64.so.7")   # the runtime already loaded under this name
    hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    hip.hipFree.argtypes = [ctypes.c_void_p]
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
    return hip


def _check(rc, what):
    assert rc == 0, f"{what}: HIP error {rc}"


def _records(pos, vel):
    return np.ascontiguousarray(np.concatenate([pos, vel], axis=1), dtype=np.float32)


# consecutive calls: (fluid cells, periodicity, n, steps); the second system is smaller
CALLS = [((24, 20, 14), (1, 1, 0), 10, 3), ((6, 5, 8), (1, 1, 1), 4, 2)]


@pytest.mark.parametrize("algorithm", ["srd", "extended"])
def test_noalloc_step_matches_the_handle_step(pympcd, algorithm):
    hip = _hip()
    for call, (fluid, perio, n, steps) in enumerate(CALLS):
        p = make_params(pympcd, fluid, perio, n=n, algorithm=algorithm, seed=9000 + call, drag=0.002)
        sim = pympcd.Simulation(p, "hip")
        h = Handle(sim)
        sim.step(2)
        start = _records(sim.get_particle_positions(), sim.get_particle_velocities())
        step_index = h.step_index
        sim.step(steps)                                     # the handle's step
        want = _records(sim.get_particle_positions(), sim.get_particle_velocities())
        N = len(start)
        print(f"[noalloc] {algorithm} call {call}: {fluid} {perio} n={n}, {N} particles, {steps} steps", flush=True)
        bufs = [ctypes.c_void_p(), ctypes.c_void_p()]
        for b in bufs:
            _check(hip.hipMalloc(ctypes.byref(b), 24 * N), "hipMalloc")
        try:
            _check(hip.hipMemcpy(bufs[0], start.ctypes.data, 24 * N, H2D), "upload")
            _check(hip.hipMemset(bufs[1], 0xFF, 24 * N), "fill")
            capi.step_noalloc(capi.c_params(p), step_index, steps, bufs[0].value, bufs[1].value, N)
            got = np.empty_like(start)
            _check(hip.hipMemcpy(got.ctypes.data, bufs[1], 24 * N, D2H), "download")
        finally:
            for b in bufs:
                hip.hipFree(b)
        assert np.array_equal(got, want), (algorithm, call)
        del sim, h


def test_noalloc_step_rejects_systems_beyond_its_array(pympcd):
    p = make_params(pympcd, (128, 128, 20), (1, 1, 0), n=10, seed=1)     # 3.3e6 particles
    c = capi.c_params(p)
    with pytest.raises(RuntimeError, match="at most"):
        capi.step_noalloc(c, 0, 1, 16, 32, int(p.N))        # rejected before any device work
