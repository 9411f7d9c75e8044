"""ctypes binding of the C-ABI in include/mpcd_hip.h (libmpcd_hip.so).

The pybind11 module covers the reference's Python surface; this module exposes
the remaining entry points (parity hooks, timing, tiling, device self-tests) on
the same handle, obtained with ``Simulation._handle()``.
"""
from __future__ import annotations

import ctypes
import re
from pathlib import Path

import numpy as np

PKG = Path(__file__).resolve().parent
LIB_PATH = PKG / "libmpcd_hip.so"
HEADER = PKG.parent / "include" / "mpcd_hip.h"

_u32p = ctypes.POINTER(ctypes.c_uint32)
_f32p = ctypes.POINTER(ctypes.c_float)
_f64p = ctypes.POINTER(ctypes.c_double)
_u64p = ctypes.POINTER(ctypes.c_uint64)
_vp = ctypes.c_void_p

_lib = None


class CParams(ctypes.Structure):
    """mpcd_params_t of include/mpcd_hip.h."""
    _fields_ = [("experiment", ctypes.c_int), ("algorithm", ctypes.c_int), ("collision_kernel", ctypes.c_int),
                ("periodicity", ctypes.c_int * 3), ("volume_size", ctypes.c_float * 3), ("N", ctypes.c_uint64),
                ("n", ctypes.c_uint32), ("device_id", ctypes.c_uint32), ("delta_t", ctypes.c_float),
                ("temperature", ctypes.c_float), ("thermal_velocity", ctypes.c_float), ("drag", ctypes.c_float),
                ("rotation_angle", ctypes.c_float), ("seed", ctypes.c_uint64), ("system_id", ctypes.c_uint32),
                ("reserved", ctypes.c_uint32)]


def c_params(p, seed=None) -> CParams:
    """The C-ABI parameter block of Params `p` (seed: p.seed unless given)."""
    c = CParams()
    c.experiment = 0 if p.experiment == "standard" else 1
    c.algorithm = 0 if p.algorithm == "srd" else 1
    c.collision_kernel = 2
    for a in range(3):
        c.periodicity[a] = int(p.periodicity[a])
        c.volume_size[a] = float(p.volume_size[a])
    c.N = int(p.N)
    c.n = int(p.n)
    c.device_id = int(p.device_id)
    c.delta_t = p.delta_t
    c.temperature = p.temperature
    c.thermal_velocity = p.thermal_velocity
    c.drag = p.drag
    c.rotation_angle = p.rotation_angle
    c.seed = int(p.seed if seed is None else seed)
    c.system_id = int(p.system_id)
    return c


def step_noalloc(params: CParams, step_index: int, n_steps: int, state_in: int, state_out: int,
                 n_particles: int) -> None:
    """mpcd_hip_step_noalloc on two device buffers (addresses of this library's HIP runtime)."""
    if lib().mpcd_hip_step_noalloc(ctypes.byref(params), step_index, n_steps, ctypes.c_void_p(state_in),
                                   ctypes.c_void_p(state_out), n_particles) != 0:
        raise RuntimeError(lib().mpcd_hip_last_error(None).decode())


def declared_symbols() -> list[str]:
    """Entry points declared in include/mpcd_hip.h."""
    text = HEADER.read_text()
    return sorted(set(re.findall(r"\b(mpcd_hip_[a-z0-9_]+)\s*\(", text)))


def lib() -> ctypes.CDLL:
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise ImportError(f"{LIB_PATH} is not built; run python cudampcd_amd/build.py")
        L = ctypes.CDLL(str(LIB_PATH))
        L.mpcd_hip_last_error.restype = ctypes.c_char_p
        L.mpcd_hip_last_error.argtypes = [_vp]
        sigs = {
            "mpcd_hip_step": [_vp, ctypes.c_int],
            "mpcd_hip_step_and_sample": [_vp, ctypes.c_int],
            "mpcd_hip_step_injected": [_vp, _f32p, _f32p, ctypes.c_size_t, _u32p, _f32p, ctypes.c_size_t,
                                       ctypes.c_size_t],
            "mpcd_hip_step_injected_extended": [_vp, _f32p, _u32p, _f32p, ctypes.c_size_t, ctypes.c_size_t,
                                                _u32p, _f32p, ctypes.c_size_t, ctypes.c_size_t],
            "mpcd_hip_get_sample_mean": [_vp, _f32p, _f32p, ctypes.c_size_t],
            "mpcd_hip_get_sample_mean_f64": [_vp, _f64p, _f64p, ctypes.c_size_t],
            "mpcd_hip_n_particles": [_vp, _u64p],
            "mpcd_hip_n_cells": [_vp, _u64p],
            "mpcd_hip_get_positions": [_vp, _f32p, ctypes.c_size_t],
            "mpcd_hip_get_velocities": [_vp, _f32p, ctypes.c_size_t],
            "mpcd_hip_set_positions": [_vp, _f32p, ctypes.c_size_t],
            "mpcd_hip_set_velocities": [_vp, _f32p, ctypes.c_size_t],
            "mpcd_hip_set_particles": [_vp, _f32p, _f32p, ctypes.c_size_t],
            "mpcd_hip_get_step_index": [_vp, _u64p],
            "mpcd_hip_set_step_index": [_vp, ctypes.c_uint64],
            "mpcd_hip_grid_shift": [_vp, ctypes.c_uint64, _f32p],
            "mpcd_hip_peek_sort": [_vp, _u32p, _u32p, _u32p],
            "mpcd_hip_record_permutation": [_vp, ctypes.c_int],
            "mpcd_hip_last_permutation": [_vp, _u32p, ctypes.c_size_t],
            "mpcd_hip_enable_timing": [_vp, ctypes.c_int],
            "mpcd_hip_last_timing": [_vp, _f64p, _f64p, ctypes.POINTER(ctypes.c_int)],
            "mpcd_hip_set_tiling": [_vp, ctypes.c_uint32, ctypes.c_uint32],
            "mpcd_hip_get_tiling": [_vp, _u32p, _u32p],
            "mpcd_hip_system_tiling": [_vp, ctypes.c_uint32, _u32p, _u32p],
            "mpcd_hip_n_systems": [_vp, _u32p],
            "mpcd_hip_system_info": [_vp, ctypes.c_uint32, _u64p, _u64p, _u64p, _u64p],
            "mpcd_hip_system_get_particles": [_vp, ctypes.c_uint32, _f32p, _f32p, ctypes.c_size_t],
            "mpcd_hip_system_sample_mean": [_vp, ctypes.c_uint32, _f32p, _f32p, ctypes.c_size_t],
            "mpcd_hip_system_sample_mean_f64": [_vp, ctypes.c_uint32, _f64p, _f64p, ctypes.c_size_t],
            "mpcd_hip_reset_sampling": [_vp],
            "mpcd_hip_use_graph": [_vp, ctypes.c_int],
            "mpcd_hip_get_sampling_state": [_vp, _f64p, ctypes.c_size_t, _u64p, ctypes.POINTER(ctypes.c_int)],
            "mpcd_hip_set_sampling_state": [_vp, _f64p, ctypes.c_size_t, ctypes.c_uint64, ctypes.c_int],
            "mpcd_hip_autotune": [_vp, ctypes.c_int, _u32p, _u32p, _f64p],
            "mpcd_hip_profile_loss": [_vp, ctypes.c_int, ctypes.c_float, ctypes.c_float, _f32p, _f64p,
                                      ctypes.c_size_t, _f64p, _f64p],
            "mpcd_hip_profile_loss_grad": [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float, _f32p,
                                           _f64p, ctypes.c_size_t, _f64p, _f64p, _f64p, _f64p, _f32p],
            "mpcd_hip_tune_capacities": [ctypes.c_int, ctypes.c_uint32, _u32p, ctypes.c_uint32, _u32p],
            "mpcd_hip_step_noalloc": [ctypes.POINTER(CParams), ctypes.c_uint64, ctypes.c_int, _vp, _vp,
                                      ctypes.c_size_t],
            "mpcd_hip_bench_copy": [ctypes.c_size_t, ctypes.c_int, _f64p],
            "mpcd_hip_selftest_philox": [ctypes.c_uint32, ctypes.c_uint32, _u32p, _u32p, ctypes.c_size_t],
            "mpcd_hip_selftest_math": [_f32p, _f32p, ctypes.c_size_t, ctypes.c_int],
        }
        for name, args in sigs.items():
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        _lib = L
    return _lib


def _p(a, t):
    return a.ctypes.data_as(t)


class Handle:
    """C-ABI calls on the handle of a live Simulation."""

    def __init__(self, sim):
        """sim: a Simulation or an Ensemble of this module."""
        self.sim = sim  # keeps the simulation alive
        self.h = ctypes.c_void_p(sim._handle())
        self.L = lib()
        n = ctypes.c_uint64()
        self._check(self.L.mpcd_hip_n_particles(self.h, ctypes.byref(n)))
        self.N = int(n.value)
        self._check(self.L.mpcd_hip_n_cells(self.h, ctypes.byref(n)))
        self.C = int(n.value)

    def _check(self, rc):
        if rc != 0:
            raise RuntimeError(self.L.mpcd_hip_last_error(self.h).decode())

    @property
    def step_index(self) -> int:
        v = ctypes.c_uint64()
        self._check(self.L.mpcd_hip_get_step_index(self.h, ctypes.byref(v)))
        return int(v.value)

    @step_index.setter
    def step_index(self, s: int):
        self._check(self.L.mpcd_hip_set_step_index(self.h, s))

    def step_injected(self, shift, axes, ghost_counts=None, ghost_velocities=None):
        """One SRD step with the caller's draws (see mpcd_hip_step_injected):
        shift (3,), axes (C, 3) in reference cell order, and with walls the
        ghost counts (2 X Y,) and velocities (2 X Y, n, 3) of the wall-layer
        cells."""
        sh = np.ascontiguousarray(shift, dtype=np.float32).reshape(3)
        ax = np.ascontiguousarray(axes, dtype=np.float32).reshape(-1)
        if ax.size != 3 * self.C:
            raise ValueError("axes must hold 3 values per cell")
        gc = None if ghost_counts is None else np.ascontiguousarray(ghost_counts, dtype=np.uint32).reshape(-1)
        gv = None if ghost_velocities is None else np.ascontiguousarray(ghost_velocities, dtype=np.float32).reshape(-1)
        if (gc is None) != (gv is None):
            raise ValueError("ghost counts and velocities go together")
        self._check(self.L.mpcd_hip_step_injected(self.h, _p(sh, _f32p), _p(ax, _f32p), self.C,
                                                  None if gc is None else _p(gc, _u32p),
                                                  None if gv is None else _p(gv, _f32p),
                                                  0 if gc is None else gc.size, 0 if gv is None else gv.size))

    def step_injected_extended(self, shift, cell_draws, thermal, ghost_counts=None, ghosts=None):
        """One extended-rule step with the caller's draws (see
        mpcd_hip_step_injected_extended): shift (3,), per cell (direction
        index, sign, gate uniform bits) (C, 3) uint32, thermal Gaussians
        (C, 4 n), and with walls the ghost counts (2 X Y,) and ghost positions
        and velocities (2 X Y, n, 6)."""
        sh = np.ascontiguousarray(shift, dtype=np.float32).reshape(3)
        cd = np.ascontiguousarray(cell_draws, dtype=np.uint32).reshape(-1)
        th = np.ascontiguousarray(thermal, dtype=np.float32).reshape(-1)
        if cd.size != 3 * self.C:
            raise ValueError("cell draws: 3 words per cell")
        gc = None if ghost_counts is None else np.ascontiguousarray(ghost_counts, dtype=np.uint32).reshape(-1)
        gg = None if ghosts is None else np.ascontiguousarray(ghosts, dtype=np.float32).reshape(-1)
        if (gc is None) != (gg is None):
            raise ValueError("ghost counts and ghosts go together")
        self._check(self.L.mpcd_hip_step_injected_extended(self.h, _p(sh, _f32p), _p(cd, _u32p), _p(th, _f32p),
                                                           self.C, th.size, None if gc is None else _p(gc, _u32p),
                                                           None if gg is None else _p(gg, _f32p),
                                                           0 if gc is None else gc.size,
                                                           0 if gg is None else gg.size))

    def grid_shift(self, step: int) -> np.ndarray:
        out = np.empty(3, np.float32)
        self._check(self.L.mpcd_hip_grid_shift(self.h, step, _p(out, _f32p)))
        return out

    def peek_sort(self):
        cell = np.empty(self.N, np.uint32)
        cnt = np.empty(self.C, np.uint32)
        perm = np.empty(self.N, np.uint32)
        self._check(self.L.mpcd_hip_peek_sort(self.h, _p(cell, _u32p), _p(cnt, _u32p), _p(perm, _u32p)))
        return cell, cnt, perm

    def record_permutation(self, on: bool = True):
        self._check(self.L.mpcd_hip_record_permutation(self.h, int(on)))

    def last_permutation(self) -> np.ndarray:
        out = np.empty(self.N, np.uint32)
        self._check(self.L.mpcd_hip_last_permutation(self.h, _p(out, _u32p), self.N))
        return out

    def enable_timing(self, on: bool = True):
        self._check(self.L.mpcd_hip_enable_timing(self.h, int(on)))

    def last_timing(self):
        a = ctypes.c_double()
        b = ctypes.c_double()
        n = ctypes.c_int()
        self._check(self.L.mpcd_hip_last_timing(self.h, ctypes.byref(a), ctypes.byref(b), ctypes.byref(n)))
        return a.value, b.value, n.value

    def set_tiling(self, cells_per_tile: int = 0, lds_capacity: int = 0):
        self._check(self.L.mpcd_hip_set_tiling(self.h, cells_per_tile, lds_capacity))

    def tiling(self):
        a = ctypes.c_uint32()
        b = ctypes.c_uint32()
        self._check(self.L.mpcd_hip_get_tiling(self.h, ctypes.byref(a), ctypes.byref(b)))
        return a.value, b.value

    def system_tiling(self, s: int):
        tc = ctypes.c_uint32()
        cap = ctypes.c_uint32()
        self._check(self.L.mpcd_hip_system_tiling(self.h, s, ctypes.byref(tc), ctypes.byref(cap)))
        return int(tc.value), int(cap.value)

    def use_graph(self, on: bool = True):
        self._check(self.L.mpcd_hip_use_graph(self.h, int(on)))

    @property
    def n_systems(self) -> int:
        v = ctypes.c_uint32()
        self._check(self.L.mpcd_hip_n_systems(self.h, ctypes.byref(v)))
        return int(v.value)

    def system_info(self, s: int):
        """(first particle, particles, first cell, cells) of system s."""
        v = [ctypes.c_uint64() for _ in range(4)]
        self._check(self.L.mpcd_hip_system_info(self.h, s, *[ctypes.byref(x) for x in v]))
        return tuple(int(x.value) for x in v)

    def reset_sampling(self):
        """End the running average: the next sampled step starts a new one."""
        self._check(self.L.mpcd_hip_reset_sampling(self.h))

    def sampling_state(self):
        """(accumulators (C, 4) float64, samples after the first, average running)."""
        acc = np.empty((self.C, 4), np.float64)
        cnt = ctypes.c_uint64()
        act = ctypes.c_int()
        self._check(self.L.mpcd_hip_get_sampling_state(self.h, _p(acc, _f64p), self.C, ctypes.byref(cnt), ctypes.byref(act)))
        return acc, int(cnt.value), bool(act.value)

    def set_sampling_state(self, acc, sample_counter: int, active: bool):
        a = np.ascontiguousarray(acc, dtype=np.float64).reshape(self.C, 4)
        self._check(self.L.mpcd_hip_set_sampling_state(self.h, _p(a, _f64p), self.C, sample_counter, int(active)))

    def autotune(self, steps_per_trial: int = 3):
        a = ctypes.c_uint32()
        b = ctypes.c_uint32()
        ms = ctypes.c_double()
        self._check(self.L.mpcd_hip_autotune(self.h, steps_per_trial, ctypes.byref(a), ctypes.byref(b), ctypes.byref(ms)))
        return a.value, b.value, ms.value


def tune_capacities(algorithm: str, nmax: int) -> list[int]:
    """LDS capacities the autotuner tries (no device needed)."""
    out = np.zeros(16, np.uint32)
    n = ctypes.c_uint32(0)
    if lib().mpcd_hip_tune_capacities(0 if algorithm == "srd" else 1, nmax, _p(out, _u32p), len(out),
                                       ctypes.byref(n)) != 0:
        raise RuntimeError("mpcd_hip_tune_capacities failed")
    return [int(v) for v in out[: n.value]]


def bench_copy(nbytes: int, reps: int = 10) -> float:
    """Milliseconds per device-to-device copy of nbytes (read nbytes + write nbytes)."""
    ms = ctypes.c_double()
    if lib().mpcd_hip_bench_copy(nbytes, reps, ctypes.byref(ms)) != 0:
        raise RuntimeError("copy benchmark failed")
    return ms.value


def selftest_philox(k0: int, k1: int, ctr: np.ndarray) -> np.ndarray:
    ctr = np.ascontiguousarray(ctr, dtype=np.uint32).reshape(-1, 4)
    out = np.empty_like(ctr)
    if lib().mpcd_hip_selftest_philox(k0, k1, _p(ctr, _u32p), _p(out, _u32p), len(ctr)) != 0:
        raise RuntimeError("device self-test failed")
    return out


def selftest_math(x: np.ndarray, which: int) -> np.ndarray:
    """which: 0 log, 1 sincospi, 2 exp, 3 sqrt, 4 fmod/wrap (L=7), 5 reciprocal / divide by 3."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    out = np.empty((len(x), 2), np.float32)
    if lib().mpcd_hip_selftest_math(_p(x, _f32p), _p(out, _f32p), len(x), which) != 0:
        raise RuntimeError("device self-test failed")
    return out
