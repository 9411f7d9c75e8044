"""ctypes access to the CPU oracle (oracle/liboracle.so) and to the reference
CPU backend compiled in place (oracle/_ref/libref.so).

Test infrastructure only: the oracle is the checker, never the thing tested.
"""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]
ORACLE_SO = ROOT / "oracle" / "liboracle.so"
REF_SO = ROOT / "oracle" / "_ref" / "libref.so"

c_float_p = ctypes.POINTER(ctypes.c_float)
c_uint_p = ctypes.POINTER(ctypes.c_uint)
c_double_p = ctypes.POINTER(ctypes.c_double)


class OrcParams(ctypes.Structure):
    _fields_ = [
        ("volume", ctypes.c_float * 3),
        ("periodic", ctypes.c_int * 3),
        ("n", ctypes.c_uint),
        ("N", ctypes.c_ulonglong),
        ("delta_t", ctypes.c_float),
        ("drag", ctypes.c_float),
        ("thermal_velocity", ctypes.c_float),
        ("rotation_deg", ctypes.c_float),
        ("algorithm", ctypes.c_int),
        ("rng_mode", ctypes.c_int),
        ("seed_a", ctypes.c_ulonglong),
        ("seed_b", ctypes.c_ulonglong),
        ("system", ctypes.c_uint),
        ("ghost_xy_order", ctypes.c_int),
        ("experiment", ctypes.c_int),
        ("sum_order", ctypes.c_int),
    ]


class RefParams(ctypes.Structure):
    _fields_ = [
        ("volume", ctypes.c_float * 3),
        ("periodic", ctypes.c_int * 3),
        ("n", ctypes.c_uint),
        ("N", ctypes.c_ulonglong),
        ("delta_t", ctypes.c_float),
        ("drag", ctypes.c_float),
        ("thermal_velocity", ctypes.c_float),
        ("algorithm", ctypes.c_int),
        ("seed_a", ctypes.c_ulonglong),
        ("seed_b", ctypes.c_ulonglong),
        ("experiment", ctypes.c_int),
    ]


def _ptr(a: np.ndarray, t):
    return a.ctypes.data_as(t)


_oracle = None
_ref = None


def oracle_lib() -> ctypes.CDLL:
    global _oracle
    if _oracle is None:
        lib = ctypes.CDLL(str(ORACLE_SO))
        lib.orc_create.restype = ctypes.c_void_p
        lib.orc_create.argtypes = [ctypes.POINTER(OrcParams)]
        vp = [ctypes.c_void_p]
        for name in ("orc_destroy", "orc_distribute"):
            getattr(lib, name).argtypes = vp
        lib.orc_n_particles.restype = ctypes.c_ulonglong
        lib.orc_n_particles.argtypes = vp
        lib.orc_n_cells.restype = ctypes.c_uint
        lib.orc_n_cells.argtypes = vp
        lib.orc_set_step.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong]
        lib.orc_get_step.restype = ctypes.c_ulonglong
        lib.orc_get_step.argtypes = vp
        lib.orc_set_particles.argtypes = [ctypes.c_void_p, c_float_p, c_float_p, ctypes.c_ulonglong]
        lib.orc_get_particles.argtypes = [ctypes.c_void_p, c_float_p, c_float_p]
        lib.orc_translate.argtypes = [ctypes.c_void_p, c_float_p]
        lib.orc_collide.argtypes = vp
        lib.orc_force_shift.argtypes = [ctypes.c_void_p, c_float_p]
        lib.orc_get_cells.restype = ctypes.c_ulonglong
        lib.orc_get_cells.argtypes = [ctypes.c_void_p, c_uint_p, c_uint_p, c_uint_p]
        lib.orc_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.orc_step_and_sample.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.orc_get_sample_mean.argtypes = [ctypes.c_void_p, c_float_p, c_float_p]
        lib.orc_get_sample_mean_f64.argtypes = [ctypes.c_void_p, c_double_p, c_double_p]
        lib.orc_sample_count.restype = ctypes.c_ulonglong
        lib.orc_sample_count.argtypes = vp
        lib.orc_philox.argtypes = [ctypes.c_uint, ctypes.c_uint, c_uint_p, c_uint_p]
        lib.orc_logf.restype = ctypes.c_float
        lib.orc_logf.argtypes = [ctypes.c_float]
        lib.orc_expf.restype = ctypes.c_float
        lib.orc_expf.argtypes = [ctypes.c_float]
        lib.orc_sincospi.argtypes = [ctypes.c_float, c_float_p, c_float_p]
        lib.orc_box_muller.argtypes = [ctypes.c_uint, ctypes.c_uint, c_float_p, c_float_p]
        lib.orc_xoshiro_uniform.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_ulonglong, c_float_p]
        lib.orc_xoshiro_gaussian.argtypes = [ctypes.c_ulonglong, ctypes.c_ulonglong, ctypes.c_ulonglong, c_float_p]
        lib.orc_counter_draws.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, c_float_p]
        lib.orc_ext_direction.argtypes = [ctypes.c_int, c_float_p]
        lib.orc_ghost_pos_draws.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                            ctypes.c_float, ctypes.c_int, ctypes.c_int, c_float_p]
        lib.orc_srd_ghost_sums.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint,
                                           c_float_p]
        c_int_p = ctypes.POINTER(ctypes.c_int)
        lib.orc_ext_axis_draws.argtypes = [ctypes.c_void_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_uint, c_int_p,
                                           c_int_p, c_float_p]
        lib.orc_sym_inverse.argtypes = [c_float_p, ctypes.c_float, c_float_p]
        lib.orc_set_force.argtypes = [ctypes.c_void_p, c_float_p]
        lib.orc_log_draws.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.orc_get_draws.argtypes = [ctypes.c_void_p, c_float_p, c_float_p, c_uint_p, c_float_p]
        lib.orc_get_ext_draws.argtypes = [ctypes.c_void_p, c_float_p, c_uint_p, c_float_p, c_uint_p, c_float_p]
        lib.orc_frozen_profile_loss.restype = ctypes.c_double
        lib.orc_frozen_profile_loss.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                                c_double_p, c_double_p, ctypes.c_uint, ctypes.c_uint, ctypes.c_int,
                                                c_double_p]
        _oracle = lib
    return _oracle


def ref_available() -> bool:
    return REF_SO.exists()


def ref_lib() -> ctypes.CDLL:
    global _ref
    if _ref is None:
        lib = ctypes.CDLL(str(REF_SO))
        lib.ref_create.restype = ctypes.c_void_p
        lib.ref_create.argtypes = [ctypes.POINTER(RefParams)]
        vp = [ctypes.c_void_p]
        lib.ref_destroy.argtypes = vp
        lib.ref_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.ref_step_and_sample.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.ref_get_sample_mean.argtypes = [ctypes.c_void_p, c_float_p, c_float_p]
        lib.ref_get_particles.argtypes = [ctypes.c_void_p, c_float_p, c_float_p]
        lib.ref_set_particles.argtypes = [ctypes.c_void_p, c_float_p, c_float_p, ctypes.c_ulonglong]
        lib.ref_reseed.argtypes = [ctypes.c_void_p, ctypes.c_ulonglong, ctypes.c_ulonglong]
        lib.ref_translate.argtypes = [ctypes.c_void_p, c_float_p]
        lib.ref_collide.argtypes = vp
        lib.ref_get_cells.restype = ctypes.c_ulonglong
        lib.ref_get_cells.argtypes = [ctypes.c_void_p, c_uint_p, c_uint_p, c_uint_p]
        lib.ref_n_particles.restype = ctypes.c_ulonglong
        lib.ref_n_particles.argtypes = vp
        _ref = lib
    return _ref


def fluid_count(volume, periodic, n):
    """Particle count the reference Params setters produce (bindings.cpp:125-131)."""
    v = [volume[i] - 2 * (1 - periodic[i]) for i in range(3)]
    return int(n * v[0] * v[1] * v[2])


class Oracle:
    """One oracle simulation.  `volume` is the stored grid (wall layers included).

    `order`: per-cell summation order, "serial" (the reference's) or "lanes"
    (the GPU's lane groups); default "lanes" with the counter stream (the GPU's
    draws) and "serial" with the reference stream."""

    def __init__(self, volume, periodic=(1, 1, 0), n=10, N=None, delta_t=0.02, drag=0.0,
                 thermal_velocity=1.0, rotation_deg=120.0, algorithm="srd", rng="counter",
                 seed=(12345, 67890), system=0, ghost_xy_order=0, distribute=True, experiment="standard",
                 order=None):
        self.lib = oracle_lib()
        p = OrcParams()
        for i in range(3):
            p.volume[i] = float(volume[i])
            p.periodic[i] = int(periodic[i])
        p.n = n
        p.N = fluid_count(volume, periodic, n) if N is None else N
        p.delta_t = delta_t
        p.drag = drag
        p.thermal_velocity = thermal_velocity
        p.rotation_deg = rotation_deg
        p.algorithm = 0 if algorithm == "srd" else 1
        p.rng_mode = 0 if rng == "reference" else 1
        p.seed_a, p.seed_b = seed
        p.system = system
        p.ghost_xy_order = ghost_xy_order
        p.experiment = 1 if experiment == "channel" else 0
        if order is None:
            order = "lanes" if rng == "counter" else "serial"
        assert order in ("serial", "lanes"), order
        p.sum_order = 1 if order == "lanes" else 0
        self.params = p
        self.h = self.lib.orc_create(ctypes.byref(p))
        self.N = int(p.N)
        self.C = int(self.lib.orc_n_cells(self.h))
        if distribute:
            self.lib.orc_distribute(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.orc_destroy(self.h)
            self.h = None

    def set_particles(self, pos, vel):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        self.N = len(pos)
        self.lib.orc_set_particles(self.h, _ptr(pos, c_float_p), _ptr(vel, c_float_p), self.N)

    def particles(self):
        pos = np.empty((self.N, 3), np.float32)
        vel = np.empty((self.N, 3), np.float32)
        self.lib.orc_get_particles(self.h, _ptr(pos, c_float_p), _ptr(vel, c_float_p))
        return pos, vel

    def set_step(self, s):
        self.lib.orc_set_step(self.h, s)

    def translate(self):
        sh = np.empty(3, np.float32)
        self.lib.orc_translate(self.h, _ptr(sh, c_float_p))
        return sh

    def collide(self):
        self.lib.orc_collide(self.h)

    def force_shift(self, shift):
        if shift is None:
            self.lib.orc_force_shift(self.h, None)
        else:
            self._forced = np.ascontiguousarray(shift, dtype=np.float32)
            self.lib.orc_force_shift(self.h, _ptr(self._forced, c_float_p))

    def draws(self, purpose, step, cell, count):
        """Counter-mode draws: purpose 7 gaussians, 2 ghost-count uniforms, 5 axes (3 per cell)."""
        out = np.empty(count * (3 if purpose == 5 else 1), np.float32)
        self.lib.orc_counter_draws(self.h, purpose, step, cell, count, _ptr(out, c_float_p))
        return out.reshape(count, 3) if purpose == 5 else out

    def ghost_positions(self, step, cell, count, shift, layer, sign):
        """`count` wall-ghost positions of one cell, each coordinate in [0, 1]
        (reference mode: the stream's rejection loop; counter mode: ghost k = i)."""
        out = np.empty((count, 3), np.float32)
        self.lib.orc_ghost_pos_draws(self.h, step, cell, count, float(shift), int(layer), int(sign),
                                     _ptr(out, c_float_p))
        return out

    def srd_ghost_sums(self, step, cell0, count, added):
        """What the SRD rule adds to the velocity sum of wall cells [cell0,
        cell0 + count) holding `added` ghosts each: counter mode the one-draw
        sum, reference mode `added` thermal velocities drawn one by one and
        summed (the reference's construction)."""
        out = np.empty((count, 3), np.float32)
        self.lib.orc_srd_ghost_sums(self.h, step, cell0, count, added, _ptr(out, c_float_p))
        return out

    def ext_axis_draws(self, step, cell0, count):
        """(direction index 0..12, sign bit, gate uniform) of cells [cell0, cell0 + count)."""
        sel = np.empty(count, np.int32)
        flip = np.empty(count, np.int32)
        gate = np.empty(count, np.float32)
        ip = ctypes.POINTER(ctypes.c_int)
        self.lib.orc_ext_axis_draws(self.h, step, cell0, count, _ptr(sel, ip), _ptr(flip, ip), _ptr(gate, c_float_p))
        return sel, flip, gate

    def cells(self):
        cell = np.empty(self.N, np.uint32)
        cnt = np.empty(self.C, np.uint32)
        lst = np.empty(self.N, np.uint32)
        k = self.lib.orc_get_cells(self.h, _ptr(cell, c_uint_p), _ptr(cnt, c_uint_p), _ptr(lst, c_uint_p))
        return cell, cnt, lst[:k]

    def step(self, k=1):
        self.lib.orc_step(self.h, k)

    def step_and_sample(self, k=1):
        self.lib.orc_step_and_sample(self.h, k)

    def sample_mean(self):
        d = np.empty(self.C, np.float32)
        v = np.empty(3 * self.C, np.float32)
        self.lib.orc_get_sample_mean(self.h, _ptr(d, c_float_p), _ptr(v, c_float_p))
        return d, v

    def sample_mean_f64(self):
        d = np.empty(self.C, np.float64)
        v = np.empty(3 * self.C, np.float64)
        self.lib.orc_get_sample_mean_f64(self.h, _ptr(d, c_double_p), _ptr(v, c_double_p))
        return d, v

    def log_draws(self, on=True):
        """Record the draws of the SRD collisions from now on (see last_draws)."""
        self.lib.orc_log_draws(self.h, int(on))

    def last_draws(self):
        """Draws of the last logged SRD collision: shift (3,), axes (C, 3) in
        reference cell order (zero where no collision), ghost counts (2 X Y,)
        and ghost velocities (2 X Y, n, 3) of the wall-layer cells
        w = (upper layer ? X Y : 0) + x + X y."""
        X, Y = int(self.params.volume[0]), int(self.params.volume[1])
        n = int(self.params.n)
        shift = np.empty(3, np.float32)
        axes = np.empty((self.C, 3), np.float32)
        gcnt = np.empty(2 * X * Y, np.uint32)
        gvel = np.empty((2 * X * Y, n, 3), np.float32)
        self.lib.orc_get_draws(self.h, _ptr(shift, c_float_p), _ptr(axes, c_float_p), _ptr(gcnt, c_uint_p),
                               _ptr(gvel, c_float_p))
        return shift, axes, gcnt, gvel

    def last_ext_draws(self):
        """Draws of the last logged extended-rule collision: shift (3,), per
        cell (direction index, sign, gate uniform as float32 bits) (C, 3), the
        members' thermal Gaussians (C, 4n), ghost counts (2 X Y,) and ghost
        positions and velocities (2 X Y, n, 6) of the wall-layer cells."""
        X, Y = int(self.params.volume[0]), int(self.params.volume[1])
        n = int(self.params.n)
        shift = np.empty(3, np.float32)
        ext = np.empty((self.C, 3), np.uint32)
        gauss = np.empty((self.C, 4 * n), np.float32)
        gcnt = np.empty(2 * X * Y, np.uint32)
        gvel = np.empty((2 * X * Y, n, 6), np.float32)
        self.lib.orc_get_ext_draws(self.h, _ptr(shift, c_float_p), _ptr(ext, c_uint_p), _ptr(gauss, c_float_p),
                                   _ptr(gcnt, c_uint_p), _ptr(gvel, c_float_p))
        return shift, ext, gauss, gcnt, gvel

    def set_force(self, force):
        """x body force per cell (reference cell order) applied in the SRD rule; None removes it."""
        if force is None:
            self.lib.orc_set_force(self.h, None)
        else:
            self._force = np.ascontiguousarray(force, dtype=np.float32)
            self.lib.orc_set_force(self.h, _ptr(self._force, c_float_p))

    def frozen_profile_loss(self, steps, drag, angle_deg, target, zlo, zhi, force=None, use_real=False):
        """Profile loss of `steps` steps from the current state with the cells, tables,
        bounces and draws of the oracle's own trajectory held fixed and the velocities
        carried in float64 with the given drag, angle and force (see mpcd_oracle.cpp).
        Returns (loss, profile); the oracle's state is unchanged."""
        t = np.ascontiguousarray(target, dtype=np.float64)
        prof = np.empty(zhi - zlo, np.float64)
        f = None if force is None else np.ascontiguousarray(force, dtype=np.float64)
        loss = self.lib.orc_frozen_profile_loss(self.h, steps, drag, angle_deg,
                                                None if f is None else _ptr(f, c_double_p),
                                                _ptr(t, c_double_p), zlo, zhi, int(use_real), _ptr(prof, c_double_p))
        return loss, prof


class Reference:
    """The reference CPU backend (cpu_backend.cpp) compiled in place, seeded."""

    def __init__(self, volume, periodic=(1, 1, 0), n=10, N=None, delta_t=0.02, drag=0.0,
                 thermal_velocity=1.0, algorithm="srd", seed=(12345, 67890), experiment="standard"):
        self.lib = ref_lib()
        p = RefParams()
        for i in range(3):
            p.volume[i] = float(volume[i])
            p.periodic[i] = int(periodic[i])
        p.n = n
        p.N = fluid_count(volume, periodic, n) if N is None else N
        p.delta_t = delta_t
        p.drag = drag
        p.thermal_velocity = thermal_velocity
        p.algorithm = 0 if algorithm == "srd" else 1
        p.seed_a, p.seed_b = seed
        p.experiment = 1 if experiment == "channel" else 0
        self.h = self.lib.ref_create(ctypes.byref(p))
        self.N = int(p.N)
        self.C = int(np.prod([int(x) for x in volume]))

    def __del__(self):
        if getattr(self, "h", None):
            self.lib.ref_destroy(self.h)
            self.h = None

    def particles(self):
        pos = np.empty((self.N, 3), np.float32)
        vel = np.empty((self.N, 3), np.float32)
        self.lib.ref_get_particles(self.h, _ptr(pos, c_float_p), _ptr(vel, c_float_p))
        return pos, vel

    def set_particles(self, pos, vel):
        pos = np.ascontiguousarray(pos, dtype=np.float32).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float32).reshape(-1, 3)
        self.N = len(pos)
        self.lib.ref_set_particles(self.h, _ptr(pos, c_float_p), _ptr(vel, c_float_p), self.N)

    def reseed(self, a, b):
        self.lib.ref_reseed(self.h, a, b)

    def translate(self):
        sh = np.empty(3, np.float32)
        self.lib.ref_translate(self.h, _ptr(sh, c_float_p))
        return sh

    def collide(self):
        self.lib.ref_collide(self.h)

    def cells(self):
        cell = np.empty(self.N, np.uint32)
        cnt = np.empty(self.C, np.uint32)
        lst = np.empty(self.N, np.uint32)
        k = self.lib.ref_get_cells(self.h, _ptr(cell, c_uint_p), _ptr(cnt, c_uint_p), _ptr(lst, c_uint_p))
        return cell, cnt, lst[:k]

    def step(self, k=1):
        self.lib.ref_step(self.h, k)

    def step_and_sample(self, k=1):
        self.lib.ref_step_and_sample(self.h, k)

    def sample_mean(self):
        d = np.empty(self.C, np.float32)
        v = np.empty(3 * self.C, np.float32)
        self.lib.ref_get_sample_mean(self.h, _ptr(d, c_float_p), _ptr(v, c_float_p))
        return d, v
