"""Pathwise adjoint of a time-averaged velocity-profile loss (no reference counterpart).

For a single-system SRD Simulation, ``steps`` steps from its current state
with drag g, rotation angle alpha (degrees) and an x body force f per
reference cell (v_x += g dt + f dt for colliding particles) define

    L = mean over fluid z layers of (<vbar_x>(z) - target(z))^2,

<vbar_x>(z) being the x velocity averaged over the particles of layer z after
each step and then over the steps.  The gradient (dL/dg, dL/dalpha per
degree, dL/df per cell) holds the trajectory's cells, cell tables, wall
bounces, rotation axes and ghosts fixed, under which one step is linear in
the velocities; HIP kernels run the reverse steps over a tape recomputed from
checkpoints every ``checkpoint_every`` steps.  The simulation's state is left
unchanged, so repeated calls follow the same trajectory (common random
numbers for finite differences).

``ProfileLoss`` exposes the same computation as a torch.autograd.Function of
(drag, angle, force) tensors.  The force field and its gradient cross the
C-ABI as host arrays (torch keeps its own HIP runtime; see include/mpcd_hip.h).
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import capi

_f64p = ctypes.POINTER(ctypes.c_double)
_f32p = ctypes.POINTER(ctypes.c_float)


@dataclass
class ProfileResult:
    loss: float
    profile: np.ndarray
    grad_drag: float | None = None
    grad_angle: float | None = None
    grad_force: np.ndarray | None = None


def fluid_layers(params) -> int:
    """Number of fluid z layers (the target's length) of a Params object."""
    z = int(params.volume_size[2])
    return z if params.periodicity[2] else z - 2


def _f64(a):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float64).reshape(-1))


def _force(h, force):
    if force is None:
        return None
    f = np.ascontiguousarray(np.asarray(force, dtype=np.float32).reshape(-1))
    if f.size != h.C:
        raise ValueError(f"force holds {f.size} values, the grid has {h.C} cells")
    return f


def profile_loss(sim, steps: int, drag: float, rotation_angle: float, target, force=None) -> ProfileResult:
    h = capi.Handle(sim)
    t = _f64(target)
    f = _force(h, force)
    prof = np.empty(t.size, np.float64)
    loss = ctypes.c_double()
    h._check(h.L.mpcd_hip_profile_loss(h.h, steps, drag, rotation_angle,
                                       None if f is None else f.ctypes.data_as(_f32p),
                                       t.ctypes.data_as(_f64p), t.size, ctypes.byref(loss), prof.ctypes.data_as(_f64p)))
    return ProfileResult(loss.value, prof)


def profile_loss_grad(sim, steps: int, checkpoint_every: int, drag: float, rotation_angle: float, target,
                      force=None, want_force_grad: bool = True) -> ProfileResult:
    h = capi.Handle(sim)
    t = _f64(target)
    f = _force(h, force)
    prof = np.empty(t.size, np.float64)
    gf = np.empty(h.C, np.float32) if want_force_grad else None
    loss, gd, ga = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
    h._check(h.L.mpcd_hip_profile_loss_grad(h.h, steps, checkpoint_every, drag, rotation_angle,
                                            None if f is None else f.ctypes.data_as(_f32p),
                                            t.ctypes.data_as(_f64p), t.size, ctypes.byref(loss),
                                            prof.ctypes.data_as(_f64p), ctypes.byref(gd), ctypes.byref(ga),
                                            None if gf is None else gf.ctypes.data_as(_f32p)))
    return ProfileResult(loss.value, prof, gd.value, ga.value, gf)


def _make_function():
    import torch

    class ProfileLoss(torch.autograd.Function):
        """loss = ProfileLoss.apply(drag, angle, force, sim, target, steps, checkpoint_every)

        drag, angle: 0-d tensors (angle in degrees); force: tensor of n_cells
        values (reference cell order) or None; target: one value per fluid z
        layer.  Returns a float64 0-d tensor on drag's device; gradients come
        back on each input's device and dtype."""

        @staticmethod
        def forward(ctx, drag, angle, force, sim, target, steps, checkpoint_every):
            f = None if force is None else force.detach().to("cpu", torch.float32).numpy()
            tg = target.detach().cpu().numpy() if torch.is_tensor(target) else target
            need = ctx.needs_input_grad[:3]
            if any(need):
                r = profile_loss_grad(sim, steps, checkpoint_every, float(drag), float(angle), tg, force=f,
                                      want_force_grad=force is not None)
                ctx.grads = (r.grad_drag, r.grad_angle, r.grad_force)
            else:
                r = profile_loss(sim, steps, float(drag), float(angle), tg, force=f)
                ctx.grads = None
            ctx.inputs = (drag, angle, force)
            dev = drag.device if torch.is_tensor(drag) else torch.device("cpu")
            return torch.tensor(r.loss, dtype=torch.float64, device=dev)

        @staticmethod
        def backward(ctx, gout):
            gd, ga, gf = ctx.grads
            drag, angle, force = ctx.inputs
            g = float(gout)
            out_d = torch.tensor(g * gd, dtype=drag.dtype, device=drag.device) if torch.is_tensor(drag) else None
            out_a = torch.tensor(g * ga, dtype=angle.dtype, device=angle.device) if torch.is_tensor(angle) else None
            out_f = None
            if force is not None and gf is not None:
                out_f = (torch.from_numpy(gf.astype(np.float64) * g)).to(force.device, force.dtype)
            return out_d, out_a, out_f, None, None, None, None

    return ProfileLoss


def __getattr__(name):
    if name == "ProfileLoss":
        fn = _make_function()
        globals()["ProfileLoss"] = fn
        return fn
    raise AttributeError(name)
