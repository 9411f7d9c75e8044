"""Shared helpers for the GPU tests: build a Simulation the way the reference
README does, and compare a GPU step with the oracle."""
from __future__ import annotations

import numpy as np


def make_params(pympcd, fluid, periodicity=(1, 1, 0), n=10, algorithm="srd", seed=1234,
                drag=0.001, delta_t=0.02, temperature=1.0, angle=120.0, experiment="standard"):
    p = pympcd.Params()
    p.n = n
    p.temperature = temperature
    p.periodicity = tuple(periodicity)
    p.volume_size = tuple(fluid)          # wall layers are added by the setter
    p.drag = drag
    p.delta_t = delta_t
    p.experiment = experiment
    p.algorithm = algorithm
    p.rotation_angle = angle
    p.seed = seed
    return p


def stored_volume(p):
    return tuple(int(v) for v in p.volume_size)


def min_image(d, L):
    """Componentwise periodic minimum-image difference."""
    L = np.asarray(L, np.float64)
    return d - L * np.round(d / L)


def vel_error(a, b, thermal=1.0):
    """max |dv| / max(|v|, sqrt(T)), the norm-wise relative metric."""
    d = np.linalg.norm(a.astype(np.float64) - b.astype(np.float64), axis=1)
    s = np.maximum(np.linalg.norm(b.astype(np.float64), axis=1), thermal)
    return float((d / s).max()) if len(d) else 0.0
