"""Closed-form SRD transport coefficients and the Poiseuille viscosity estimate.

Formulas:
  collisional viscosity  eta_c = (1 - cos a) / (18 dt) * (n - 1 + exp(-n))
      (the formula the reference prints, README.md:98 of azantop/cudaMPCD)
  kinetic viscosity      eta_k = n T dt [ 5 n / ((n - 1 + exp(-n)) (4 - 2 cos a - 2 cos 2a)) - 1/2 ]
      (standard SRD kinetic contribution, Kikuchi et al. 2003; Gompper et al. 2009)
  Poiseuille estimate    eta = rho g L^2 / (8 v_max), rho = n
      (README.md:80-102)
"""
from __future__ import annotations

import math

import numpy as np


def srd_viscosity(n: float, delta_t: float, angle_deg: float = 120.0, temperature: float = 1.0,
                  kinetic: bool = False) -> float:
    a = math.radians(angle_deg)
    eta_c = (1.0 - math.cos(a)) / (18.0 * delta_t) * (n - 1.0 + math.exp(-n))
    if not kinetic:
        return eta_c
    eta_k = n * temperature * delta_t * (
        5.0 * n / ((n - 1.0 + math.exp(-n)) * (4.0 - 2.0 * math.cos(a) - 2.0 * math.cos(2.0 * a))) - 0.5)
    return eta_c + eta_k


def poiseuille_viscosity(velocity: np.ndarray, n: float, drag: float, wall_gap: float) -> float:
    """eta from the x-velocity profile across z (cells along the last spatial axis)."""
    profile = velocity[..., 0].mean(axis=(0, 1))
    return wall_gap * wall_gap * n * drag / (8.0 * float(profile.max()))
