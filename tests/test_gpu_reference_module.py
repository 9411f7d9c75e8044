"""The reference's own pybind11 module, with the one-branch factory patch of
INTEGRATION.md and HipBackend linked in (oracle/_ref/hip/, built by
oracle/Makefile), runs the HIP backend on the GPU: Simulation(params, "hip")
through the reference's SimulationHandle and bindings, beside
Simulation(params, "cpu") of the same module on the same parameters.

The two backends draw different random numbers, so their results agree in
distribution: the same particle count and array shapes, fluid density n
(times the reference's S/(S-1) sample normalisation), temperature, particles
inside the walls.
"""
from __future__ import annotations

import pytest

from test_reference_integration import HIP_MODULE, run_module

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not HIP_MODULE, reason="oracle/_ref/hip module not built (reference sources absent)")]


def test_reference_module_runs_hip_beside_cpu():
    out = run_module("hip", "cpu")
    hip, cpu = out["hip"], out["cpu"]
    assert hip["ok"], hip
    assert cpu["ok"], cpu
    assert hip["n"] == cpu["n"] == 5 * 8 * 8 * 10
    assert hip["shape_d"] == cpu["shape_d"] == [8, 8, 12]
    assert hip["shape_v"] == cpu["shape_v"] == [8, 8, 12, 3]
    want = 5 * 30 / 29                     # 30 samples divided by 29 (cpu_backend.cpp:111-122)
    for r in (hip, cpu):
        assert abs(r["fluid_density"] - want) < 0.02 * want
        assert abs(r["var_v"] - 1.0) < 0.08
        # inside the walls (z_wall = 5), or half a cell beyond for a particle
        # left at its shifted position alone in a wall-layer cell
        assert r["max_abs_z"] <= 5.5
        assert abs(r["mean_vx"]) < 0.1         # thermal noise of 3200 particles is 0.018
