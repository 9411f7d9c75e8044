"""The drop-in route inside the reference's tree:
HipBackend (cudampcd_amd/csrc/hip_backend/) compiled against the reference's
own mpcd::Backend and SimulationParameters, registered by one branch in the
reference's backend factory (src/libmpcd/common/backend_factory.hpp:15-27),
and reached through the reference's own SimulationHandle and pybind11 module.

oracle/Makefile builds that module into oracle/_ref/hip/ from the reference's
sources where they lie (bindings.cpp, simulator_handle.cpp,
simulation_parameters.cpp, the CPU backend) plus hip_backend.cpp, with the
factory patched by oracle/patch_factory.py.  Each module runs in its own
subprocess (two pybind11 modules registering classes of the same C++ names
cannot share one interpreter).
"""
from __future__ import annotations

import difflib
import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
REF = Path("/root/reference")
FACTORY = REF / "src" / "libmpcd" / "common" / "backend_factory.hpp"
HIP_MODULE = sorted((ROOT / "oracle" / "_ref" / "hip").glob("_pympcd*.so"))

needs_ref_sources = pytest.mark.skipif(not FACTORY.exists(), reason="reference sources absent")
needs_module = pytest.mark.skipif(not HIP_MODULE, reason="oracle/_ref/hip module not built (reference sources absent)")

DRIVER = r"""
import importlib.util, json, sys
spec = importlib.util.spec_from_file_location("_pympcd", sys.argv[1])
m = importlib.util.module_from_spec(spec)
spec.loader.exec_module(m)
out = {}
def make():
    p = m.Params()
    p.n = 5
    p.temperature = 1.0
    p.volume_size = (8, 8, 10)
    p.periodicity = (1, 1, 0)
    p.drag = 0.01
    p.delta_t = 0.02
    return p
for backend in sys.argv[2:]:
    try:
        s = m.Simulation(make(), backend)
        s.step(20)
        s.step_and_sample(30)
        d, v = s.get_mean_fields()
        pos = s.get_particle_positions()
        vel = s.get_particle_velocities()
        out[backend] = dict(ok=True, n=len(pos), shape_d=list(d.shape), shape_v=list(v.shape),
                            fluid_density=float(d[:, :, 1:-1].mean()), mean_vx=float(vel[:, 0].mean()),
                            var_v=float(vel.var()), max_abs_z=float(abs(pos[:, 2]).max()))
    except RuntimeError as e:
        out[backend] = dict(ok=False, error=str(e))
print(json.dumps(out))
"""


def run_module(*backends):
    res = subprocess.run([sys.executable, "-c", DRIVER, str(HIP_MODULE[0]), *backends], capture_output=True,
                         text=True, timeout=300, cwd="/tmp")
    assert res.returncode == 0, res.stderr[-2000:]
    return json.loads(res.stdout.strip().splitlines()[-1])


@needs_ref_sources
def test_hip_backend_compiles_against_the_reference_headers():
    """hip_backend.cpp, unchanged, against include/mpcd/api/simulation_parameters.hpp
    and src/libmpcd/common/backend.hpp of the reference (which lack
    rotation_angle, seed and system_id: the accessors fall back to the
    reference's fixed choices)."""
    cmd = ["g++", "-std=c++17", "-fsyntax-only", "-Wall", "-Werror", f"-I{REF / 'include'}",
           f"-I{REF / 'src' / 'libmpcd'}", f"-I{ROOT / 'include'}",
           str(ROOT / "cudampcd_amd" / "csrc" / "hip_backend" / "hip_backend.cpp")]
    res = subprocess.run(cmd, capture_output=True, text=True)
    assert res.returncode == 0, res.stderr


@needs_ref_sources
def test_factory_patch_adds_only_the_hip_branch():
    sys.path.insert(0, str(ROOT / "oracle"))
    from patch_factory import patch

    before = FACTORY.read_text().splitlines()
    after = patch(FACTORY.read_text()).splitlines()
    diff = [d for d in difflib.ndiff(before, after) if d[:1] in "+-"]
    assert all(d.startswith("+ ") for d in diff), diff
    added = "\n".join(d[2:] for d in diff)
    assert '#include "hip_backend/hip_backend.hpp"' in added
    assert 'backend_type == "hip"' in added and "std::make_unique<hip::HipBackend>(params)" in added
    assert len(diff) == 3 + 5


@needs_module
def test_reference_module_dispatches_to_the_hip_backend():
    """In the reference's own module, "cpu" still runs the reference CPU
    backend, an unknown name raises the reference's message, and "hip" reaches
    HipBackend: without a GPU its device query fails (the GPU run of the same
    module is tests/test_gpu_reference_module.py)."""
    from test_api_cpu import _has_gpu

    out = run_module("cpu", "cuda", "hip")
    assert out["cpu"]["ok"] and out["cpu"]["n"] == 5 * 8 * 8 * 10
    assert out["cuda"] == {"ok": False, "error": "Unsupported backend type"}
    if not _has_gpu():
        assert not out["hip"]["ok"] and "device" in out["hip"]["error"].lower()
