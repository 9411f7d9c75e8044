"""Host-side API and C-ABI checks that need no GPU.

* Params property semantics (setter side effects) against the reference's own
  pybind11 module (src/bindings/bindings.cpp:91-174) compiled in place into
  oracle/_ref, each module driven in its own subprocess.
* libmpcd_hip.so loads and exports every entry point include/mpcd_hip.h declares.
* The factory rejects unknown backends with the reference's message
  (backend_factory.hpp:25-26); invalid configurations fail with a clear error
  before any device work.
"""
from __future__ import annotations

import json
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]

SCRIPT = r"""
import importlib.util, json, sys
which = sys.argv[1]
if which == "ours":
    sys.path.insert(0, sys.argv[2])
    import cudampcd_amd._pympcd as m
else:
    spec = importlib.util.spec_from_file_location("_pympcd", sys.argv[2])
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)

def snap(p, with_n=True):
    d = dict(volume_size=list(p.volume_size), periodicity=list(p.periodicity), n=p.n,
             temperature=p.temperature, experiment=p.experiment, algorithm=p.algorithm,
             collision_kernel=p.collision_kernel, delta_t=p.delta_t, drag=p.drag,
             device_id=p.device_id)
    if with_n:
        d["N"] = p.N
    return d

out = []
p = m.Params()
out.append(snap(p, with_n=False))
p.n = 10; out.append(snap(p))
p.volume_size = (10, 10, 100); out.append(snap(p))
p.periodicity = (1, 1, 0); out.append(snap(p))
p.temperature = 2.0; out.append(snap(p))
p.experiment = "standart"; p.algorithm = "extended"; p.collision_kernel = "sorting"; out.append(snap(p))
p.experiment = "standard"; p.algorithm = "srd"; p.collision_kernel = "nonsense"; out.append(snap(p))
p.periodicity = (1, 0, 0); out.append(snap(p))
p.volume_size = (7, 9, 20); out.append(snap(p))
p.n = 3; out.append(snap(p))
q = m.Params(); q.periodicity = (1, 1, 0); q.volume_size = (12, 8, 30); q.n = 20; out.append(snap(q))
q.N = 12345; q.delta_t = 0.05; q.drag = 0.003; q.device_id = 0; out.append(snap(q))
try:
    q.volume_size = (1, 2)
    out.append("no error")
except Exception as e:
    out.append(type(e).__name__)
print(json.dumps(out))
"""


def _run(which, target):
    res = subprocess.run([sys.executable, "-c", SCRIPT, which, str(target)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr
    return json.loads(res.stdout.strip().splitlines()[-1])


def test_params_semantics_match_reference_module():
    ref = sorted((ROOT / "oracle" / "_ref").glob("_pympcd*.so"))
    if not ref:
        pytest.skip("reference pybind module not built (reference sources absent)")
    ours = _run("ours", ROOT)
    theirs = _run("ref", ref[0])
    assert ours == theirs


def test_c_abi_exports_every_declared_symbol():
    from cudampcd_amd import capi

    lib = capi.lib()
    names = capi.declared_symbols()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def _has_gpu():
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
        n = ctypes.c_int(0)
        return hip.hipGetDeviceCount(ctypes.byref(n)) == 0 and n.value > 0
    except OSError:
        return False


def test_unknown_backend_rejected(pympcd):
    p = pympcd.Params()
    for name in ("cpu", "cuda", "HIP", ""):
        with pytest.raises(RuntimeError, match="Unsupported backend type"):
            pympcd.Simulation(p, name)
    with pytest.raises(TypeError):
        pympcd.Simulation(p)


def test_invalid_configurations_fail_cleanly(pympcd):
    p = pympcd.Params()
    p.experiment = "channel"
    p.periodicity = (1, 1, 1)
    p.volume_size = (2, 6, 6)
    with pytest.raises(RuntimeError, match="more than 2 cells along x"):
        pympcd.Simulation(p, "hip")
    p = pympcd.Params()
    p.periodicity = (0, 1, 1)
    with pytest.raises(RuntimeError, match="walls in z only"):
        pympcd.Simulation(p, "hip")
    p = pympcd.Params()
    p.volume_size = (10.5, 10, 10)
    with pytest.raises(RuntimeError, match="whole numbers"):
        pympcd.Simulation(p, "hip")
    p = pympcd.Params()
    p.volume_size = (2 ** 21 + 1, 1, 10)   # the wrap's exact fast path needs 2.5 L below 2^23
    with pytest.raises(RuntimeError, match="2\\^21"):
        pympcd.Simulation(p, "hip")
    p = pympcd.Params()
    p.N = 0                                  # an empty system: rejected, not launched with empty grids
    with pytest.raises(RuntimeError, match="N must be positive"):
        pympcd.Simulation(p, "hip")


@pytest.mark.skipif(_has_gpu(), reason="a GPU is present")
def test_no_device_is_an_error_not_a_fallback(pympcd):
    p = pympcd.Params()
    p.n = 2
    p.volume_size = (4, 4, 4)
    with pytest.raises(RuntimeError, match="device"):
        pympcd.Simulation(p, "hip")


def test_params_extensions(pympcd):
    p = pympcd.Params()
    assert p.rotation_angle == 120.0
    assert p.seed is None
    p.seed = 2 ** 63 + 5
    assert p.seed == 2 ** 63 + 5
    p.seed = None
    assert p.seed is None
    p.system_id = 7
    assert p.system_id == 7


def test_theory_numbers():
    """BASELINE theory table: eta_c(n=10)=37.500, eta_c(n=20)=79.167, with kinetic 37.585 / 79.318."""
    from cudampcd_amd.physics import srd_viscosity

    assert abs(srd_viscosity(10, 0.02) - 37.5) < 5e-3
    assert abs(srd_viscosity(20, 0.02) - 79.167) < 5e-3
    assert abs(srd_viscosity(5, 0.02) - 16.695) < 5e-3
    assert abs(srd_viscosity(10, 0.02, kinetic=True) - 37.585) < 5e-3
    assert abs(srd_viscosity(20, 0.02, kinetic=True) - 79.318) < 5e-3


def test_ensemble_validation(pympcd):
    a = pympcd.Params()
    a.n = 2
    a.volume_size = (4, 4, 4)
    b = pympcd.Params()
    b.n = 2
    b.volume_size = (4, 4, 4)
    b.algorithm = "extended"
    with pytest.raises(RuntimeError, match="same algorithm"):
        pympcd.Ensemble([a, b])
    with pytest.raises(RuntimeError, match="at least one system"):
        pympcd.Ensemble([])
    c = pympcd.Params()
    c.periodicity = (0, 1, 1)
    with pytest.raises(RuntimeError, match="system 1: .*walls in z only"):
        pympcd.Ensemble([a, c])


@pytest.mark.skipif(_has_gpu(), reason="a GPU is present")
def test_ensemble_without_device_is_an_error(pympcd):
    a = pympcd.Params()
    a.n = 2
    a.volume_size = (4, 4, 4)
    with pytest.raises(RuntimeError, match="device"):
        pympcd.Ensemble([a, a])


def test_autotuner_tries_only_capacities_that_hold_a_collision_table():
    """A trial capacity below 4n would meet a cell fuller than the batch whose
    table the big-cell launch cannot collide either, and fail the step; the
    tuner skips such capacities (and the extended rule's capacities below 512
    or without room for one wall cell's n ghosts)."""
    from cudampcd_amd import capi

    compiled = [256, 384, 448, 512, 640, 1152, 2048]
    assert capi.tune_capacities("srd", 10) == compiled
    assert capi.tune_capacities("srd", 100) == [448, 512, 640, 1152, 2048]
    assert capi.tune_capacities("srd", 150) == [640, 1152, 2048]
    assert capi.tune_capacities("srd", 512) == [2048]
    for n in (1, 10, 64, 100, 150, 300, 512):
        for algo in ("srd", "extended"):
            caps = capi.tune_capacities(algo, n)
            assert caps and all(c >= 4 * n and c in compiled for c in caps), (algo, n, caps)
            if algo == "extended":
                assert min(caps) >= 512
    assert capi.tune_capacities("extended", 10) == [512, 640, 1152, 2048]
