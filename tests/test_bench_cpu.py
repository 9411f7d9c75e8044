"""The benchmark's host-side contract, without a GPU: it refuses more than one
GPU before touching the device, the ensemble workload's channel set is the
seeded one its docstring describes, and the measured HBM traffic it reports
comes from committed profiles of the same workloads."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))

import bench  # noqa: E402


@pytest.mark.parametrize("argv,world", [(["--gpus", "2"], "1"), (["--gpus", "8"], "1"), ([], "2")],
                         ids=["gpus2", "gpus8", "world2"])
def test_bench_measures_one_gpu_only(argv, world):
    env = dict(os.environ, WORLD_SIZE=world)
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), *argv], env=env, capture_output=True, text=True,
                       timeout=60)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_ensemble_channel_set():
    a = bench.c4_members(0x5EED0004)
    assert a == bench.c4_members(0x5EED0004)
    assert a != bench.c4_members(0x5EED0005)
    assert len(a) == 64
    assert len({m["seed"] for m in a}) == 64
    assert [m["system_id"] for m in a] == list(range(64))
    for m in a:
        x, y, z = m["fluid"]
        assert x == y and x in (32, 64, 128)
        assert 20 <= z <= 200
        assert m["n"] in (5, 10, 20)
        assert 5e-4 <= m["drag"] <= 2e-3
    # every cross-section and density occurs in the default set
    assert {m["fluid"][0] for m in a} == {32, 64, 128}
    assert {m["n"] for m in a} == {5, 10, 20}


def test_algorithmic_bytes_per_step():
    b = bench.algorithmic_bytes(400_000_000, 1000 * 1000 * 22)
    assert b["collide"] == 56 * 400_000_000 + 12 * 22_000_000
    assert b["perm"] == 8 * 400_000_000 + 8 * 22_000_000
    assert b["scan"] == 20 * 22_000_000
    assert b["step"] == b["collide"] + b["perm"] + b["scan"]


@pytest.mark.parametrize("config", ["c2", "c3"])
def test_committed_traffic_matches_the_workload(config):
    cfg = bench.CONFIGS[config]
    x, y, z = cfg["fluid"]
    particles = x * y * z * cfg["n"]
    t = bench.load_traffic(config)
    assert t is not None, f"profiles/k_collide_{config}_traffic.md has no machine-readable lines"
    assert t["particles"] == particles
    # measured HBM bytes of one launch: at least the compulsory bytes (the
    # working set is far beyond the caches) and at most half again as much
    alg = bench.algorithmic_bytes(particles, x * y * (z + 2))["collide"]
    assert 1.0 <= t["bytes_per_launch"] / alg <= 1.5
