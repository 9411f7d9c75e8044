"""Writes tests/golden/c1_reference_cpu.json: the plain SRD Poiseuille case on
the reference CPU backend (cpu_backend.cpp + simulation_functions.cpp compiled
in place into oracle/_ref/libref.so by oracle/Makefile).

Case: 10x10x100 fluid cells (10x10x102 stored), walls in z, n = 10, T = 1,
g = 1e-3, dt = 0.02, rotation 120 degrees (the reference README's snippet,
README.md:52-67), seeded Xoshiro128+ stream.

Parts (each in its own process):
  c1      1000 warm-up + 10,000 sampled steps on the reference, timed on one
          core (the reference backend is single-threaded): particle-steps/s,
          viscosity from the reference's own getSampleMean.
  readme  1000 warm-up + 100,000 sampled steps on the reference: the README's
          protocol, whose printed result is 37.27 against 37.50.
  acc_c1, acc_readme
          the same two runs on the oracle in reference mode, which follows the
          reference trajectory bit for bit and accumulates the per-step means
          both as the reference does (FP32 Kahan, cpu_backend.cpp:85-107) and
          in FP64: the difference between the two is the error of the
          reference's FP32 accumulation at that length.

Each reference run is repeated for six seeds (the first also drives the
accumulation parts); timings are those of the first seed.

Run from the repository root (about 70 minutes on 6 cores):
    python tests/golden/make_c1_fixture.py
"""
from __future__ import annotations

import json
import os
import platform
import sys
import time
from multiprocessing import Pool
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path[:0] = [str(ROOT / "tests"), str(ROOT)]

VOLUME = (10, 10, 102)
PERIODIC = (1, 1, 0)
N_DENS, DRAG, DT = 10, 0.001, 0.02
SEED = (0x5EED0001, 0x0C1C1C1C)
# further realisations of the reference runs: the viscosity estimate of one run
# scatters by about 4% (1000 + 10,000 steps) and 0.6% (1000 + 100,000)
SEEDS = [SEED] + [(0x5EED0001 + k, 0x0C1C1C1C ^ (k * 0x9E3779B9 & 0xFFFFFFFF)) for k in range(1, 6)]
WARMUP = 1000
SAMPLES = {"c1": 10_000, "readme": 100_000}
OUT = Path(__file__).resolve().parent / "c1_reference_cpu.json"


def eta_of(vel_flat):
    """README.md:80-102: eta = L^2 n g / (8 max_z <v_x>_{x,y}), L = 100."""
    X, Y, Z = VOLUME
    prof = vel_flat.reshape(Z, Y, X, 3)[..., 0].astype(np.float64).mean(axis=(1, 2))
    L = Z - 2
    return L * L * N_DENS * DRAG / (8.0 * prof.max()), prof


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def run(task):
    from oracle_lib import Oracle, Reference

    part, seed = task
    kw = dict(n=N_DENS, drag=DRAG, delta_t=DT, algorithm="srd", seed=seed)
    key = part.split("_")[-1]
    samples = SAMPLES[key]
    if part in SAMPLES:
        r = Reference(VOLUME, PERIODIC, **kw)
        t0 = time.perf_counter()
        r.step(WARMUP)
        r.step_and_sample(samples)
        t1 = time.perf_counter()
        d, v = r.sample_mean()
        eta, prof = eta_of(v)
        steps = WARMUP + samples
        return task, {"eta": eta, "profile_x": prof.tolist(), "mean_density_fluid": float(d.reshape(VOLUME[::-1])[1:-1].mean()),
                      "seconds": t1 - t0, "steps": steps, "particles": r.N,
                      "particle_steps_per_s": r.N * steps / (t1 - t0)}
    o = Oracle(VOLUME, PERIODIC, rng="reference", **kw)
    o.step(WARMUP)
    o.step_and_sample(samples)
    d32, v32 = o.sample_mean()
    d64, v64 = o.sample_mean_f64()
    eta32, prof32 = eta_of(v32)
    eta64, prof64 = eta_of(v64)
    dv = np.abs(v32.astype(np.float64) - v64)
    return task, {"eta_fp32_kahan": eta32, "eta_fp64": eta64,
                  "max_abs_velocity_diff": float(dv.max()), "rms_velocity_diff": float(np.sqrt((dv ** 2).mean())),
                  "max_abs_density_diff": float(np.abs(d32.astype(np.float64) - d64).max()),
                  "max_abs_profile_diff": float(np.abs(prof32 - prof64).max()),
                  "profile_x_fp32_kahan": prof32.tolist()}


def main():
    tasks = [("readme", sd) for sd in SEEDS] + [("acc_readme", SEED), ("acc_c1", SEED)] + [("c1", sd) for sd in SEEDS]
    with Pool(6) as pool:
        res = dict(pool.map(run, tasks))
    from cudampcd_amd.physics import srd_viscosity

    def runs(part):
        rs = [res[(part, sd)] for sd in SEEDS]
        etas = [r["eta"] for r in rs]
        first = dict(rs[0])
        first.update({"seeds": [list(sd) for sd in SEEDS], "etas": etas, "eta_mean": float(np.mean(etas)),
                      "eta_std": float(np.std(etas, ddof=1)),
                      "profile_x_mean": np.mean([r["profile_x"] for r in rs], axis=0).tolist()})
        return first

    out = {
        "case": {"volume_stored": VOLUME, "periodicity": PERIODIC, "n": N_DENS, "drag": DRAG, "delta_t": DT,
                 "temperature": 1.0, "rotation_deg": 120.0, "seed": list(SEED), "warmup": WARMUP},
        "theory_eta": srd_viscosity(N_DENS, DT),
        "host": {"cpu_model": cpu_model(), "nproc": os.cpu_count(), "threads_used": 1,
                 "compiler": "g++ -O3 -DNDEBUG -ffp-contract=off (oracle/Makefile)"},
        "reference_c1": runs("c1"), "reference_readme": runs("readme"),
        "accumulation_c1": res[("acc_c1", SEED)], "accumulation_readme": res[("acc_readme", SEED)],
    }
    # the oracle in reference mode reproduces the reference's own FP32 mean
    for k in ("c1", "readme"):
        assert np.array_equal(np.float64(out[f"reference_{k}"]["eta"]), np.float64(out[f"accumulation_{k}"]["eta_fp32_kahan"])), k
    OUT.write_text(json.dumps(out, indent=1) + "\n")
    print(json.dumps({k: (v if not isinstance(v, dict) else {a: b for a, b in v.items() if "profile" not in a})
                      for k, v in out.items()}, indent=1))


if __name__ == "__main__":
    main()
