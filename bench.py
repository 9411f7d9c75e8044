#!/usr/bin/env python3
"""Throughput of the MPCD time step on one MI355X.

    python bench.py [--steps K] [--warmup W] [--config c2|c3|c4|c5] [--no-cpu]

Default workload (c2): plain SRD, 1000x1000x20 fluid cells with walls in z
(1000x1000x22 stored cells), n=20 -> 4e8 particles, T=1, dt=0.02, g=1e-3,
rotation 120 deg, counting sort into cell order every step; seeded uniform
positions and Maxwell-Boltzmann velocities (synthetic).  c3: extended MPC,
500x500x40 fluid cells, n=10 (1e8 particles).  c4: 64 independent SRD
Poiseuille channels drawn from the seed (cross-section 32^2, 64^2 or 128^2
cells, wall gap uniform in [20, 200], n in {5, 10, 20}, g uniform in
[5e-4, 2e-3], distinct seeds) advanced as one Ensemble (autotuned tiles,
captured step graph), timed against a plain loop over 64 Simulation objects
with the same seeds (identical per-channel results, so equal accuracy).
c5: pathwise adjoint gradient of a time-averaged velocity-profile loss (SRD
64x64x50 cells, n=10, 200 steps, checkpoint every 20) against one loss
evaluation, with the whole tape resident as well, and the accuracy study of
the drag gradient against central differences.

A step is one shift + stream + sort + collide pass over all particles, with
the particle state resident in HBM.  Prints one JSON line; the default (c2)
line also carries the sampled c2 step, c3, c4 and c5 under "extra", so that
every workload is timed in the same run.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
# reference optimized step, README.md:35-41 (94 ms/step, 1000x1000x20, n=20), counted over 4e8 fluid particles
REFERENCE_PSTEPS = 4.0e8 / 0.094

CONFIGS = {
    "c2": dict(fluid=(1000, 1000, 20), n=20, algorithm="srd",
               workload="SRD Poiseuille channel 1000x1000x20 cells, n=20, walls in z, sort every step"),
    "c3": dict(fluid=(500, 500, 40), n=10, algorithm="extended",
               workload="extended MPC 500x500x40 cells, n=10, walls in z, sort every step"),
    "c5": dict(fluid=(64, 64, 50), n=10, algorithm="srd", steps=200, checkpoint=20,
               workload="pathwise adjoint of the time-averaged velocity-profile loss, SRD 64x64x50 cells, n=10, "
                        "200 steps, checkpoint every 20"),
    "c4": dict(channels=64, algorithm="srd",
               workload="ensemble of 64 SRD Poiseuille channels, cross-section {32,64,128}^2, "
                        "wall gap U[20,200], n in {5,10,20}, g U[5e-4,2e-3], one launch per kernel per step"),
}


def c4_members(seed: int, channels: int = 64) -> list[dict]:
    """The seeded channel set of the ensemble workload."""
    import numpy as np
    rng = np.random.default_rng(seed)
    out = []
    for i in range(channels):
        a = int(rng.choice([32, 64, 128]))
        lz = int(rng.integers(20, 201))
        n = int(rng.choice([5, 10, 20]))
        g = float(rng.uniform(5e-4, 2e-3))
        out.append(dict(fluid=(a, a, lz), n=n, drag=g, seed=seed + 1 + i, system_id=i))
    return out


def make_params(pympcd, fluid, n, algorithm, drag, seed, system_id=0):
    p = pympcd.Params()
    p.n = n
    p.temperature = 1.0
    p.periodicity = (1, 1, 0)
    p.volume_size = fluid
    p.drag = drag
    p.delta_t = 0.02
    p.experiment = "standard"
    p.algorithm = algorithm
    p.seed = seed
    p.system_id = system_id
    return p


def algorithmic_bytes(n_particles: int, n_cells: int) -> dict:
    """Compulsory HBM bytes of each kernel of one step: every input read once and every output written once."""
    collide = 56 * n_particles + 12 * n_cells   # perm 4 + gathered state 24 + state out 24 + next key 4; offsets 4 + count RMW 8
    scan = 20 * n_cells                         # counts read twice, offsets + cursors written, counts cleared
    perm = 8 * n_particles + 8 * n_cells        # key 4 + perm slot 4; cursor RMW 8
    return {"collide": collide, "scan": scan, "perm": perm, "step": collide + scan + perm}


def host_cpu() -> dict:
    """CPU model and core counts of the host the baseline runs on."""
    model = "unknown"
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        usable = os.cpu_count()
    return {"cpu_model": model, "nproc": os.cpu_count(), "usable_cpus": usable}


def _time_cpu(sim, seconds_target: float):
    sim.step(1)
    t0 = time.perf_counter()
    steps = 0
    while True:
        sim.step(1)
        steps += 1
        if time.perf_counter() - t0 >= seconds_target:
            break
    return sim.N * steps / (time.perf_counter() - t0), steps


def cpu_baseline(n: int, seconds_target: float = 15.0) -> dict:
    """Reference CPU backend (oracle/_ref, compiled from the reference sources;
    single-threaded, as the reference is) on one host core: a bounded sample of
    the workload, a cross-section slice of the same channel (100x100 of the
    1000x1000 cells, the same gap, walls, n and thermodynamics: the CPU cost per
    particle-step does not depend on the cross-section), plus the plain SRD
    Poiseuille case of the reference README (10x10x100 cells, n = 10)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle_lib import Oracle, Reference, ref_available  # checker libraries, timed here only

    kind = "reference" if ref_available() else "port"

    def make(volume, nn):
        if kind == "reference":
            return Reference(volume, (1, 1, 0), n=nn, drag=0.001, delta_t=0.02, algorithm="srd", seed=(11, 13))
        return Oracle(volume, (1, 1, 0), n=nn, drag=0.001, delta_t=0.02, algorithm="srd", rng="reference", seed=(11, 13))

    fluid = (100, 100, 20)
    sim = make((fluid[0], fluid[1], fluid[2] + 2), n)
    value, steps = _time_cpu(sim, seconds_target)
    n_slice = sim.N
    del sim
    c1 = make((10, 10, 102), 10)
    c1_value, c1_steps = _time_cpu(c1, seconds_target / 2)
    what = "reference CPU backend (cpu_backend.cpp + simulation_functions.cpp, g++ -O3)" if kind == "reference" \
        else "oracle restatement"
    return {"value": value, "unit": "particle-steps/s", "cores": 1, "kind": kind,
            "sample": f"{what}: SRD {fluid[0]}x{fluid[1]}x{fluid[2]} cells (a cross-section slice of the "
                      f"1000x1000x20 channel), n={n}, walls in z ({n_slice} particles), {steps} steps after 1 "
                      f"warm-up, single thread",
            "host": host_cpu(),
            "c1": {"value": c1_value, "unit": "particle-steps/s", "steps": c1_steps,
                   "workload": "SRD Poiseuille 10x10x100 cells, n=10, walls in z (1e5 particles), single thread"}}


def load_traffic(config: str):
    """HBM bytes per k_collide launch measured by rocprofv3 FETCH_SIZE / WRITE_SIZE
    (calibrated against a copy kernel), from the committed profile summary."""
    f = ROOT / "profiles" / f"k_collide_{config}_traffic.md"
    if not f.exists():
        return None
    out = {}
    for line in f.read_text().splitlines():
        key, _, val = line.strip().partition(":")
        if key in ("particles", "bytes_per_launch") and val.strip().isdigit():
            out[key] = int(val)
    return out if len(out) == 2 else None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--config", default="c2", choices=sorted(CONFIGS))
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--seed", type=int, default=None, help="default 0x5EED0000 + config number")
    ap.add_argument("--windows", type=int, default=32, help="adjoint accuracy study: windows (c5)")
    ap.add_argument("--no-loop", action="store_true", help="c4: skip the plain per-channel loop (profiling)")
    ap.add_argument("--tiling", default="",
                    help="cells_per_tile,lds_capacity for k_collide (results do not depend on it)")
    ap.add_argument("--sample", action="store_true",
                    help="time step_and_sample (a mean-field sample every step) instead of step")
    ap.add_argument("--single", action="store_true",
                    help="c2: time only the plain step (skip the sampled-step and extended-rule lines)")
    ap.add_argument("--copy-gb", type=float, default=0.0,
                    help="only run the plain copy kernel over this many GB (profiling reference) and exit")
    args = ap.parse_args()

    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: this library has no multi-GPU path "
              "(the step would shard as x-slabs with a particle halo exchange between neighbours)", file=sys.stderr)
        sys.exit(2)

    import cudampcd_amd as pympcd
    from cudampcd_amd import capi
    from cudampcd_amd.capi import Handle

    if args.copy_gb > 0:
        nbytes = int(args.copy_gb * 1e9) // 16 * 16
        ms = capi.bench_copy(nbytes, 5)
        print(json.dumps({"copy_bytes": nbytes, "ms_per_copy": ms, "GBps": 2 * nbytes / ms / 1e6}), flush=True)
        return

    cfg = CONFIGS[args.config]
    if args.seed is None:
        args.seed = 0x5EED0000 + int(args.config[1:])
    if args.config == "c4":
        print(json.dumps(bench_ensemble(args, cfg)), flush=True)
        return
    if args.config == "c5":
        print(json.dumps(bench_adjoint(args, cfg)), flush=True)
        return
    tiling = tuple(int(x) for x in args.tiling.split(",")) if args.tiling else None
    out = measure(args.config, args.seed, args.steps, args.warmup, args.sample, tiling)
    if args.config == "c2" and not args.sample and not args.single:
        # the sampled step (the reference's viscosity workflow) and the extended
        # rule, timed in the same run so their numbers carry the driver's clock
        out["extra"] = {
            "c2_sampled": summary(measure("c2", args.seed, args.steps, args.warmup, True, None)),
            "c3": summary(measure("c3", 0x5EED0003, args.steps, args.warmup, False, None)),
        }
        # the ensemble against its plain loop, and the adjoint gradient against
        # the loss, in the same run (their own configurations' step counts)
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup, sub.seed, sub.no_loop, sub.no_cpu = 50, 5, 0x5EED0004, False, True
        out["extra"]["c4"] = bench_ensemble(sub, CONFIGS["c4"])
        sub = argparse.Namespace(**vars(args))
        sub.steps, sub.warmup, sub.seed, sub.no_cpu = 5, 10, 0x5EED0005, True
        out["extra"]["c5"] = bench_adjoint(sub, CONFIGS["c5"])
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(CONFIGS[args.config].get("n", 10))
    print(json.dumps(out), flush=True)


def summary(r: dict) -> dict:
    """The figures of a secondary measurement kept in the main line."""
    return {k: r[k] for k in ("value", "unit", "ms_per_step", "steps", "warmup", "config", "roofline", "device_ms")}


def measure(config, seed, steps, warmup, sample, tiling) -> dict:
    """One single-system workload: build it, warm up, time `steps` steps (or
    sampled steps) with the state resident in HBM, and the collide launches
    with HIP events on the handle's stream."""
    import gc

    import cudampcd_amd as pympcd
    from cudampcd_amd import capi
    from cudampcd_amd.capi import Handle

    cfg = CONFIGS[config]
    p = make_params(pympcd, cfg["fluid"], cfg["n"], cfg["algorithm"], 0.001, seed)
    print(f"[bench] creating {cfg['workload']} ({p.N} particles)", file=sys.stderr, flush=True)
    sim = pympcd.Simulation(p, "hip")
    h = Handle(sim)
    N, C = h.N, h.C
    if tiling:
        h.set_tiling(*tiling)

    print(f"[bench] warmup {warmup} steps", file=sys.stderr, flush=True)
    sim.step(warmup)
    run = sim.step
    if sample:
        sim.step_and_sample(2)      # the first sample of an average runs eagerly
        run = sim.step_and_sample
    h.enable_timing(True)
    print(f"[bench] timing {steps} {'sampled ' if sample else ''}steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    run(steps)                      # returns after the device finished
    wall = time.perf_counter() - t0
    collide_ms, total_ms, launches = h.last_timing()
    tiles = list(h.tiling())
    volume = [int(v) for v in p.volume_size]
    del sim, h
    gc.collect()

    copy_bytes = 4 * 10 ** 9
    copy_ms = capi.bench_copy(copy_bytes, 5)
    copy_gbps = 2 * copy_bytes / copy_ms / 1e6

    ms_per_step = wall * 1e3 / steps
    value = N * steps / wall
    ab = algorithmic_bytes(N, C)
    collide_avg_ms = collide_ms / steps
    achieved = ab["collide"] / (collide_avg_ms * 1e-3) / 1e9
    traffic = load_traffic(config)
    roofline = {
        "bound": "hbm",
        "kernel": "k_collide (gather + stream + sort-in-cell + collide + next-step histogram"
                  + (" + mean-field sums)" if sample else ")"),
        "achieved": round(achieved, 1),
        "peak": HBM_PEAK_GBS,
        "unit": "GB/s",
        "frac": round(achieved / HBM_PEAK_GBS, 4),
        "traffic": traffic.get("bytes_per_launch") if traffic and traffic.get("particles") == N and not sample else None,
        "algorithmic_bytes_per_launch": ab["collide"],
        "avg_launch_ms": round(collide_avg_ms, 4),
        "copy_kernel_GBps": round(copy_gbps, 1),
        "frac_of_copy": round(achieved / copy_gbps, 4),
    }
    per_step = 120 + 20 / cfg["n"]
    step_roofline = {
        "basis": f"{per_step:.0f} B/particle-step pipeline figure (stream+histogram 24, scatter 48, collision 48, cells 20/n)",
        "achieved": round(value * per_step / 1e9, 1),
        "frac": round(value * per_step / 1e9 / HBM_PEAK_GBS, 4),
        "this_build_bytes_per_particle_step": round(ab["step"] / N, 2),
        "this_build_GBps": round(ab["step"] * steps / wall / 1e9, 1),
    }
    return {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": steps,
        "warmup": warmup,
        "ms_per_step": round(ms_per_step, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": round(value / REFERENCE_PSTEPS, 3) if config == "c2" and not sample else None,
        "dtype": "f32",
        "data": "synthetic (seeded uniform positions, Maxwell-Boltzmann velocities)",
        "config": {"workload": cfg["workload"], "cells": volume, "particles": N,
                   "algorithm": cfg["algorithm"], "n": cfg["n"], "delta_t": 0.02, "drag": 0.001,
                   "tiling": tiles, "parallelism": "single GPU",
                   "sampling": "a mean-field sample every step (step_and_sample)" if sample else "none (step)"},
        "roofline": roofline,
        "step_roofline": step_roofline,
        "device_ms": {"collide_total": round(collide_ms, 3), "step_total": round(total_ms, 3), "launches": launches},
    }


def bench_ensemble(args, cfg):
    """Batched ensemble vs a plain loop of Simulation objects over the same channels and seeds."""
    import cudampcd_amd as pympcd
    from cudampcd_amd import capi
    from cudampcd_amd.capi import Handle

    members = c4_members(args.seed, cfg["channels"])
    params = [make_params(pympcd, m["fluid"], m["n"], cfg["algorithm"], m["drag"], m["seed"], m["system_id"])
              for m in members]
    print(f"[bench] creating ensemble of {len(params)} channels", file=sys.stderr, flush=True)
    ens = pympcd.Ensemble(params)
    h = Handle(ens)
    N, C = h.N, h.C
    tc, cap, tune_ms = ens.autotune(4)   # steps per trial: two ranked the tile shapes unreliably
    shapes = {}
    for i, m in enumerate(members):       # the shape each density class runs (cells per tile vary with the grid)
        shapes.setdefault(f"n={m['n']}", sorted({tuple(h.system_tiling(j)) for j, q in enumerate(members)
                                                 if q["n"] == m["n"]}))
    print(f"[bench] autotuned tiles: {shapes}, {tune_ms:.2f} ms/step", file=sys.stderr, flush=True)
    ens.step(args.warmup)
    print(f"[bench] timing {args.steps} batched steps", file=sys.stderr, flush=True)
    t0 = time.perf_counter()
    ens.step(args.steps)
    wall = time.perf_counter() - t0
    # kernel time of the collide launches (eager launches with HIP events)
    h.enable_timing(True)
    ens.step(5)
    collide_ms, _, _ = h.last_timing()
    h.enable_timing(False)
    collide_avg_ms = collide_ms / 5
    del ens, h

    wall_loop = None
    if not args.no_loop:
        print(f"[bench] plain loop over {len(params)} Simulation objects", file=sys.stderr, flush=True)
        sims = [pympcd.Simulation(q, "hip") for q in params]
        for sm in sims:
            sm.step(args.warmup)
        t0 = time.perf_counter()
        for sm in sims:
            sm.step(args.steps)
        wall_loop = time.perf_counter() - t0
        del sims

    copy_bytes = 4 * 10 ** 9
    copy_gbps = 2 * copy_bytes / capi.bench_copy(copy_bytes, 5) / 1e6
    value = N * args.steps / wall
    loop_value = N * args.steps / wall_loop if wall_loop else None
    ab = algorithmic_bytes(N, C)
    achieved = ab["collide"] / (collide_avg_ms * 1e-3) / 1e9
    n_mean = N / sum(m["fluid"][0] * m["fluid"][1] * m["fluid"][2] for m in members)
    per_step = 120 + 20 / n_mean
    out = {
        "metric": "particle-steps/s",
        "value": value,
        "unit": "particle-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall * 1e3 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded channel set, uniform positions, Maxwell-Boltzmann velocities)",
        "config": {"workload": cfg["workload"], "channels": len(members), "particles": N, "cells": C,
                   "algorithm": cfg["algorithm"], "delta_t": 0.02,
                   "tiling": {k: [list(x) for x in v] for k, v in shapes.items()}, "parallelism": "single GPU"},
        "plain_loop": None if wall_loop is None else {
            "value": loop_value, "ms_per_step": round(wall_loop * 1e3 / args.steps, 4),
            "what": f"for each of {len(members)} Simulation objects (same seeds): step({args.steps})"},
        "batched_over_loop": None if wall_loop is None else round(value / loop_value, 4),
        "equal_accuracy": "per-channel results are bit-identical to the loop's (same seeds), so the viscosity "
                          "samples and their standard errors are identical; GPU-seconds per sample scale as 1/value",
        "roofline": {
            "bound": "hbm",
            "kernel": "k_collide over all channels",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": None,
            "algorithmic_bytes_per_launch": ab["collide"],
            "avg_launch_ms": round(collide_avg_ms, 4),
            "copy_kernel_GBps": round(copy_gbps, 1),
        },
        "step_roofline": {
            "basis": f"{per_step:.0f} B/particle-step pipeline figure (mean n = {n_mean:.2f})",
            "achieved": round(value * per_step / 1e9, 1),
            "frac": round(value * per_step / 1e9 / HBM_PEAK_GBS, 4),
        },
    }
    if not args.no_cpu:
        print("[bench] cpu baseline", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline(10)
    return out


def bench_adjoint(args, cfg):
    """Time per gradient of the profile loss against one loss evaluation, with
    the configuration's checkpoint interval (forward, recompute of every segment
    but the last two, reverse) and with the whole tape resident in HBM (forward and
    reverse only); then the accuracy study: the drag gradient of consecutive
    windows from the adjoint and from central differences of the loss (common
    random numbers within a window), their means with standard errors, and the
    cost of each estimator per unit of variance."""
    import numpy as np

    import cudampcd_amd as pympcd
    from cudampcd_amd import adjoint

    p = make_params(pympcd, cfg["fluid"], cfg["n"], cfg["algorithm"], 0.001, args.seed)
    sim = pympcd.Simulation(p, "hip")
    steps, K = cfg["steps"], cfg["checkpoint"]
    nl = adjoint.fluid_layers(p)
    zc = (np.arange(nl) + 0.5) / nl * 2.0 - 1.0
    a = float(np.random.default_rng(args.seed).uniform(0.02, 0.05))
    target = a * (1.0 - zc ** 2)
    print(f"[bench] {sim.n_particles} particles, warm-up {args.warmup * 10} steps", file=sys.stderr, flush=True)
    sim.step(args.warmup * 10)
    reps = max(1, min(args.steps, 5))

    def timed(f):
        f()   # allocates the tape
        t0 = time.perf_counter()
        for _ in range(reps):
            out = f()
        return out, (time.perf_counter() - t0) / reps

    r, t_grad = timed(lambda: adjoint.profile_loss_grad(sim, steps, K, p.drag, p.rotation_angle, target))
    r_res, t_res = timed(lambda: adjoint.profile_loss_grad(sim, steps, steps, p.drag, p.rotation_angle, target))
    lr, t_loss = timed(lambda: adjoint.profile_loss(sim, steps, p.drag, p.rotation_angle, target))
    assert r_res.grad_drag == r.grad_drag and np.array_equal(r_res.grad_force, r.grad_force)

    def fd_drag(nsteps, frac):
        h = frac * p.drag
        return (adjoint.profile_loss(sim, nsteps, p.drag + h, p.rotation_angle, target).loss -
                adjoint.profile_loss(sim, nsteps, p.drag - h, p.rotation_angle, target).loss) / (2 * h)

    short = cfg["checkpoint"]
    r_short = adjoint.profile_loss_grad(sim, short, K, p.drag, p.rotation_angle, target, want_force_grad=False)
    fd_short = {f"h={f:g}g": fd_drag(short, f) for f in (0.01, 0.05, 0.2)}
    # accuracy study over consecutive windows (each starts where the previous ended)
    windows, fracs = args.windows, (0.05, 0.2)
    print(f"[bench] accuracy study: {windows} windows of {steps} steps", file=sys.stderr, flush=True)
    adj_g, fd_g = [], {f: [] for f in fracs}
    for _ in range(windows):
        adj_g.append(adjoint.profile_loss_grad(sim, steps, steps, p.drag, p.rotation_angle, target,
                                               want_force_grad=False).grad_drag)
        for f in fracs:
            fd_g[f].append(fd_drag(steps, f))
        sim.step(steps)

    def stats(x):
        x = np.asarray(x, np.float64)
        return float(x.mean()), float(x.std(ddof=1)), float(x.std(ddof=1) / np.sqrt(len(x)))

    am, asd, ase = stats(adj_g)
    study = {"windows": windows, "adjoint": {"mean": am, "std": asd, "se": ase}}
    for f in fracs:
        m, sd, se = stats(fd_g[f])
        # GPU-ms per estimate of d/d(drag) with standard error eps: (sd / eps)^2 samples
        study[f"central_fd_h={f:g}g"] = {
            "mean": m, "std": sd, "se": se, "z_vs_adjoint": (m - am) / float(np.hypot(se, ase)),
            "adjoint_speedup_at_equal_se_one_scalar": round((sd ** 2 * 2 * t_loss) / (asd ** 2 * t_res), 2),
        }
    C = int(np.prod([int(v) for v in p.volume_size]))
    N = sim.n_particles
    out = {
        "metric": "ms per gradient",
        "value": round(t_grad * 1e3, 3),
        "unit": "ms",
        "n_gpus": 1,
        "steps": reps,
        "warmup": args.warmup * 10,
        "ms_per_step": round(t_grad * 1e3 / steps, 4),
        "higher_is_better": False,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 state, f64 gradient accumulation",
        "data": "synthetic (seeded initial state, target a (1 - (2z/L)^2) with seeded a)",
        "config": {"workload": cfg["workload"], "particles": N, "cells": C, "steps": steps,
                   "checkpoint_every": K, "parallelism": "single GPU"},
        "loss": lr.loss,
        "ms_per_loss": round(t_loss * 1e3, 3),
        "loss_evaluations_per_gradient": round(t_grad / t_loss, 3),
        "resident_tape": {"ms_per_gradient": round(t_res * 1e3, 3),
                          "loss_evaluations_per_gradient": round(t_res / t_loss, 3),
                          "what": f"checkpoint interval {steps}: the forward pass tapes all {steps} steps "
                                  f"(bit-identical gradient), nothing is recomputed"},
        "gradient": {"d_drag": r.grad_drag, "d_angle_per_degree": r.grad_angle,
                     "d_force_norm": float(np.linalg.norm(r.grad_force))},
        "central_fd_cost": {
            "ms_two_scalars": round(4 * t_loss * 1e3, 3),
            "ms_full_field": round(2 * (2 + C) * t_loss * 1e3, 1),
            "adjoint_speedup_two_scalars": round(4 * t_loss / t_grad, 3),
            "adjoint_speedup_full_field": round(2 * (2 + C) * t_loss / t_grad, 1),
        },
        "central_fd_d_drag_short_window": {f"{short}_steps": {"adjoint": r_short.grad_drag, **fd_short}},
        "accuracy_study_d_drag": study,
        "particle_steps_per_s_forward": N * steps / t_loss,
    }
    del sim
    return out


if __name__ == "__main__":
    main()
