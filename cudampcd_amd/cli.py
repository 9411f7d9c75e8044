"""Command-line driver: ``python -m cudampcd_amd <input_file>``.

Runs a simulation described by a parameter file in the reference's format
(simulation_parameters.cpp:15-183) the way the reference's program does
(main.cpp:11-71): equilibration, then the sampling run, with fluid samples,
backups and the SIGUSR2 time-out of status_io.hpp:55-63.

Schedule (probing.hpp:14-31), with t the number of steps taken so far:
  * a sample is taken after loop step ``s`` when ``s % sample_every == 0``;
  * it is written alone when ``average_samples == 1``; otherwise
    ``(t // sample_every) % average_samples`` == 1 starts an average, 0 ends
    it (the sum is divided by ``average_samples // sample_every``, as
    cuda_backend_impl.cu:58 does), anything else adds to it.

Differences, on purpose:
  * samples are written as ``fluid/<t>.npy`` (the (Z, Y, X, 4) float32 record
    array the reference writes as the HDF5 dataset ``fluid/<t>``; HDF5 is not
    available here), read back with ``cudampcd_amd.io.read_fluid_snapshot``;
  * averages are summed in float64 from exact fixed-point cell sums (the
    reference sums float32 with Kahan compensation): equal to float rounding;
  * backups are real (``backup.npz``, cudampcd_amd.checkpoint; the reference's
    writeBackupFile is empty), and ``read_backup yes`` restores one;
  * the reference's last ``writeSample`` after the loop re-adds the final state
    to an already finished average; here it writes the open average, if any,
    divided by the samples it holds;
  * ``--seed S`` fixes the random stream (the reference seeds from the clock).
"""
from __future__ import annotations

import os
import signal
import sys
from dataclasses import dataclass
from pathlib import Path

import numpy as np

BACKUP = "backup.npz"
TIME_FILE = "last_timestep"


class _Stream:
    """Whitespace-token and line reads over a text, as ``operator>>`` and
    ``std::getline`` read an ifstream."""

    def __init__(self, text: str):
        self.t = text
        self.i = 0

    def token(self) -> str:
        t, i = self.t, self.i
        while i < len(t) and t[i].isspace():
            i += 1
        j = i
        while j < len(t) and not t[j].isspace():
            j += 1
        if i == j:
            raise ValueError("parameter file ends early")
        self.i = j
        return t[i:j]

    def getline(self) -> str:
        j = self.t.find("\n", self.i)
        j = len(self.t) if j < 0 else j
        line = self.t[self.i:j]
        self.i = min(j + 1, len(self.t))
        return line

    def pair(self) -> str:
        self.token()              # the parameter's name
        return self.token()


@dataclass
class RunSettings:
    output_directory: str
    algorithm: str
    collision_kernel: str
    volume_size: tuple          # fluid cells, as in the file
    periodicity: tuple
    n: int
    temperature: float
    delta_t: float
    drag: float
    equilibration_steps: int
    steps: int
    sample_every: int
    average_samples: int
    write_backup: bool
    read_backup: bool
    load_directory: str
    device_id: int
    text: str


def parse_input(text: str) -> RunSettings:
    """Fields of a parameter file, read in the reference's order
    (simulation_parameters.cpp:30-93): a token stream after the word
    ``general``, three skipped lines before the fluid and simulation blocks."""
    s = _Stream(text)
    while s.token() != "general":
        pass
    s.getline()
    directory = s.pair()
    s.pair()                                   # experiment: only "standard" is read (simulation_parameters.cpp:38-39)
    s.token()
    vol = tuple(float(s.token()) for _ in range(3))
    s.token()
    perio = tuple(int(s.token()) for _ in range(3))
    for _ in range(3):
        s.getline()
    algorithm = s.pair()
    kernel = s.pair()
    n = int(s.pair())
    temperature = float(s.pair())
    delta_t = float(s.pair())
    drag = float(s.pair())
    for _ in range(3):
        s.getline()
    eq = int(s.pair())
    steps = int(s.pair())
    every = int(s.pair())
    avg = int(s.pair())
    write_backup = s.pair() == "yes"
    rb = s.pair()
    read_backup, load_dir = rb == "yes", "./"
    if rb not in ("yes", "no"):                # a directory to load the status from (simulation_parameters.cpp:80-85)
        read_backup, load_dir = True, rb
    device = int(s.pair())
    if algorithm not in ("srd", "extended"):
        algorithm = "srd"                      # the reference keeps its default (simulation_parameters.cpp:50-53)
    if every < 1 or avg < 1:
        raise ValueError("sample_every and average_samples_over must be positive")
    return RunSettings(directory, algorithm, kernel, vol, perio, n, temperature, delta_t, drag, eq, steps, every,
                       avg, write_backup, read_backup, load_dir, device, text)


def make_params(pympcd, r: RunSettings, seed: int | None = None):
    p = pympcd.Params()
    p.n = r.n
    p.temperature = r.temperature
    # the parameter file reader takes the square root of the float temperature
    # (simulation_parameters.cpp:182), not of the double the Params setter sees
    p.thermal_velocity = float(np.sqrt(np.float32(r.temperature)))
    p.periodicity = r.periodicity
    p.volume_size = r.volume_size              # fluid cells; the setter adds the wall layers
    p.drag = r.drag
    p.delta_t = r.delta_t
    p.algorithm = r.algorithm
    p.experiment = "standard"
    p.device_id = r.device_id
    if seed is not None:
        p.seed = seed
    return p


def what_to_do(t: int, r: RunSettings) -> str:
    """probing.hpp:18-31."""
    if r.average_samples == 1:
        return "snapshot"
    at = (t // r.sample_every) % r.average_samples
    return "start" if at == 1 else ("finish" if at == 0 else "accumulate")


def backup_every(r: RunSettings) -> int:
    """main.cpp:48: every 1e5 steps above 1e8 particles (stored volume), else 5e5."""
    stored = [v + (0 if q else 2) for v, q in zip(r.volume_size, r.periodicity)]
    return 100_000 if stored[0] * stored[1] * stored[2] * r.n > 10e7 else 500_000


class _Flag:
    time_out = False


def _on_usr2(signum, frame):
    _Flag.time_out = True


def _enter_directory(directory: str):
    """Create the output directory (and its parents) and change into it
    (simulation_parameters.cpp:102-158)."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    os.chdir(d)


def _write_sample(sim, h, t: int, divisor: int, out):
    from . import io
    acc, _, _ = h.sampling_state()
    stored = [int(v) for v in sim.get_parameters().volume_size]
    X, Y, Z = stored
    a = (acc / float(divisor)).reshape(Z, Y, X, 4).transpose(2, 1, 0, 3)
    io.write_fluid_snapshot(".", t, a[..., 0], a[..., 1:])
    h.reset_sampling()
    out(f"\nsample fluid/{t}.npy written")


def _next_backup(step: int, every: int) -> int:
    b = -(-step // every) * every
    return b if b > 0 else every


def run(r: RunSettings, backend: str = "hip", seed: int | None = None, out=print, make_sim=None, make_handle=None) -> int:
    """Run the parameter file's simulation inside its output directory
    (``make_sim(params)`` / ``make_handle(sim)`` replace the Simulation and its
    C-ABI handle, for tests of the schedule)."""
    previous = signal.signal(signal.SIGUSR2, _on_usr2)
    _Flag.time_out = False
    try:
        return _run(r, backend, seed, out, make_sim, make_handle)
    finally:
        signal.signal(signal.SIGUSR2, previous)


def _run(r, backend, seed, out, make_sim, make_handle) -> int:
    from . import checkpoint

    _enter_directory(r.output_directory)
    load_dir = Path(r.load_directory).resolve()      # relative to the output directory, as the reference reads it
    Path("simulation.info").write_text("simulation parameters where: \n" + r.text)
    out(f"using {r.algorithm} MPCD algorithm")
    out(f"using {r.collision_kernel} collision kernel")

    start = 0
    if r.read_backup:
        sim = checkpoint.restore(load_dir / BACKUP, backend)
        if r.load_directory == "./":
            tf = load_dir / TIME_FILE
            last = int(tf.read_text().split()[0]) if tf.exists() else -1
            if last < 0:
                out("no time file found! exiting...")
                return 255
            start = last + 1
    elif make_sim is not None:
        sim = make_sim(r)
    else:
        import cudampcd_amd as pympcd
        sim = pympcd.Simulation(make_params(pympcd, r, seed), backend)
    if make_handle is not None:
        h = make_handle(sim)
    else:
        from . import capi
        h = capi.Handle(sim)

    if not r.read_backup:
        out("equilibration")
        done = 0
        while done < r.equilibration_steps:
            k = min(1000, r.equilibration_steps - done)
            sim.step(k)
            done += k
        out("equilibration done.")

    out("sampling")
    every = backup_every(r)
    window_open = h.sampling_state()[2]
    step = start
    while step < r.steps:
        if step % r.sample_every:
            # plain steps up to the next sample, backup step or time-out check
            nxt = min(r.steps, (step // r.sample_every + 1) * r.sample_every, step + 1000)
            if r.write_backup:
                nxt = min(nxt, _next_backup(step, every) + 1)
            sim.step(nxt - step)
            step = nxt - 1
        else:
            t = h.step_index + 1
            probe = what_to_do(t, r)
            if probe in ("snapshot", "start"):
                h.reset_sampling()
            k = 1
            if probe == "accumulate" and r.sample_every == 1:
                # a run of samples that only add to the average, in one call
                last = min(r.steps, step + 1000)
                if r.write_backup:
                    last = min(last, _next_backup(step, every) + 1)
                while step + k < last and what_to_do(t + k, r) == "accumulate":
                    k += 1
            sim.step_and_sample(k)
            step += k - 1
            window_open = True
            if probe == "snapshot":
                _write_sample(sim, h, t, 1, out)
                window_open = False
            elif probe == "finish":
                _write_sample(sim, h, t, max(1, r.average_samples // r.sample_every), out)
                window_open = False
        if _Flag.time_out or (r.write_backup and step % every == 0 and step != 0):
            Path(TIME_FILE).write_text(str(step))
            checkpoint.save(sim, BACKUP)
            if _Flag.time_out:
                out("time out! backup written, bye!")
                return 0
        step += 1
    out("sampling done.")
    Path(TIME_FILE).write_text(str(r.steps))
    if r.steps and start < r.steps and window_open:
        acc, count, active = h.sampling_state()
        if active:
            _write_sample(sim, h, h.step_index, count + 1, out)
    if r.write_backup:
        checkpoint.save(sim, BACKUP)
    return 0


def main(argv=None) -> int:
    argv = sys.argv[1:] if argv is None else argv
    seed = None
    if len(argv) == 3 and argv[1] == "--seed":
        seed = int(argv[2])
        argv = argv[:1]
    if len(argv) != 1:
        print("Please start the program with the name of an input file as parameter. The program is now exiting.")
        return 255
    path = Path(argv[0])
    if not path.exists():
        print(f"no such input_file: {path}")
        return 0
    return run(parse_input(path.read_text()), seed=seed)
