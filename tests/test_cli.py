"""Command-line driver: parameter file, schedule of samples and averages
(main.cpp:11-71, probing.hpp:14-31, simulation_parameters.cpp:15-183).

The schedule runs against a stand-in simulation whose per-step sample is the
step count t itself, so every written file holds the mean of the t's its
window collected; the expected windows are worked out by hand from the
reference's rules in each test's comment."""
from pathlib import Path

import numpy as np
import pytest

from cudampcd_amd import cli, io

GOLDEN = Path(__file__).parent / "golden"


def test_parse_reference_format():
    r = cli.parse_input((GOLDEN / "cli_input.txt").read_text())
    assert r.output_directory == "out/channel_a/"
    assert r.volume_size == (6.0, 5.0, 8.0) and r.periodicity == (1, 1, 0)
    assert (r.algorithm, r.collision_kernel) == ("extended", "sorting")
    assert (r.n, r.temperature, r.delta_t, r.drag) == (5, 0.5, 0.05, 0.002)
    assert (r.equilibration_steps, r.steps, r.sample_every, r.average_samples) == (3, 12, 1, 4)
    assert r.write_backup is True
    # a directory in place of yes/no means: restore from there
    assert r.read_backup is True and r.load_directory == "runs/old/"
    assert r.device_id == 0


def test_parse_rejects_truncated_file():
    text = (GOLDEN / "cli_input.txt").read_text()
    with pytest.raises(ValueError):
        cli.parse_input(text[: text.index("steps")])


def test_arguments_and_missing_file(capsys, tmp_path):
    assert cli.main([]) == 255
    assert "input file" in capsys.readouterr().out
    assert cli.main([str(tmp_path / "nope.txt")]) == 0
    assert "no such input_file" in capsys.readouterr().out


def test_backup_cadence():
    r = cli.parse_input((GOLDEN / "cli_input.txt").read_text())
    assert cli.backup_every(r) == 500_000
    r.volume_size = (1000.0, 1000.0, 18.0)          # 1000 * 1000 * 20 * 5 = 1e8 particles: not above 1e8
    assert cli.backup_every(r) == 500_000
    r.n = 6
    assert cli.backup_every(r) == 100_000


class _Sim:
    def __init__(self, shape):
        self.shape = shape
        self.C = int(np.prod(shape))
        self.t = 0
        self.acc = np.zeros((self.C, 4))
        self.count = 0
        self.active = False
        self.calls = []

    def step(self, k):
        self.calls.append(("step", k))
        self.t += k

    def step_and_sample(self, k):
        self.calls.append(("sample", k))
        for _ in range(k):
            self.t += 1
            if not self.active:
                self.acc[:] = self.t
                self.count, self.active = 0, True
            else:
                self.acc += self.t
                self.count += 1

    def get_parameters(self):
        class P:
            volume_size = tuple(float(v) for v in self.shape)
        return P()


class _Handle:
    def __init__(self, sim):
        self.sim = sim

    @property
    def step_index(self):
        return self.sim.t

    def reset_sampling(self):
        self.sim.active = False

    def sampling_state(self):
        return self.sim.acc.copy(), self.sim.count, self.sim.active


def _run(tmp_path, monkeypatch, eq, steps, every, avg):
    monkeypatch.chdir(tmp_path)
    r = cli.parse_input((GOLDEN / "cli_input.txt").read_text())
    r.output_directory, r.read_backup, r.write_backup = "run", False, False
    r.equilibration_steps, r.steps, r.sample_every, r.average_samples = eq, steps, every, avg
    sims = []

    def make(_):
        sims.append(_Sim((2, 3, 4)))
        return sims[0]

    assert cli.run(r, make_sim=make, make_handle=_Handle, out=lambda *_: None) == 0
    root = tmp_path / "run"
    got = {}
    for t in io.snapshot_steps(root):
        d, v = io.read_fluid_snapshot(root, t)
        assert d.shape == (2, 3, 4) and v.shape == (2, 3, 4, 3)
        assert np.all(d == d.flat[0]) and np.all(v == d[..., None])
        got[t] = float(d.flat[0])
    assert sims[0].t == eq + steps
    assert (root / "simulation.info").read_text().startswith("simulation parameters where: \n")
    assert (root / "last_timestep").read_text() == str(steps)
    return got, sims[0].calls


def test_schedule_windows_every_step(tmp_path, monkeypatch):
    # eq 3, 10 steps, every step sampled, windows of 4: t = 4..13 and
    # t % 4 = 0 ends a window, 1 starts one.  t = 4 ends a window holding
    # only itself (4 / 4); 5..8 -> 26 / 4; 9..12 -> 42 / 4; 13 opens a window
    # the run leaves open, written at the end over its one sample.
    got, calls = _run(tmp_path, monkeypatch, 3, 10, 1, 4)
    assert got == {4: 1.0, 8: 6.5, 12: 10.5, 13: 13.0}
    assert ("sample", 2) in calls            # 6 and 7 only add: one call


def test_schedule_sparse_samples(tmp_path, monkeypatch):
    # eq 0, 9 steps, a sample every second loop step (t = 1, 3, 5, 7, 9),
    # windows by (t // 2) % 3: t = 1 ends one (1 / (3 // 2)); 3 starts,
    # 5 adds, 7 ends (3 + 5 + 7) / 1; 9 starts the window left open.
    got, calls = _run(tmp_path, monkeypatch, 0, 9, 2, 3)
    assert got == {1: 1.0, 7: 15.0, 9: 9.0}
    assert calls.count(("step", 1)) == 4


def test_schedule_snapshots(tmp_path, monkeypatch):
    # average_samples_over 1: every sample is written alone
    got, _ = _run(tmp_path, monkeypatch, 2, 3, 1, 1)
    assert got == {3: 3.0, 4: 4.0, 5: 5.0}


def test_thermal_velocity_follows_the_file_reader():
    """The reference's file reader sets thermal_velocity = sqrtf(float T)
    (simulation_parameters.cpp:182); the Params temperature setter takes the
    square root of the double, which differs in the last bit for T = 0.9."""
    import cudampcd_amd as pympcd

    text = (GOLDEN / "cli_input.txt").read_text().replace("temperature                  0.5",
                                                         "temperature                  0.9")
    r = cli.parse_input(text)
    p = cli.make_params(pympcd, r)
    want = np.sqrt(np.float32(0.9))
    assert np.float32(p.thermal_velocity) == want
    assert np.float32(np.sqrt(0.9)) != want            # the case the reader's order matters for
    q = pympcd.Params()
    q.temperature = 0.9
    assert np.float32(q.thermal_velocity) == np.float32(np.sqrt(0.9))


def test_checkpoint_params_round_trip_bitwise():
    """A checkpoint restores the thermal velocity it was saved with (at T = 0.9
    the float32 temperature's square root differs from the saved value)."""
    import cudampcd_amd as pympcd
    from cudampcd_amd import checkpoint

    p = pympcd.Params()
    p.n = 7
    p.temperature = 0.9
    p.periodicity = (1, 1, 0)
    p.volume_size = (6, 5, 8)
    p.drag = 0.003
    p.delta_t = 0.05
    p.rotation_angle = 97.0
    p.algorithm = "extended"
    p.seed = 4321
    rec = checkpoint.param_record(p, 4321)
    q = checkpoint.params_of(rec)
    for k in ("n", "temperature", "thermal_velocity", "drag", "delta_t", "rotation_angle", "N",
              "volume_size", "periodicity", "algorithm", "experiment", "seed"):
        assert getattr(q, k) == getattr(p, k), k
    legacy = {k: v for k, v in rec.items() if k != "thermal_velocity"}
    assert checkpoint.params_of(legacy).thermal_velocity != p.thermal_velocity
