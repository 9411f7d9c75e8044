"""Command-line driver on the GPU: the written averages equal those of the
same run driven through the Python API, and a run stopped by SIGUSR2 and
restarted from its backup writes the same files bit for bit."""
import os
import signal
from pathlib import Path

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

INPUT = """
        -------- general settings ------

        directory  run/

        experiment              standard
        volume_size            6   5   8
        periodicity            1   1   0

        ------------ fluid -------------

        algorithm               {algo}
        collision_kernel       optimized
        density                        5
        temperature                  1.0
        delta_t                     0.02
        drag                        0.01

        ---------- simulation ----------

        equilibration_steps            3
        steps                         10
        sample_every                   1
        average_samples_over           4
        write_backup                 {wb}
        read_backup                  {rb}
        cuda_device                    0
"""


def _files(root):
    from cudampcd_amd import io
    return {t: io.read_fluid_snapshot(root, t) for t in io.snapshot_steps(root)}


@pytest.mark.parametrize("algo", ["srd", "extended"])
def test_cli_averages_and_restart(pympcd, tmp_path, monkeypatch, algo):
    from cudampcd_amd import capi, cli
    # uninterrupted run
    monkeypatch.chdir(tmp_path)
    r = cli.parse_input(INPUT.format(algo=algo, wb="no", rb="no"))
    assert cli.run(r, seed=77, out=lambda *_: None) == 0
    full = _files(tmp_path / "run")
    assert sorted(full) == [4, 8, 12, 13]

    # the same windows through the API: t = 4 alone, 5..8, 9..12, and the
    # window 13 opens, each sum over the averaged samples / 4 (13: / 1)
    sim = pympcd.Simulation(cli.make_params(pympcd, r, 77), "hip")
    h = capi.Handle(sim)
    sim.step(3)
    X, Y, Z = (int(v) for v in sim.get_parameters().volume_size)
    for t, k, div in ((4, 1, 4), (8, 4, 4), (12, 4, 4), (13, 1, 1)):
        h.reset_sampling()
        sim.step_and_sample(k)
        acc = h.sampling_state()[0] / div
        a = acc.reshape(Z, Y, X, 4).transpose(2, 1, 0, 3).astype(np.float32)
        d, v = full[t]
        assert np.array_equal(d, a[..., 0]) and np.array_equal(v, a[..., 1:]), t
    assert np.isclose(full[8][0].mean(dtype=np.float64), h.N / h.C, rtol=1e-6)   # every particle counted once

    # stopped by SIGUSR2 after the sample at t = 8, restarted from its backup
    monkeypatch.chdir(tmp_path)
    other = tmp_path / "second"
    other.mkdir()
    monkeypatch.chdir(other)

    def out(msg, *_):
        if "fluid/8.npy" in str(msg):
            os.kill(os.getpid(), signal.SIGUSR2)

    r2 = cli.parse_input(INPUT.format(algo=algo, wb="yes", rb="no"))
    assert cli.run(r2, seed=77, out=out) == 0
    assert (other / "run" / "last_timestep").read_text() == "4"
    assert sorted(_files(other / "run")) == [4, 8]
    monkeypatch.chdir(other)
    r3 = cli.parse_input(INPUT.format(algo=algo, wb="yes", rb="yes"))
    assert cli.run(r3, seed=123, out=lambda *_: None) == 0    # the backup's seed wins
    resumed = _files(other / "run")
    assert sorted(resumed) == sorted(full)
    for t in full:
        assert np.array_equal(resumed[t][0], full[t][0]) and np.array_equal(resumed[t][1], full[t][1]), t
