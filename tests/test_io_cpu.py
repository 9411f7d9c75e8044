"""Fluid snapshot layout (CPU): the reference writes fluid/<step> as float32
(Z, Y, X, 4), records {density, vx, vy, vz} in cell order x + X (y + Y z)
(cuda_backend_impl.cu:63-71; FluidState, mpc_cell.hpp:105-108)."""
from __future__ import annotations

import numpy as np

from cudampcd_amd import io


def test_snapshot_layout_and_round_trip(tmp_path):
    X, Y, Z = 5, 3, 4
    rng = np.random.default_rng(0)
    d = rng.uniform(0, 20, size=(X, Y, Z)).astype(np.float32)
    v = rng.normal(size=(X, Y, Z, 3)).astype(np.float32)
    f = io.write_fluid_snapshot(tmp_path, 1200, d, v)
    assert f.name == "1200.npy" and f.parent.name == "fluid"
    rec = np.load(f)
    assert rec.dtype == np.float32 and rec.shape == (Z, Y, X, 4)
    flat = rec.reshape(-1, 4)
    for x, y, z in [(0, 0, 0), (4, 2, 3), (1, 2, 0), (3, 0, 2)]:
        c = x + X * (y + Y * z)
        assert flat[c, 0] == d[x, y, z]
        assert np.array_equal(flat[c, 1:], v[x, y, z])
    d2, v2 = io.read_fluid_snapshot(tmp_path, 1200)
    assert np.array_equal(d2, d) and np.array_equal(v2, v)
    io.write_fluid_snapshot(tmp_path, 100, d, v)
    assert io.snapshot_steps(tmp_path) == [100, 1200]
