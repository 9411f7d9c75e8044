"""Fluid snapshots in the reference's layout (HDF5 is not available here).

The reference writes each sample as the float32 dataset ``fluid/<step>`` of
``simulation_data.h5`` with dimensions (Z, Y, X, 4): per cell (reference
index x + X (y + Y z)) the record {density, vx, vy, vz}
(cuda_backend_impl.cu:27-75, adapters/hdf5/h5cpp.cpp:29-37, FluidState in
mpc_cell.hpp:105-108).  This module writes the same array, same shape and
order, as ``<root>/fluid/<step>.npy`` and reads it back.
"""
from __future__ import annotations

from pathlib import Path

import numpy as np


def fluid_record(density, velocity) -> np.ndarray:
    """(X, Y, Z) density and (X, Y, Z, 3) velocity -> (Z, Y, X, 4) float32 records."""
    d = np.asarray(density, dtype=np.float32)
    v = np.asarray(velocity, dtype=np.float32)
    if d.ndim != 3 or v.shape != d.shape + (3,):
        raise ValueError("density must be (X, Y, Z) and velocity (X, Y, Z, 3)")
    rec = np.empty(d.shape[::-1] + (4,), np.float32)
    rec[..., 0] = d.transpose(2, 1, 0)
    rec[..., 1:] = v.transpose(2, 1, 0, 3)
    return rec


def write_fluid_snapshot(root, step: int, density, velocity) -> Path:
    """Write one sample; returns the file written (``<root>/fluid/<step>.npy``)."""
    out = Path(root) / "fluid"
    out.mkdir(parents=True, exist_ok=True)
    f = out / f"{int(step)}.npy"
    np.save(f, fluid_record(density, velocity), allow_pickle=False)
    return f


def read_fluid_snapshot(root, step: int):
    """(density (X, Y, Z), velocity (X, Y, Z, 3)) of a written sample."""
    rec = np.load(Path(root) / "fluid" / f"{int(step)}.npy", allow_pickle=False)
    return rec[..., 0].transpose(2, 1, 0).copy(), rec[..., 1:].transpose(2, 1, 0, 3).copy()


def snapshot_steps(root) -> list[int]:
    """Steps with a written sample, ascending."""
    d = Path(root) / "fluid"
    return sorted(int(p.stem) for p in d.glob("*.npy")) if d.exists() else []
