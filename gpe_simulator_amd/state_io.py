"""Slab-wise wavefunction snapshots (gpe/state_io.py:11-58 of the reference).

Same file format: a complex128 .npy cube written 16 x-planes at a time, a JSON
metadata file beside it and the shared grid.npz; complex64 fields are widened
to complex128 on the device before the copy.  Loading validates shape, dtype,
simulation range, reference N_atom and axes like the reference.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from .field import DeviceField

SLAB_SIZE = 16


def save_wavefunction(path, stage, step):
    path = Path(path)
    path.parent.mkdir(parents=True, exist_ok=True)
    n = stage.N_grids
    saved = np.lib.format.open_memmap(path, mode="w+", dtype=np.complex128, shape=(n, n, n))
    psi = stage.psi
    for start in range(0, n, SLAB_SIZE):
        stop = min(start + SLAB_SIZE, n)
        saved[start:stop] = psi.slab(start, stop) if isinstance(psi, DeviceField) else np.asarray(psi[start:stop])
    saved.flush()
    del saved
    np.savez(path.parent / "grid.npz", x_array=np.asarray(stage.x_array), k_array=np.asarray(stage.k_array))
    metadata = dict(step=step, time=step * stage.dt, evolution=stage.type, N_atom=stage.N_atoms,
                    simulation_range=stage.simulation_range, N_grids=stage.N_grids, stage=stage.config)
    path.with_suffix(".json").write_text(json.dumps(metadata, indent=2) + "\n")


def load_wavefunction(path, stage):
    """Load a reusable state into the stage's solve without renormalising it."""
    path = Path(path)
    saved = np.load(path, mmap_mode="r", allow_pickle=False)
    if saved.shape != (stage.N_grids,) * 3 or saved.dtype != np.complex128:
        raise ValueError("The loaded state must be a complex128 cube matching N_grids.")
    metadata_path = path.with_suffix(".json")
    if metadata_path.exists():
        metadata = json.loads(metadata_path.read_text())
        if metadata["simulation_range"] != stage.simulation_range:
            raise ValueError("The loaded state has a different simulation range.")
        if metadata["N_atom"] != stage.N_atoms:
            raise ValueError("The loaded state uses a different reference N_atom.")
    grid_path = path.with_name(path.stem + "_grid.npz")
    if not grid_path.exists():
        grid_path = path.parent / "grid.npz"
    if grid_path.exists():
        with np.load(grid_path, allow_pickle=False) as grid:
            for name in ("x_array", "k_array"):
                expected = np.asarray(getattr(stage, name))
                if grid[name].shape != expected.shape or not np.allclose(grid[name], expected):
                    raise ValueError("The loaded state's grid does not match this simulation.")
    solve = stage.solve
    if solve is None:
        return np.array(saved)
    n = stage.N_grids
    for start in range(0, n, SLAB_SIZE):
        stop = min(start + SLAB_SIZE, n)
        solve.upload(np.ascontiguousarray(saved[start:stop]), offset=start * n * n)
    return DeviceField(solve)
