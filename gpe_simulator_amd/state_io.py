"""Wavefunction snapshots in the reference's on-disk format (gpe/state_io.py:11-58).

The format, which files from either engine must share:

* ``psi_<step>.npy`` -- the field as a complex128 (N, N, N) cube, C order;
* ``psi_<step>.json`` -- step, time, evolution kind, reference N_atom,
  simulation range, N_grids and the stage's configuration;
* ``grid.npz`` in the same folder (older exports: ``psi_<step>_grid.npz``
  beside the snapshot) -- the x and k axes.

The field crosses between device and host in slabs of SLAB_SIZE x-planes, so
host memory holds one slab at a time; a complex64 field is widened to
complex128 on the device as it is read.  Loading checks the cube against the
stage (shape and dtype), and the companions when present (range, N_atom,
axes); the error texts are the reference's, because callers match on them.
"""

from __future__ import annotations

import json
from pathlib import Path

import numpy as np

from .field import DeviceField

SLAB_SIZE = 16

_CUBE_ERROR = "The loaded state must be a complex128 cube matching N_grids."
_RANGE_ERROR = "The loaded state has a different simulation range."
_ATOMS_ERROR = "The loaded state uses a different reference N_atom."
_GRID_ERROR = "The loaded state's grid does not match this simulation."


def _slabs(n):
    """(first, end) x-plane ranges of at most SLAB_SIZE planes covering 0..n."""
    for first in range(0, n, SLAB_SIZE):
        yield first, min(n, first + SLAB_SIZE)


def _snapshot_metadata(stage, step):
    return {"step": step, "time": step * stage.dt, "evolution": stage.type, "N_atom": stage.N_atoms,
            "simulation_range": stage.simulation_range, "N_grids": stage.N_grids, "stage": stage.config}


def _host_planes(psi, first, end):
    """x-planes first..end of the stage field as complex128 host memory."""
    if isinstance(psi, DeviceField):
        return psi.slab(first, end)
    return np.asarray(psi[first:end], dtype=np.complex128)


def save_wavefunction(path, stage, step):
    """Write the stage's field at `step`, its metadata and the shared axes."""
    target = Path(path)
    target.parent.mkdir(parents=True, exist_ok=True)
    n = stage.N_grids
    cube = np.lib.format.open_memmap(target, mode="w+", dtype=np.complex128, shape=(n, n, n))
    for first, end in _slabs(n):
        cube[first:end] = _host_planes(stage.psi, first, end)
    cube.flush()
    del cube                                       # closes the mapping before the companions
    np.savez(target.parent / "grid.npz", x_array=np.asarray(stage.x_array),
             k_array=np.asarray(stage.k_array))
    text = json.dumps(_snapshot_metadata(stage, step), indent=2)
    target.with_suffix(".json").write_text(text + "\n")


def _axes_file(snapshot):
    """The axes written with `snapshot`: its own _grid.npz, else the folder's."""
    own = snapshot.with_name(f"{snapshot.stem}_grid.npz")
    return own if own.exists() else snapshot.parent / "grid.npz"


def _check_companions(snapshot, stage):
    meta_file = snapshot.with_suffix(".json")
    if meta_file.exists():
        meta = json.loads(meta_file.read_text())
        if meta["simulation_range"] != stage.simulation_range:
            raise ValueError(_RANGE_ERROR)
        if meta["N_atom"] != stage.N_atoms:
            raise ValueError(_ATOMS_ERROR)
    axes_file = _axes_file(snapshot)
    if not axes_file.exists():
        return
    with np.load(axes_file, allow_pickle=False) as axes:
        for key in ("x_array", "k_array"):
            mine = np.asarray(getattr(stage, key))
            theirs = axes[key]
            if theirs.shape != mine.shape or not np.allclose(theirs, mine):
                raise ValueError(_GRID_ERROR)


def load_wavefunction(path, stage):
    """The saved field for `stage`, uploaded slab by slab into the stage's solve
    (a host array when the stage has no solve yet).  The norm is kept as saved
    and no clock is carried over."""
    snapshot = Path(path)
    cube = np.load(snapshot, mmap_mode="r", allow_pickle=False)
    n = stage.N_grids
    if cube.dtype != np.complex128 or cube.shape != (n, n, n):
        raise ValueError(_CUBE_ERROR)
    _check_companions(snapshot, stage)
    solve = stage.solve
    if solve is None:
        return np.array(cube)
    plane = n * n
    for first, end in _slabs(n):
        solve.upload(np.ascontiguousarray(cube[first:end]), offset=first * plane)
    return DeviceField(solve)
