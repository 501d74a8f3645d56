"""DeviceField: the wavefunction resident in device memory.

Plays the role of the reference's CuPy array `stage.psi`: propagators return
it, diagnostics read it, and `get()` / `numpy.asarray(field)` copy it to the
host as complex128 (the reference's cp.asnumpy)."""

from __future__ import annotations

import numpy as np


class DeviceField:
    def __init__(self, solve):
        self.solve = solve

    @property
    def shape(self):
        return (self.solve.n,) * 3

    @property
    def size(self):
        return self.solve.n ** 3

    @property
    def dtype(self):
        return np.dtype(np.complex64 if self.solve.precision == 0 else np.complex128)

    def get(self):
        """Host copy as complex128."""
        return self.solve.to_numpy()

    def __array__(self, dtype=None, copy=None):
        out = self.get()
        return out if dtype is None else out.astype(dtype)

    def slab(self, start, stop):
        """Host copy of x-planes [start, stop) (flat C-order slab)."""
        n = self.solve.n
        return self.solve.download(start * n * n, (stop - start) * n * n).reshape(stop - start, n, n)

    def set(self, values):
        self.solve.upload(np.asarray(values))

    def __repr__(self):
        return f"DeviceField(shape={self.shape}, dtype={self.dtype})"


def asnumpy(value):
    """cp.asnumpy equivalent for DeviceField and host arrays."""
    if isinstance(value, DeviceField):
        return value.get()
    return np.asarray(value)
