"""|psi(k)|^2 shell histogram (gpe/nk_extraction.py:16-45 of the reference).

The device bins every mode of fftn(psi) by shell floor(|k|/dk + 1/2) with an
integer-exact index and sums |F|^2 per shell in FP64; this module turns the
per-shell sums into the reference's output dictionary.

The same passes also bin the spectrum into VIEW_BINS (128) equal-width bins
over [0, k_max], k_max = sqrt(3) (n/2) dk the corner wavenumber, again with
an integer-exact index: a fixed-size view for comparing grids of different
sides (no reference counterpart)."""

from __future__ import annotations

import math

import numpy as np

from ._lib import VIEW_BINS
from .field import DeviceField


def shell_summary(counts, occupation, x_array, k_array, n):
    """Reference output fields from raw per-shell counts and sum |F|^2."""
    dx = float(x_array[1] - x_array[0])
    dk = abs(float(k_array[1] - k_array[0]))
    nbins = int(math.floor(math.sqrt(3) * (n // 2) + 0.5)) + 1
    counts = np.asarray(counts, dtype=np.int64)
    Nk = np.asarray(occupation, dtype=np.float64) * (dx ** 3 / n ** 3)
    nk = Nk / (np.maximum(counts, 1) * dk ** 3)
    k = np.arange(nbins, dtype=np.float64) * dk
    edges = (np.arange(nbins + 1, dtype=np.float64) - 0.5) * dk
    edges[0] = 0
    complete_shells = edges[1:] <= n * dk / 2
    return dict(k=k, nk=nk, Nk=Nk, counts=counts, edges=edges, complete_shells=complete_shells)


def _solve_of(psi_k):
    from .helper import DeviceSpectrum
    if isinstance(psi_k, DeviceSpectrum):
        return psi_k.field.solve
    if isinstance(psi_k, DeviceField):
        return psi_k.solve
    raise TypeError("nk_extraction takes fftn(psi) of the engine's field (gpe_simulator_amd.helper.fftn) "
                    "or the DeviceField itself; the wavefunction stays on the device")


def nk_extraction(psi_k, x_array, k_array, slab_size=16, t=0.0):
    """Shell averages of |fftn(psi)|^2; Nk sums to the wavefunction norm.

    `psi_k` is fftn(psi) as the reference passes it (a DeviceSpectrum from
    helper.fftn), or the real-space DeviceField: the transform is fused into
    the histogram passes either way, and slab_size is accepted for the
    reference's signature."""
    solve = _solve_of(psi_k)
    diag = solve.diagnostics(t, spectrum=True)
    return shell_summary(diag["counts"], diag["occupation"], x_array, k_array, solve.n)


def view_summary(counts, occupation, x_array, k_array, n):
    """Output fields of the equal-width view: bin centres k, edges, k_max,
    counts, Nk = sum |F|^2 dx^3/N^3 and nk = Nk / (count dk^3), the shell
    normalisation of the reference applied to these bins."""
    dx = float(x_array[1] - x_array[0])
    dk = abs(float(k_array[1] - k_array[0]))
    counts = np.asarray(counts, dtype=np.int64)
    Nk = np.asarray(occupation, dtype=np.float64) * (dx ** 3 / n ** 3)
    k_max = math.sqrt(3) * (n // 2) * dk
    width = k_max / VIEW_BINS
    return dict(k=(np.arange(VIEW_BINS) + 0.5) * width, edges=np.arange(VIEW_BINS + 1) * width,
                k_max=k_max, nk=Nk / (np.maximum(counts, 1) * dk ** 3), Nk=Nk, counts=counts)


def nk_view(psi_k, x_array, k_array, t=0.0):
    """The 128-bin equal-width view of |fftn(psi)|^2 (see view_summary)."""
    solve = _solve_of(psi_k)
    diag = solve.diagnostics(t, spectrum=True)
    return view_summary(diag["view_counts"], diag["view_occupation"], x_array, k_array, solve.n)
