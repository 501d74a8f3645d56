"""Norm, renormalisation, energies, angular momentum and schedules on the device.

Counterparts of gpe/helper.py:81-156 of the reference; every reduction runs in
libgpe_hip (FP64 partial sums, deterministic finish)."""

from __future__ import annotations

import numpy as np

from .field import DeviceField

A0 = 5.29177210544e-11
LENGTH_UNIT = 1e-6


def _field(psi):
    if not isinstance(psi, DeviceField):
        raise TypeError("psi must be a DeviceField (the engine keeps the wavefunction on the device)")
    return psi


def get_N(psi, x_array=None):
    """FP64 norm dx^3 * sum |psi|^2 (gpe/helper.py:81-88)."""
    return _field(psi).solve.norm()


def normalise(psi, x_array=None):
    """psi /= sqrt(norm) on the device (gpe/helper.py:91-93)."""
    _field(psi).solve.normalise()


def get_beta(N_atoms, a):
    """beta = 8*pi*N_atom*a*a0/l (gpe/helper.py:109-113)."""
    return 8 * np.pi * N_atoms * (a * A0) / LENGTH_UNIT


def is_due(i, interval):
    """Output schedule predicate (gpe/helper.py:131-132)."""
    return i > 0 and interval is not None and interval > 0 and i % interval == 0


class DeviceSpectrum:
    """fftn(psi) of a device field, as the reference's callers hold it
    (psi_k = cp.fft.fftn(psi), gpe/save.py:31).  The transform itself runs
    inside the engine's diagnostic passes when a consumer (get_E,
    nk_extraction) reads it, so no second field is materialised; `get()`
    returns the raw unnormalised transform on the host."""

    def __init__(self, field):
        self.field = _field(field)

    @property
    def shape(self):
        return self.field.shape

    @property
    def size(self):
        return self.field.size

    def get(self):
        return self.field.solve.spectrum()

    def __array__(self, dtype=None, copy=None):
        out = self.get()
        return out if dtype is None else out.astype(dtype)


def fftn(psi):
    """The engine's counterpart of cp.fft.fftn(psi) for a DeviceField."""
    return DeviceSpectrum(psi)


def _solve_of(psi_or_k, name):
    if isinstance(psi_or_k, DeviceSpectrum):
        return psi_or_k.field.solve
    if isinstance(psi_or_k, DeviceField):
        return psi_or_k.solve
    raise TypeError(f"{name} takes fftn(psi) of the engine's field (gpe_simulator_amd.helper.fftn) "
                    "or the DeviceField itself; the wavefunction stays on the device")


def get_E(psi_k, x_array=None, k_array=None, t=0.0):
    """Kinetic energy dx^3/N^3 * sum k^2 |fftn(psi)|^2 of the completed state
    (gpe/helper.py:135-140).  Takes fftn(psi) as the reference does (a
    DeviceSpectrum), or the real-space DeviceField: the forward FFT runs inside
    the engine's diagnostic passes either way."""
    return _solve_of(psi_k, "get_E").diagnostics(t, spectrum=False)["E_kin"]


def get_L(psi, x_array=None, k_array=None, slab_size=16):
    """Lz = Re int psi^* (x p_y - y p_x) psi dV with spectral derivatives
    (gpe/helper.py:143-156), from FP64 device sums (slab_size is accepted for
    the reference's signature; the engine needs no slabs)."""
    return _field(psi).solve.angular_momentum()


def energies(psi, t=0.0):
    """norm, E_kin, E_pot, E_int, E, mu and the Parseval norm in one pass set."""
    return _field(psi).solve.diagnostics(t, spectrum=False)
