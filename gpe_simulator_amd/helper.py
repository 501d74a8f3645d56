"""Norm, renormalisation, energies and schedules on the device.

Counterparts of gpe/helper.py:81-140 of the reference; every reduction runs in
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


def get_E(psi, x_array=None, k_array=None, t=0.0):
    """Kinetic energy dx^3/N^3 * sum k^2 |fftn(psi)|^2 of the completed state
    (gpe/helper.py:135-140).  Takes the real-space field: the forward FFT runs
    inside the engine's diagnostic passes."""
    return _field(psi).solve.diagnostics(t, spectrum=False)["E_kin"]


def energies(psi, t=0.0):
    """norm, E_kin, E_pot, E_int, E, mu and the Parseval norm in one pass set."""
    return _field(psi).solve.diagnostics(t, spectrum=False)
