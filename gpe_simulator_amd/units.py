"""Harmonic-oscillator units <-> the reference's units.

Workloads are often stated for H' = -1/2 lap + 1/2 sum w^2 x^2 + g|psi|^2;
the reference integrates H = -lap + V + beta|psi|^2 (gpe/kernels.py:15-17,
gpe/helper.py:109-113).  H = 2 H' with lengths unchanged, so
V = sum w^2 x^2, beta = 2 g, dt = dt'/2, t = t'/2 (a drive at frequency W'
runs at W = 2 W'), and E = 2 E', mu = 2 mu'.
"""

from __future__ import annotations

import math

A0 = 5.29177210544e-11
LENGTH_UNIT = 1e-6


def beta_from_g(g):
    return 2.0 * g


def dt_from_oscillator(dt):
    return dt / 2.0


def energy_to_oscillator(e):
    return e / 2.0


def harmonic_expression(wx=1.0, wy=1.0, wz=1.0):
    """Reference-unit trap sum w^2 x^2 as a schema-2 expression."""
    def term(w, var):
        c = float(w) * float(w)
        return f"{var}**2" if c == 1.0 else f"{repr(c)}*{var}**2"
    return " + ".join((term(wx, "x"), term(wy, "y"), term(wz, "z")))


def scattering_length_for_beta(beta, n_atom=150000):
    """`a` (Bohr radii) giving `beta` at reference N_atom (inverse of get_beta)."""
    return beta * LENGTH_UNIT / (8 * math.pi * n_atom * A0)


def stirrer_expression(amplitude=5.0, radius=4.0, omega=0.5, width2=2.0):
    """Gaussian beam amplitude*exp(-((x-R cos W t)^2+(y-R sin W t)^2)/width2)
    moving on a circle, given in oscillator units and returned in reference
    units (potential doubled, angular speed doubled)."""
    a, w, radius, width2 = 2.0 * float(amplitude), 2.0 * float(omega), float(radius), float(width2)
    return (f"{repr(a)}*exp(-((x-{repr(radius)}*cos({repr(w)}*t))**2"
            f"+(y-{repr(radius)}*sin({repr(w)}*t))**2)/{repr(width2)})")
