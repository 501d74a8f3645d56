"""Build device solves from reference-style parameters.

`make_solve` is what `stage.prepare` uses: it creates a libgpe_hip solve
(after tuning the pass variants of its shape once per context, see tune.py),
plans the static and time-dependent potential expressions into per-target
programs (see expressions.plan_potential) and registers them.
"""

from __future__ import annotations

import numpy as np

from . import _lib, tune
from .expressions import TGT_VOL, TGT_XY, TGT_Z, plan_potential

PRECISIONS = {"c64": _lib.C64, "complex64": _lib.C64, "c128": _lib.C128, "complex128": _lib.C128}


def precision_code(precision):
    if isinstance(precision, int):
        return precision
    try:
        return PRECISIONS[str(precision).lower()]
    except KeyError:
        raise ValueError(f"precision must be one of {sorted(PRECISIONS)}") from None


def make_solve(n, x_array, k_array, kind, dt, beta, static=None, dynamic=None, kcut=None,
               precision="c128", ctx=None):
    """Create a solve and load its potential programs.

    kind: "real" or "imaginary"; static / dynamic: schema-2 expressions (or
    None); kcut: cutoff wavenumber (real time only) or None."""
    kind_code = _lib.IMAG if kind == "imaginary" else _lib.REAL
    ctx = ctx or _lib.default_context()
    # pass variants: tuned once per context, shape and step plan, then the
    # context default (with a cutoff on the plane plan, x_kin alone: the
    # plane passes have one shape per side)
    plan_code = _lib.default_plan(precision_code(precision), kind_code, masked=kcut is not None, n=n)
    if kcut is not None and plan_code == _lib.PLAN_PLANE:
        plan_code = tune.PLAN_PLANE_CUT
    tune.ensure_tuned(n, precision_code(precision), kind_code, ctx, plan=plan_code)
    solve = _lib.Solve(n, precision_code(precision), kind_code, dt, beta, x_array, k_array, kcut=kcut, ctx=ctx)
    plan = plan_potential(static, dynamic if kind_code == _lib.REAL else None)
    for target, program in plan.static.items():
        solve.potential_static({TGT_XY: _lib.TGT_XY, TGT_Z: _lib.TGT_Z, TGT_VOL: _lib.TGT_VOL}[target], program)
    for target, program in plan.dynamic.items():
        solve.potential_dynamic({TGT_XY: _lib.TGT_XY, TGT_Z: _lib.TGT_Z, TGT_VOL: _lib.TGT_VOL}[target], program)
    return solve


def midpoint_times(first_index, nsteps, dt):
    """t_mid of steps first..first+nsteps-1: float(i*dt) + dt/2, computed with
    the same float64 operations as gpe/launch.py:10 and gpe/make_propagator.py:94."""
    return np.array([float(i * dt) + dt / 2 for i in range(first_index, first_index + nsteps)],
                    dtype=np.float64)
