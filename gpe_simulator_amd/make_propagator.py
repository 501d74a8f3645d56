"""Propagator factories: the drop-in for gpe/make_propagator.py.

`make_real_propagator(stage)` / `make_imag_propagator(stage)` return the same
callables as the reference (gpe/make_propagator.py:5-152): `propagator(stage)`
for static and imaginary stages, `propagator(stage, t)` when a dynamic
potential is active, each advancing one Strang step and returning the
wavefunction.  The step runs in libgpe_hip on the stage's device solve.  The
callable also has `advance(stage, first_index, count)`, which `launch_stage`
uses to run `count` steps in one call (the engine then fuses the second
half-step of each step with the first of the next).
"""

from __future__ import annotations

import numpy as np

from .engine import midpoint_times
from .field import DeviceField


def _check_contract(stage, solve):
    """The reference's pre-launch assertions (gpe/make_propagator.py:20-36,
    66-81, 117-133), restated for the device field."""
    assert stage.solve is solve, "the stage was prepared again; build a new propagator"
    psi = stage.psi
    assert isinstance(psi, DeviceField) and psi.solve is solve, "stage.psi must be the stage's device field"
    assert psi.shape == (stage.N_grids,) * 3
    assert stage.x_array.dtype == np.float64 and stage.x_array.shape == (stage.N_grids,)
    assert stage.k_array.dtype == np.float64 and stage.k_array.shape == (stage.N_grids,)
    assert float(stage.dt) == solve.dt, "stage.dt changed after prepare()"
    assert float(stage.beta) == solve.beta, "stage.beta changed after prepare()"


def make_real_propagator(stage):
    solve = stage.solve
    dynamic = bool(stage.potential_params["dynamic"]["activate"])

    if not dynamic:
        def time_independent_propagator(stage):
            _check_contract(stage, solve)
            solve.step(1)
            return stage.psi
        propagator = time_independent_propagator
    else:
        def time_dependent_propagator(stage, t):
            _check_contract(stage, solve)
            solve.step(1, np.array([float(t) + float(stage.dt) / 2]))
            return stage.psi
        propagator = time_dependent_propagator

    def advance(stage, first_index, count):
        _check_contract(stage, solve)
        solve.step(count, midpoint_times(first_index, count, stage.dt) if dynamic else None)
        return stage.psi

    propagator.advance = advance
    return propagator


def make_imag_propagator(stage):
    solve = stage.solve

    def standard_imaginary_step(stage):
        _check_contract(stage, solve)
        solve.step(1)
        return stage.psi

    def advance(stage, first_index, count):
        _check_contract(stage, solve)
        solve.step(count)
        return stage.psi

    standard_imaginary_step.advance = advance
    return standard_imaginary_step
