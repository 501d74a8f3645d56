"""Stage driver: gpe/launch.py:4-14 of the reference.

`launch_stage` keeps the reference's contract (outputs at step 0 and after each
step; stage clock t_i = i*dt; returns the final wavefunction).  Steps between
two output-due steps are handed to the engine in one call when the propagator
supports it; outputs are identical because save_outputs writes nothing on a
step where nothing is due.
"""

from __future__ import annotations

from .helper import is_due
from .save import save_outputs

MAX_STEPS_PER_CALL = 4096


def _intervals(stage):
    save, monitor = stage.save_params, stage.monitor_params
    return [save["psi_interval"], save["scalar_interval"], save["nk_interval"], *monitor.values()]


def next_due(stage, after):
    """First step s in (after, steps] at which save_outputs has work."""
    intervals = [iv for iv in _intervals(stage) if iv]
    limit = min(stage.steps, after + MAX_STEPS_PER_CALL)
    for s in range(after + 1, limit + 1):
        if s == stage.steps or any(is_due(s, iv) for iv in intervals):
            return s
    return limit


def launch_stage(stage):
    """Run a prepared stage and return its final device wavefunction."""
    save_outputs(stage, 0, initial=True)
    dynamic = stage.potential_params["dynamic"]["activate"]
    advance = getattr(stage.propagator, "advance", None)
    i = 0
    while i < stage.steps:
        if advance is not None:
            j = next_due(stage, i)
            stage.psi = advance(stage, i, j - i)
            i = j
        else:
            stage.psi = stage.propagator(stage, i * stage.dt) if dynamic else stage.propagator(stage)
            i += 1
        save_outputs(stage, i, final=(i == stage.steps))
    return stage.psi
