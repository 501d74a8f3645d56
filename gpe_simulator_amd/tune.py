"""Autotuner for the pass tile variants.

Every pass kind of a step (separable plan: zk_first, zk_mid, yk, xk, xk_last;
four-sweep plan: z_first, z_mid, z_last, y_fwd, y_inv, x_kin; the plane
plan's plane passes have one shape per side, so it tunes xk and xk_last, or
with a momentum cutoff x_kin alone) reads and writes
the field in the same canonical layout, so each kind can run any of the
compiled tile variants of its (side, precision): threads per pencil, pencils
per block and the waves per SIMD its register budget allows (gpe_shapes.h).
Each step plan is tuned on its own, when a solve of that plan is first made.
The tuner works on scratch solves of the real shape (alone,
or as a batch of `batch` members, since small grids behave differently when
many solves share a launch) in two stages:

1. every variant is timed with HIP events around every pass, giving the
   fastest variant per kind;
2. that choice and the uniform choices (one variant for every kind) are timed
   as whole steps, and the fastest is refined by coordinate descent, one kind
   at a time, on whole-step time.

The second stage is what decides.  The per-pass events cost a few µs per
launch, which is most of a pass at 64^3-128^3, and a pass's speed depends on
what ran before it.  There the first stage alone chose up to 10% slower steps
than the descent.  The choice becomes the context's default for new solves
of that (side, precision, kind).

Results are cached on the context; variants a caller set on the context
explicitly are never replaced.  Set the environment variable
GPE_HIP_AUTOTUNE=0 to keep variant 0 everywhere.
"""

from __future__ import annotations

import os
import threading
import time
from dataclasses import dataclass, field

from . import _lib
from .grids import get_grids

# The plane plan of a solve with a momentum cutoff (a tuner key only; the
# library calls it PLAN_PLANE): its ZY plane passes have one shape per side,
# so x_kin is the one kind to tune, in scratch solves that carry a cutoff.
PLAN_PLANE_CUT = "plane_cut"
STEP_KINDS = {_lib.PLAN_SEPARABLE: ("zk_first", "zk_mid", "yk", "xk", "xk_last"),
              _lib.PLAN_FOUR_SWEEP: ("z_first", "z_mid", "z_last", "y_fwd", "y_inv", "x_kin"),
              # (the plane passes have one shape per side; only the x passes vary)
              _lib.PLAN_PLANE: ("xk", "xk_last"),
              PLAN_PLANE_CUT: ("x_kin",)}
# pass kinds of a step in the order the descent visits them (most time first)
STEP_SEQUENCE = {
    _lib.PLAN_SEPARABLE: {_lib.REAL: ("zk_mid", "xk", "yk", "zk_first", "xk_last"),
                          _lib.IMAG: ("xk_last", "zk_first", "yk")},
    _lib.PLAN_FOUR_SWEEP: {_lib.REAL: ("z_mid", "x_kin", "y_fwd", "y_inv", "z_first", "z_last"),
                           _lib.IMAG: ("z_first", "z_last", "x_kin", "y_fwd", "y_inv")},
    _lib.PLAN_PLANE: {_lib.REAL: ("xk", "xk_last"), _lib.IMAG: ("xk_last",)},
    PLAN_PLANE_CUT: {_lib.REAL: ("x_kin",)},
}
PLAN_NAMES = {_lib.PLAN_SEPARABLE: "separable", _lib.PLAN_FOUR_SWEEP: "four-sweep", _lib.PLAN_PLANE: "plane",
              PLAN_PLANE_CUT: "plane with cutoff"}
GAIN = 0.005        # a descent move must shorten the step by more than this fraction
_lock = threading.Lock()


@dataclass
class Tuning:
    n: int
    precision: int
    kind: int
    batch: int
    plan: int
    variants: list                       # [(tp, pencils per block, waves)]
    choice: list                         # variant per pass kind (PASS_KINDS order)
    timings: dict = field(default_factory=dict)   # variant -> {kind: avg us}
    event_choice: list = field(default_factory=list)   # fastest variant per kind by pass events
    step_us: float = float("nan")        # whole-step time of `choice`
    event_step_us: float = float("nan")  # whole-step time of `event_choice`

    def report(self):
        kinds = STEP_KINDS[self.plan]
        lines = [f"n={self.n} precision={'c64' if self.precision == _lib.C64 else 'c128'} "
                 f"kind={'imaginary' if self.kind == _lib.IMAG else 'real'} batch={self.batch} "
                 f"plan={PLAN_NAMES[self.plan]}"]
        for v, geo in enumerate(self.variants):
            t = self.timings.get(v, {})
            lines.append(f"  v{v} tp={geo[0]} b={geo[1]} waves={geo[2]}: " +
                         ", ".join(f"{k} {t[k]:.2f}" for k in kinds if k in t))
        lines.append("  choice: " + ", ".join(f"{k}=v{self.choice[_lib.PASS_KINDS.index(k)]}" for k in kinds)
                     + f" ({self.step_us:.2f} us/step; per-pass choice {self.event_step_us:.2f})")
        return "\n".join(lines)


def enabled():
    return os.environ.get("GPE_HIP_AUTOTUNE", "1") != "0"


def _plan_code(plan):
    """A tuner plan key as the library's plan (or the tuner-only key)."""
    return plan if plan == PLAN_PLANE_CUT else int(plan)


def _scratch(n, precision, kind, ctx, plan):
    length = 20.0
    x, k = get_grids(n, length)
    cut = plan == PLAN_PLANE_CUT
    s = _lib.Solve(n, precision, kind, 5e-4, 100.0, x, k, kcut=0.6 * k.max() if cut else None, ctx=ctx)
    s.set_plan(_lib.PLAN_PLANE if cut else plan)
    from .expressions import plan_potential
    for target, program in plan_potential("x**2 + y**2 + z**2", None).static.items():
        s.potential_static(target, program)
    s.gaussian(1.5)
    return s


def measure(n, precision, kind, steps=12, batch=1, rounds=3, ctx=None, plan=_lib.PLAN_SEPARABLE):
    """Average µs per launch of every step pass kind of `plan`, for every
    variant: {variant: {kind: us}} (the minimum over `rounds` timed runs)."""
    ctx = ctx or _lib.default_context()
    variants = _lib.pass_variants(n, precision)
    solves = [_scratch(n, precision, kind, ctx, plan) for _ in range(batch)]
    group = _lib.Batch(solves, ctx=ctx) if batch > 1 else None
    out = {}
    try:
        for v in range(len(variants)):
            for s in solves:
                s.set_variants([v] * len(_lib.PASS_KINDS))
            best = {}
            for _ in range(rounds):
                if group is not None:
                    group.step(2, graph=False)
                    timing = group.step_timed(steps)
                else:
                    solves[0].step(2)
                    timing = solves[0].step_timed(steps)
                for name, (ms, count) in timing.items():
                    if count and name in STEP_KINDS[plan]:
                        us = ms / count * 1e3
                        best[name] = min(best.get(name, us), us)
            out[v] = best
    finally:
        if group is not None:
            group.close()
        for s in solves:
            s.close()
    return variants, out


class StepClock:
    """Whole-step time of scratch solves under a given choice of variants:
    device time between two marks on a solo solve's stream, or host time
    around synchronous graph replays of a batch."""

    def __init__(self, n, precision, kind, batch, ctx, steps=None, rounds=3, plan=_lib.PLAN_SEPARABLE):
        self.solves = [_scratch(n, precision, kind, ctx, plan) for _ in range(batch)]
        self.group = _lib.Batch(self.solves, ctx=ctx) if batch > 1 else None
        # about 2e8 grid-point-steps (a few ms) per timed round
        self.steps = steps or max(8, min(100, int(2e8 / (n ** 3 * batch))))
        self.rounds = rounds

    def __call__(self, choice):
        for s in self.solves:
            s.set_variants(choice)
        best = float("inf")
        if self.group is None:
            s = self.solves[0]
            s.step(2)
            for _ in range(self.rounds):
                s.mark(0)
                s.step(self.steps)
                s.mark(1)
                best = min(best, s.marked_ms() * 1e3 / self.steps)
        else:
            self.group.step(2, graph=True)          # captures the graph of this choice
            for _ in range(self.rounds):
                t0 = time.perf_counter()
                self.group.step(self.steps, graph=True)
                best = min(best, (time.perf_counter() - t0) * 1e6 / self.steps)
        return best

    def close(self):
        if self.group is not None:
            self.group.close()
        for s in self.solves:
            s.close()


def refine(clock, start, kind, nvariants, plan=_lib.PLAN_SEPARABLE):
    """Fastest whole step from `start` and the uniform choices (over the plan's
    pass kinds; other kinds keep `start`), improved by coordinate descent over
    the step's pass kinds (two sweeps, moves that gain more than GAIN).
    Returns (choice, us per step)."""
    idxs = [_lib.PASS_KINDS.index(k) for k in STEP_KINDS[plan]]

    def uniform(v):
        c = list(start)
        for i in idxs:
            c[i] = v
        return c

    candidates = [list(start)] + [uniform(v) for v in range(nvariants)]
    timed = [(clock(c), c) for c in candidates]
    best_us, best = min(timed, key=lambda item: item[0])
    for _ in range(2):
        moved = False
        for name in STEP_SEQUENCE[plan][kind]:
            idx = _lib.PASS_KINDS.index(name)
            for v in range(nvariants):
                if v == best[idx]:
                    continue
                trial = list(best)
                trial[idx] = v
                us = clock(trial)
                if us < best_us * (1 - GAIN):
                    best, best_us, moved = trial, us, True
        if not moved:
            break
    return best, best_us


def autotune(n, precision, kind, steps=12, batch=1, rounds=3, ctx=None, plan=None, start=None):
    """Time every variant per pass of `plan` (default: the plan new solves of
    this precision and kind step with), then decide on whole-step time (a
    Tuning).  Kinds outside the plan keep `start` (default: variant 0)."""
    precision = int(precision)
    kind = int(kind)
    plan = _lib.default_plan(precision, kind, n=n) if plan is None else _plan_code(plan)
    ctx = ctx or _lib.default_context()
    variants, timings = measure(n, precision, kind, steps=steps, batch=batch, rounds=rounds, ctx=ctx, plan=plan)
    event_choice = list(start) if start is not None else [0] * len(_lib.PASS_KINDS)
    for name in STEP_KINDS[plan]:
        measured = [(timings[v][name], v) for v in timings if name in timings[v]]
        if measured:
            event_choice[_lib.PASS_KINDS.index(name)] = min(measured)[1]
    clock = StepClock(n, precision, kind, batch, ctx, plan=plan)
    try:
        event_us = clock(event_choice)
        choice, step_us = refine(clock, event_choice, kind, len(variants), plan=plan)
    finally:
        clock.close()
    return Tuning(n=n, precision=precision, kind=kind, batch=batch, plan=plan, variants=variants, choice=choice,
                  timings=timings, event_choice=event_choice, step_us=step_us, event_step_us=event_us)


def ensure_tuned(n, precision, kind, ctx=None, batch=1, plan=None):
    """Tune (side, precision, kind) for `plan` once per context and, for solo
    tuning, make the choice the context's default for new solves (the plan's
    kinds only; a tuning of the other plan keeps its own) unless the caller
    has set that default.  Returns the Tuning, or None when tuning is
    disabled, the side has a single variant or the default was set."""
    if not enabled():
        return None
    ctx = ctx or _lib.default_context()
    n, precision, kind = int(n), int(precision), int(kind)
    plan = _lib.default_plan(precision, kind, n=n) if plan is None else _plan_code(plan)
    key = (n, precision, kind, int(batch)) if plan == _lib.default_plan(precision, kind, n=n) else \
        (n, precision, kind, int(batch), plan)
    with _lock:
        if key in ctx.tunings:
            return ctx.tunings[key]
        if len(_lib.pass_variants(n, precision)) < 2 or (batch == 1 and key[:3] in ctx.variant_defaults):
            ctx.tunings[key] = None
            return None
        result = autotune(n, precision, kind, batch=batch, ctx=ctx, plan=plan,
                          start=ctx.default_variants(n, precision, kind))
        if batch == 1:
            ctx.set_variants(n, precision, kind, result.choice, tuned=True)
        ctx.tunings[key] = result
        return result
