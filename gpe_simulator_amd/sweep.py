"""Batched parameter sweeps: imaginary-time ground states and stirring runs.

Ground states: many independent solves advance together, one batch per
(phase, grid side): every pass is one launch over all members of a batch
(gpe_batch_*), and the check_every-step launch sequence is replayed from a
captured hipGraph.  The bulk of the relaxation runs with complex64 fields
and FP64 reductions; when its Switch says so, a member is copied into a
complex128 field and polished, still batched, until |dmu| < tol.  All
groups are launched on their own streams each round and overlap on the
device.

Stirring runs (stir_batched): real-time runs whose stirrers differ (speed,
radius, amplitude) advance as one batch; each member evaluates its own
time-dependent beam before every chunk of steps.

Convergence uses the per-step norm the engine records before renormalising:
starting from a normalised state, one imaginary-time step of length dtau
multiplies the norm by exp(-2 mu dtau) near a stationary state, so
mu_est = -ln(N) / (2 dtau) costs nothing extra.  The converged result also
reports mu and E from the FP64 energy functional.  The reference has no
convergence test (imaginary stages run a fixed step count,
gpe/launch.py:8); this module is new.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field

import numpy as np

from . import _lib, tune
from .engine import make_solve, precision_code
from .grids import get_grids


@dataclass
class GroundState:
    """One imaginary-time request (reference units)."""
    n: int
    length: float
    beta: float
    static: str
    sigma: float
    dt: float
    tol: float = 1e-9
    max_steps: int = 400_000


@dataclass
class Result:
    mu: float = float("nan")          # FP64 energy functional
    energy: float = float("nan")
    mu_estimate: float = float("nan")  # norm-decay estimate at convergence
    dmu: float = float("nan")
    steps: int = 0
    steps_c64: int = 0
    converged: bool = False
    seconds: float = 0.0
    history: list = field(default_factory=list)


def mu_from_norms(norms, dt):
    norms = np.asarray(norms)
    if not np.all(np.isfinite(norms) & (norms > 0)):
        raise FloatingPointError(f"recorded norms must be finite and positive, got {norms[~(norms > 0)][:4]} "
                                 f"at positions {np.flatnonzero(~(norms > 0))[:4]} of {norms.size}")
    return -np.log(norms) / (2.0 * dt)


class FieldFloor:
    """Where a complex64 field stops relaxing, as a complex128 |dmu|.

    complex64 rounding keeps re-exciting the slow modes, so a complex64 field
    settles at a level that the complex128 estimate reads as |dmu| ~ F, far
    below the complex64 estimate's own floor.  It is measured on the first
    members to switch, all switched late (once their fitted decay reaches
    tol): the complex128 |dmu| at their second check after the switch is
    that level.  The median of `samples` such readings is used, and no later
    reading (from members switched at F itself) feeds back into it."""

    def __init__(self, samples=4):
        self.samples, self.readings = samples, []

    @property
    def value(self):
        return float(np.median(self.readings)) if len(self.readings) >= self.samples else None

    def add(self, dmu):
        if len(self.readings) < self.samples:
            self.readings.append(dmu)


class Switch:
    """When to move a complex64 imaginary-time solve to complex128.

    The norm-decay estimate is exact up to rounding, so in complex64 |dmu|
    falls exponentially (the slowest excited mode decays as exp(-2 gap t))
    until it meets a rounding floor (about 1e-6 in reference units at
    dtau = 5e-4), while the field itself keeps relaxing below that floor.
    Once the floor shows (no new minimum in `plateau` checks), the clean
    part of the history (|dmu| > 30x the floor) is fitted by a line in
    log|dmu|.  The solve stays in complex64 until the line reaches `tol`, or
    the complex64 field's own floor (a FieldFloor, once measured) if that is
    higher: from there on complex64 steps no longer bring the field closer.
    The complex128 polish then runs until its own, noise-free |dmu| < tol.
    Without a usable fit the switch happens at once.  A fit is trusted for
    at most `horizon` times the steps it took to see the floor: a line that
    predicts a longer complex64 phase was fitted to a transient (a slowly
    relaxing member whose clean history is short), and the solve switches
    at that bound instead."""

    def __init__(self, tol, plateau=10, clean_factor=30.0, fit_points=6, field_floor=None, horizon=1.5):
        self.tol, self.plateau, self.clean_factor, self.fit_points = tol, plateau, clean_factor, fit_points
        self.field_floor, self.horizon = field_floor, horizon
        self.steps, self.dmu = [], []
        self.fit = None                  # (slope, intercept) of log|dmu| against steps
        self.detected = None             # step at which the floor was recognised
        self.late = False                # switched where the line reached tol

    def update(self, steps, dmu):
        """Record the check at `steps`; True when the solve should switch."""
        self.steps.append(steps)
        self.dmu.append(dmu)
        if dmu < self.tol:
            self.late = True
            return True
        if self.detected is None:
            p = self.plateau
            if len(self.dmu) <= p or min(self.dmu[-p:]) < min(self.dmu[:-p]):
                return False
            self.detected = steps
            self.fit = self.predict()
            if self.fit is None:
                return True
        floor = self.field_floor.value if self.field_floor is not None else None
        level = max(self.tol, floor) if floor is not None else self.tol
        self.late = floor is None
        return steps >= min(self.step_at(level), self.horizon * self.detected)

    def step_at(self, level):
        slope, icpt = self.fit
        return int(np.ceil((np.log(level) - icpt) / slope))

    @property
    def target(self):
        """Step where the fitted line reaches tol (None before the fit)."""
        return self.step_at(self.tol) if self.fit else None

    def predict(self):
        floor = float(np.median(self.dmu[-self.plateau:]))
        clean = [(s, d) for s, d in zip(self.steps, self.dmu) if d > self.clean_factor * floor][-self.fit_points:]
        if len(clean) < 3:
            return None
        s, d = np.array(clean, dtype=float).T
        slope, icpt = np.polyfit(s, np.log(d), 1)
        return (slope, icpt) if slope < 0 else None


def _new_solve(req, precision, ctx=None):
    x, k = get_grids(req.n, req.length)
    return make_solve(req.n, x, k, "imaginary", req.dt, req.beta, static=req.static, precision=precision, ctx=ctx)


def _finish(solve, res, t0):
    diag = solve.diagnostics(0.0, spectrum=False)
    res.mu, res.energy = diag["mu"], diag["E"]
    res.seconds = time.perf_counter() - t0


def solve_sequential(req, precision="c128", check_every=50, ctx=None):
    """One solve alone, chunks of check_every steps, no batching or graphs."""
    t0 = time.perf_counter()
    solve = _new_solve(req, precision, ctx)
    solve.gaussian(req.sigma)
    res = Result()
    while res.steps < req.max_steps:
        solve.step(check_every)
        res.steps += check_every
        mu = mu_from_norms(solve.norm_history(check_every), req.dt)
        res.mu_estimate, res.dmu = float(mu[-1]), float(abs(mu[-1] - mu[-2]))
        if res.dmu < req.tol:
            res.converged = True
            break
    _finish(solve, res, t0)
    solve.close()
    return res


def sweep_batched(reqs, check_every=50, precision="c64", polish="c128", graph=True, ctx=None, plateau_checks=10):
    """Relax every request to |dmu| < its tol; returns one Result per request.

    With `polish` set, each member runs in `precision` until its Switch says
    so, then in `polish` until |dmu| < tol; without it, in `precision` only.
    Requests may have different grid sides: members are grouped by (phase,
    side), each group is one batch on its own stream, and every round
    advances every group by check_every steps."""
    t0 = time.perf_counter()
    ctx = ctx or _lib.default_context()
    phases = [precision] + ([polish] if polish else [])
    sides = sorted({r.n for r in reqs})
    # pass variants tuned for each batch shape (many members per launch)
    batch_variants = {}
    if tune.enabled():
        for n in sides:
            count = sum(r.n == n for r in reqs)
            for prec in phases:
                t = tune.ensure_tuned(n, precision_code(prec), _lib.IMAG, ctx=ctx, batch=min(count, 64))
                batch_variants[(prec, n)] = t.choice if t else None

    def new_member(req, prec):
        s = _new_solve(req, prec, ctx)
        if batch_variants.get((prec, req.n)):
            s.set_variants(batch_variants[(prec, req.n)])
        return s

    results = [Result() for _ in reqs]
    field_floor = {n: FieldFloor() for n in sides}
    switches = [Switch(r.tol, plateau_checks, field_floor=field_floor[r.n]) for r in reqs]
    phase_of = [0] * len(reqs)
    polish_checks = [0] * len(reqs)
    solves = []
    for req in reqs:
        s = new_member(req, precision)
        s.gaussian(req.sigma)
        solves.append(s)
    active = list(range(len(reqs)))
    batches = {}                       # (phase, side) -> (member tuple, Batch)
    rounds_of = {}                     # (phase, side) -> rounds its current batch has run
    keys = [(ph, n) for ph in range(len(phases)) for n in sides]
    while active:
        groups = {key: tuple(i for i in active if (phase_of[i], reqs[i].n) == key) for key in keys}
        # Retire stale batches before anything is launched: freeing device
        # memory waits for the whole device, so a batch closed while another
        # group runs would serialise the groups.
        for key in list(batches):
            if batches[key][0] != groups[key]:
                batches.pop(key)[1].close()
        launched = [(key, members) for key, members in groups.items() if members]
        for key, members in launched:
            if key not in batches:
                batches[key] = (members, _lib.Batch([solves[i] for i in members], ctx=ctx))
                rounds_of[key] = 0
        # A batch is captured into a graph once it has outlived one round:
        # capturing for a membership that changes again next round costs more
        # (graph instantiation) than the launches it saves.
        for key, _ in launched:
            batches[key][1].launch(check_every, graph=graph and rounds_of[key] > 0)
            rounds_of[key] += 1
        for key, _ in launched:
            batches[key][1].wait()
        still, promote = [], []
        for key, members in launched:
            ph = key[0]
            bulk = ph + 1 < len(phases)
            norms = batches[key][1].norm_history(check_every)
            for row, i in enumerate(members):
                req, res = reqs[i], results[i]
                res.steps += check_every
                if ph == 0 and bulk:
                    res.steps_c64 += check_every
                mu = mu_from_norms(norms[row], req.dt)
                res.mu_estimate, res.dmu = float(mu[-1]), float(abs(mu[-1] - mu[-2]))
                res.history.append(res.dmu)
                if res.steps >= req.max_steps:
                    continue
                if bulk and switches[i].update(res.steps, res.dmu):
                    promote.append(i)
                elif not bulk:
                    polish_checks[i] += 1
                    if polish_checks[i] == 2 and switches[i].late:
                        field_floor[req.n].add(res.dmu)
                    if res.dmu < req.tol:
                        continue
                still.append(i)
        for i in promote:
            for key in [k for k, (m, _) in batches.items() if i in m]:
                batches.pop(key)[1].close()
            phase_of[i] += 1
            s = new_member(reqs[i], phases[phase_of[i]])
            s.copy_from(solves[i])
            solves[i].close()
            solves[i] = s
        keep = set(still)
        active = [i for i in active if i in keep]
    for _, b in batches.values():
        b.close()
    for i, req in enumerate(reqs):
        results[i].converged = results[i].dmu < req.tol and phase_of[i] == len(phases) - 1
        _finish(solves[i], results[i], t0)
        solves[i].close()
    return results


def config_sweep(count=64, seed=0, n=96, length=20.0, dt_osc=1e-3, tol_osc=1e-9):
    """Seed-drawn sweep in oscillator units mapped to reference units:
    g ~ U[50, 1000], omega_z ~ U[0.5, 4] (omega_x = omega_y = 1),
    Gaussian width ~ U[0.8, 1.5]; beta = 2g, V = x^2 + y^2 + omega_z^2 z^2,
    dtau = dtau'/2, and the mu tolerance doubles with mu."""
    from .units import harmonic_expression
    rng = np.random.default_rng(seed)
    g = rng.uniform(50, 1000, count)
    wz = rng.uniform(0.5, 4, count)
    sigma = rng.uniform(0.8, 1.5, count)
    return [GroundState(n=n, length=length, beta=2 * g[i], static=harmonic_expression(1, 1, wz[i]),
                        sigma=float(sigma[i]), dt=dt_osc / 2, tol=2 * tol_osc) for i in range(count)]


# --------------------------------------------------------- stirring sweeps --
@dataclass
class StirRun:
    """One real-time run with a moving stirrer (reference units)."""
    n: int
    length: float
    beta: float
    static: str
    dynamic: str
    dt: float
    steps: int


def stirrer_runs(omegas, n=128, length=24.0, g=500.0, amplitude=5.0, radius=3.0, width2=2.0, dt_osc=5e-4,
                 steps=2000):
    """Stirring runs at the stirrer angular speeds `omegas` (oscillator
    units), otherwise identical: isotropic trap and a Gaussian beam
    amplitude*exp(-d^2/width2) on a circle of `radius`.  Mapped to reference
    units (V and beam doubled, beta = 2g, dt = dt'/2, Omega = 2 Omega')."""
    from .units import harmonic_expression, stirrer_expression
    return [StirRun(n=n, length=length, beta=2 * g, static=harmonic_expression(1, 1, 1),
                    dynamic=stirrer_expression(amplitude, radius, omega, width2), dt=dt_osc / 2, steps=steps)
            for omega in omegas]


def _stir_solve(run, precision, ctx):
    x, k = get_grids(run.n, run.length)
    return make_solve(run.n, x, k, "real", run.dt, run.beta, static=run.static, dynamic=run.dynamic,
                      precision=precision, ctx=ctx)


def _stir_report(solve, run):
    t = run.steps * run.dt
    diag = solve.diagnostics(t, spectrum=False)
    return {"E": diag["E"], "mu": diag["mu"], "norm": diag["norm"], "Lz": solve.angular_momentum()}


def stir_batched(runs, psi0, precision="c64", ctx=None):
    """Advance every run from the same initial field psi0 (host complex128)
    in one batch (one launch per pass for all runs; each run evaluates its
    own stirrer).  Returns one report (E, mu, norm, Lz at the end) per run."""
    from .engine import midpoint_times
    ctx = ctx or _lib.default_context()
    if len({(r.n, r.length, r.steps) for r in runs}) != 1:
        raise ValueError("a stirring batch shares one grid and step count")
    solves = [_stir_solve(r, precision, ctx) for r in runs]
    solves[0].upload(psi0)
    for s in solves[1:]:
        s.copy_from(solves[0])
    batch = _lib.Batch(solves, ctx=ctx)
    t_mid = np.stack([midpoint_times(0, r.steps, r.dt) for r in runs])
    batch.step(runs[0].steps, t_mid=t_mid)
    batch.close()
    out = [_stir_report(s, r) for s, r in zip(solves, runs)]
    for s in solves:
        s.close()
    return out


def stir_sequential(run, psi0, precision="c64", ctx=None):
    """The same run alone (one solve, one gpe_step call)."""
    from .engine import midpoint_times
    solve = _stir_solve(run, precision, ctx or _lib.default_context())
    solve.upload(psi0)
    solve.step(run.steps, midpoint_times(0, run.steps, run.dt))
    out = _stir_report(solve, run)
    solve.close()
    return out
