"""Batched imaginary-time ground states (parameter sweeps).

Many independent solves of one grid side advance together: every pass is one
launch over all members (gpe_batch_*), and the check_every-step launch
sequence is replayed from a captured hipGraph.  The bulk of the relaxation
runs with complex64 fields and FP64 reductions; a member whose chemical
potential has settled below the complex64 noise floor (switch_tol) is copied
into a complex128 field and polished, still batched, until |dmu| < tol.

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

from . import _lib
from .engine import make_solve
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
    return -np.log(np.asarray(norms)) / (2.0 * dt)


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


def sweep_batched(reqs, check_every=50, switch_tol=1e-6, precision="c64", polish="c128", graph=True, ctx=None,
                  plateau_checks=20):
    """Relax every request to |dmu| < its tol; returns one Result per request."""
    t0 = time.perf_counter()
    ctx = ctx or _lib.default_context()
    if len({(r.n, r.length) for r in reqs}) != 1:
        raise ValueError("a sweep batches one grid; group requests by side and box first")
    results = [Result() for _ in reqs]
    solves = []
    for req in reqs:
        s = _new_solve(req, precision, ctx)
        s.gaussian(req.sigma)
        solves.append(s)
    phases = [(precision, switch_tol if polish else None)]
    if polish:
        phases.append((polish, None))
    active = list(range(len(reqs)))
    for phase, (prec, phase_tol) in enumerate(phases):
        batch, members = None, None
        while active:
            if members != active:
                if batch is not None:
                    batch.close()
                members = list(active)
                batch = _lib.Batch([solves[i] for i in members], ctx=ctx)
            batch.step(check_every, graph=graph)
            still = []
            for i in members:
                req, res = reqs[i], results[i]
                res.steps += check_every
                if phase == 0:
                    res.steps_c64 += check_every
                mu = mu_from_norms(solves[i].norm_history(check_every), req.dt)
                res.mu_estimate, res.dmu = float(mu[-1]), float(abs(mu[-1] - mu[-2]))
                limit = phase_tol if phase_tol is not None else req.tol
                if res.dmu < limit or res.steps >= req.max_steps:
                    continue
                if phase_tol is not None:
                    # leave the complex64 phase once dmu stops falling (its noise floor)
                    res.history.append(res.dmu)
                    recent = res.history[-plateau_checks:]
                    if len(res.history) > plateau_checks and min(recent) >= min(res.history[:-plateau_checks]):
                        continue
                still.append(i)
            active = still
        if batch is not None:
            batch.close()
        if phase + 1 < len(phases):
            nxt = phases[phase + 1][0]
            for i in range(len(reqs)):
                s = _new_solve(reqs[i], nxt, ctx)
                s.copy_from(solves[i])
                solves[i].close()
                solves[i] = s
            active = [i for i in range(len(reqs)) if results[i].steps < reqs[i].max_steps]
    for i, req in enumerate(reqs):
        results[i].converged = results[i].dmu < req.tol
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
