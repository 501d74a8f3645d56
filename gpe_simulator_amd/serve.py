"""Continuous-batching solve server.

Requests (imaginary-time ground states to a mu tolerance, or fixed-length
real-time runs) arrive over time.  The scheduler keeps every admitted field
resident in HBM and loops:

  1. admits waiting requests while free HBM (gpe_mem_info) minus a reserve
     covers the request's resident + diagnostics footprint (gpe_solve_footprint);
  2. groups the idle active solves by (grid side, precision, kind, steps this
     tick) and enqueues each group's tick as one batch (one launch per pass
     for all members, replayed from a hipGraph) on its own HIP stream, so
     groups overlap on the device.  A tick is `check_every` steps, or fewer
     for a real-time run's last tick, which stops on its requested step
     count.  A large shape's tick goes to the device in pieces of about
     `launch_budget_s` each, each piece launched when the one before has
     finished, so that the launches of small groups, which share the
     process's few hardware queues, are not queued behind a whole large
     tick (while other groups are in flight; a group alone sends the rest
     of its tick at once).  At most one group of a shape is in flight: a
     solve admitted while its shape's group runs joins the group's next tick;
  3. polls the groups in flight (gpe_batch_done) and, for each one that has
     finished, reads each member's per-step norms, decides convergence
     (imaginary time: |dmu| < tol on the norm-decay estimate, with complex64
     members moving to a complex128 field when their Switch says so) or
     completion (real time: the requested step count), computes the FP64
     energies of finished solves, returns them and frees their memory.

Each group moves at its own pace: a 64^3 group's tick does not wait for a
256^3 group's, so a small request's latency follows its own work, not the
largest solve on the device.  A solve's arithmetic does not depend on its
batch-mates: every member has its own field, tables and reductions, and
every tick advances it by the same number of steps, so a result is the same
alone and in the mix.
"""

from __future__ import annotations

import time
from collections import deque
from dataclasses import dataclass

import numpy as np

from . import _lib
from .engine import make_solve
from .grids import get_grids
from .sweep import Switch, mu_from_norms
from .units import harmonic_expression


@dataclass
class Request:
    id: int
    arrival: float             # seconds after the stream starts
    n: int
    kind: str                  # "imaginary" or "real"
    precision: str             # "c64" or "c128"
    beta: float
    length: float = 20.0
    static: str = "x**2 + y**2 + z**2"
    dt: float = 5e-4
    sigma: float = 1.0
    tol: float = 2e-9          # imaginary time: |dmu| tolerance (reference units)
    steps: int = 2000          # real time: steps to run
    max_steps: int = 400_000


@dataclass
class Outcome:
    id: int
    n: int
    kind: str
    precision: str
    latency: float             # arrival -> result on the host, seconds
    queued: float              # arrival -> admission, seconds
    steps: int
    mu: float
    energy: float
    norm: float
    dmu: float
    converged: bool


@dataclass
class _Slot:
    req: Request
    solve: object
    precision: str
    admitted: float
    steps: int = 0
    dmu: float = float("inf")
    switch: Switch = None
    busy: bool = False         # a member of a group in flight


def _footprint(req, precision):
    resident, diag = _lib.footprint(req.n, _lib.C64 if precision == "c64" else _lib.C128,
                                    _lib.IMAG if req.kind == "imaginary" else _lib.REAL)
    return resident + diag


class Server:
    # seconds per grid point and step of a complex64 step (256^3 runs at about
    # 1.1e11 grid-point-steps/s), for sizing a tick's launches; complex128 twice
    STEP_SECONDS_PER_POINT = 9e-12

    def __init__(self, ctx=None, check_every=50, reserve_bytes=2 << 30, graph=True, launch_budget_s=2.5e-3):
        self.ctx = ctx or _lib.default_context()
        self.check_every = int(check_every)
        self.launch_budget_s = launch_budget_s
        self.reserve = reserve_bytes
        self.graph = graph
        self._batches = {}
        self.grid_point_steps = 0
        self.max_active = 0

    def warm(self, shapes):
        """Tune the pass variants of every (side, precision, kind) in `shapes`
        on the server's context before requests arrive, so that no request
        waits for tuning (make_solve tunes a shape the first time it sees it).
        precision: "c64" / "c128"; kind: "imaginary" / "real".  Complex64
        ground states also need the complex128 shape of their polish.
        Returns the seconds spent."""
        from . import tune
        from .engine import precision_code
        t0 = time.perf_counter()
        for n, precision, kind in shapes:
            tune.ensure_tuned(int(n), precision_code(precision), _lib.IMAG if kind == "imaginary" else _lib.REAL,
                              ctx=self.ctx)
        return time.perf_counter() - t0

    # -- solve lifecycle
    def _new_solve(self, req, precision):
        x, k = get_grids(req.n, req.length)
        return make_solve(req.n, x, k, req.kind, req.dt, req.beta, static=req.static, precision=precision,
                          ctx=self.ctx)

    def _admit(self, req, now):
        solve = self._new_solve(req, req.precision)
        solve.gaussian(req.sigma)
        return _Slot(req=req, solve=solve, precision=req.precision, admitted=now, switch=Switch(req.tol))

    def _batch_for(self, slots):
        key = tuple(id(s.solve) for s in slots)
        batch = self._batches.get(key)
        if batch is None:
            batch = _lib.Batch([s.solve for s in slots], ctx=self.ctx)
            self._batches[key] = batch
        return batch

    def _drop_batches(self, solve):
        for key in [k for k in self._batches if id(solve) in k]:
            self._batches.pop(key).close()

    # -- main loop
    def run(self, requests, realtime=True, log=None, poll_s=2e-5):
        """Serve `requests` (arrival-ordered) and return (outcomes, stats).
        realtime=False runs on a virtual clock: arrivals are not waited for,
        and the clock advances by the host time the loop spends."""
        for req in requests:
            if req.kind == "real" and req.steps < 1:
                raise ValueError(f"request {req.id}: a real-time run needs at least one step")
        start = time.perf_counter()
        incoming = deque(sorted(requests, key=lambda r: r.arrival))
        waiting, active, outcomes = deque(), [], []
        inflight = []              # [shape, tick steps, steps launched, last launch, members, batch]
        virtual = 0.0
        last = time.perf_counter()

        def now():
            nonlocal virtual, last
            if realtime:
                return time.perf_counter() - start
            t = time.perf_counter()
            virtual += t - last
            last = t
            return virtual

        while incoming or waiting or active:
            t = now()
            while incoming and incoming[0].arrival <= t:
                waiting.append(incoming.popleft())
            if not active and not waiting:
                if realtime:
                    time.sleep(max(0.0, incoming[0].arrival - t))
                else:
                    virtual = max(virtual, incoming[0].arrival)
                continue
            # admission by free device memory
            if waiting:
                free, _total = self.ctx.mem_info()
                for req in list(waiting):
                    need = _footprint(req, req.precision)
                    if req.precision == "c64" and req.kind == "imaginary":
                        need += _footprint(req, "c128")   # room for the complex128 polish
                    if free - self.reserve >= need or not active:
                        waiting.remove(req)
                        active.append(self._admit(req, t))
                        free -= need
                self.max_active = max(self.max_active, len(active))
            # launch the idle members of every shape without a group in flight;
            # a real-time member steps at most its remaining count, so a run of
            # any length ends on its requested step, and members with equal
            # step counts share a group
            running = {g[0] for g in inflight}
            groups = {}
            for slot in active:
                shape = (slot.req.n, slot.precision, slot.req.kind)
                if not slot.busy and shape not in running:
                    groups.setdefault(shape + (self._tick_steps(slot),), []).append(slot)
            # every batch is built (device allocation) before any group runs
            ready = [(key[:3], key[3], members, self._batch_for(members)) for key, members in groups.items()]
            for shape, steps, members, batch in ready:
                first = self._launch_steps(shape, steps)
                batch.launch(first, graph=self.graph)
                for slot in members:
                    slot.busy = True
                inflight.append([shape, steps, first, first, members, batch])
            # poll: launch the next piece of every group whose piece has finished
            # and handle every finished tick; sleep briefly when nothing has
            done = []
            for group in inflight:
                if not group[5].done():
                    continue
                shape, steps = group[0], group[1]
                if group[2] < steps:
                    # pieces one at a time while other groups share the device;
                    # alone, the rest of the tick at once (no gaps)
                    contended = len(inflight) > 1 or bool(waiting)
                    while group[2] < steps:
                        piece = min(self._launch_steps(shape, steps), steps - group[2])
                        group[5].launch(piece, graph=self.graph)
                        group[2], group[3] = group[2] + piece, piece
                        if contended:
                            break
                else:
                    done.append(group)
            if not done:
                time.sleep(poll_s)
                continue
            for group in done:
                inflight.remove(group)
                _shape, steps, _launched, last, members, batch = group
                # the last piece's norms: the estimate needs the tick's last two steps
                norms = batch.norm_history(last) if members[0].req.kind == "imaginary" else None
                finished = []
                for j, slot in enumerate(members):
                    slot.busy = False
                    slot.steps += steps
                    self.grid_point_steps += slot.req.n ** 3 * steps
                    req = slot.req
                    if req.kind == "imaginary":
                        try:
                            mu = mu_from_norms(norms[j], req.dt)
                        except FloatingPointError as err:
                            raise FloatingPointError(f"request {req.id} (n={req.n}, {slot.precision}, "
                                                     f"{slot.steps} steps): {err}") from None
                        slot.dmu = float(abs(mu[-1] - mu[-2]))
                        if slot.steps >= req.max_steps:
                            finished.append(slot)
                        elif slot.precision == "c64":
                            if slot.switch.update(slot.steps, slot.dmu):
                                self._promote(slot)
                        elif slot.dmu < req.tol:
                            finished.append(slot)
                    elif slot.steps >= req.steps:
                        finished.append(slot)
                for slot in finished:
                    outcomes.append(self._complete(slot, now()))
                    active.remove(slot)
                    if log:
                        log(outcomes[-1])
        elapsed = time.perf_counter() - start
        lat = np.array([o.latency for o in outcomes]) * 1e3
        stats = dict(requests=len(outcomes), seconds=elapsed,
                     grid_point_steps_per_s=self.grid_point_steps / elapsed if elapsed > 0 else 0.0,
                     p50_ms=float(np.percentile(lat, 50)) if len(lat) else None,
                     p99_ms=float(np.percentile(lat, 99)) if len(lat) else None,
                     max_active=self.max_active)
        for batch in self._batches.values():
            batch.close()
        self._batches.clear()
        return outcomes, stats

    def _launch_steps(self, shape, tick):
        """Steps of one launch of a `tick`-step tick of `shape` (side,
        precision, kind): the whole tick when it fits the launch budget, else
        the largest divisor of the tick that does (so the pieces are equal and
        one captured graph serves them), at least 2 (the mu estimate reads the
        last piece's last two norms).  A function of the shape alone, so a
        member's launches are the same alone and in the mix."""
        n, precision, _kind = shape
        step_s = n ** 3 * self.STEP_SECONDS_PER_POINT * (2.0 if precision == "c128" else 1.0)
        cap = max(2, int(self.launch_budget_s / step_s)) if self.launch_budget_s else tick
        if tick <= cap:
            return tick
        best = max(d for d in range(1, cap + 1) if tick % d == 0)
        return best if best >= max(2, cap // 2) else cap

    def _tick_steps(self, slot):
        """Steps `slot` takes this tick: check_every, or a real-time run's
        remaining steps when fewer are left."""
        if slot.req.kind == "real":
            return min(self.check_every, slot.req.steps - slot.steps)
        return self.check_every

    def _promote(self, slot):
        """Move a complex64 imaginary-time member to a complex128 field."""
        fresh = self._new_solve(slot.req, "c128")
        fresh.copy_from(slot.solve)
        self._drop_batches(slot.solve)
        slot.solve.close()
        slot.solve, slot.precision = fresh, "c128"

    def _complete(self, slot, now):
        req = slot.req
        diag = slot.solve.diagnostics(req.steps * req.dt if req.kind == "real" else 0.0, spectrum=False)
        self._drop_batches(slot.solve)
        slot.solve.close()
        converged = req.kind == "real" or (slot.precision == "c128" and slot.dmu < req.tol)
        return Outcome(id=req.id, n=req.n, kind=req.kind, precision=req.precision, latency=now - req.arrival,
                       queued=slot.admitted - req.arrival, steps=slot.steps, mu=diag["mu"], energy=diag["E"],
                       norm=diag["norm"], dmu=slot.dmu, converged=converged)


def request_stream(count, rate_per_s=2.0, seed=0, sides=(64, 96, 128, 192, 256),
                   side_p=(0.3, 0.25, 0.2, 0.15, 0.1), imag_fraction=0.7, c64_fraction=0.8,
                   g_range=(50.0, 2000.0), length=20.0, dt_osc=1e-3, tol_osc=1e-9, real_steps=2000):
    """Seeded Poisson stream of mixed requests (oscillator-unit parameters
    mapped to reference units: beta = 2g, dtau = dtau'/2, tol doubled)."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    for i in range(count):
        t += rng.exponential(1.0 / rate_per_s)
        n = int(rng.choice(sides, p=side_p))
        kind = "imaginary" if rng.random() < imag_fraction else "real"
        precision = "c64" if rng.random() < c64_fraction else "c128"
        g = rng.uniform(*g_range)
        out.append(Request(id=i, arrival=t, n=n, kind=kind, precision=precision, beta=2 * g, length=length,
                           static=harmonic_expression(1, 1, 1), dt=dt_osc / 2, sigma=1.0, tol=2 * tol_osc,
                           steps=real_steps))
    return out
