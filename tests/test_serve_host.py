"""The solve server's scheduling on a simulated device (no GPU).

A fake batch "runs" a tick for steps x (seconds per step of its side) of
wall time and advances its members' step counts; a fake solve reports the
norms of an imaginary-time relaxation whose mu converges exponentially.  The
checks: a real-time run ends on its requested step count whatever the tick;
each shape's group moves at its own pace, so small requests finish while a
large one is still running; at most one group of a shape is in flight; every
request completes once.
"""

import math
import time

import numpy as np
import pytest

from gpe_simulator_amd import serve
from gpe_simulator_amd.serve import Request, Server

SECONDS_PER_STEP = {64: 2e-5, 96: 4e-5, 256: 4e-4}


class FakeSolve:
    def __init__(self, n, kind, dt, precision):
        self.n, self.kind, self.dt, self.precision = n, kind, dt, precision
        self.steps = 0
        self.closed = False

    def gaussian(self, sigma):
        pass

    def mu(self, k):
        return 10.0 + 5.0 * math.exp(-k / 30.0)

    def copy_from(self, other):
        self.steps = other.steps

    def diagnostics(self, t, spectrum=False):
        return {"mu": self.mu(self.steps), "E": 7.5, "norm": 1.0}

    def close(self):
        self.closed = True


class FakeBatch:
    in_flight = {}          # shape -> groups running now
    most_in_flight = {}     # shape -> most groups of the shape running at once
    launches = []

    def __init__(self, solves, ctx=None):
        self.solves = list(solves)
        self.shape = (self.solves[0].n, self.solves[0].precision, self.solves[0].kind)
        self.end = None

    def launch(self, steps, graph=True):
        # launches queue on the batch's stream: a second one starts when the first ends
        for s in self.solves:
            assert not s.closed
            s.steps += steps
        queued = self.end is not None
        start = self.end if queued else time.perf_counter()
        self.end = start + steps * SECONDS_PER_STEP[self.shape[0]]
        if queued:
            FakeBatch.launches.append((self.shape, steps, len(self.solves)))
            return
        self.steps = steps
        running = FakeBatch.in_flight.get(self.shape, 0) + 1
        FakeBatch.in_flight[self.shape] = running
        FakeBatch.most_in_flight[self.shape] = max(FakeBatch.most_in_flight.get(self.shape, 0), running)
        FakeBatch.launches.append((self.shape, steps, len(self.solves)))

    def done(self):
        if self.end is not None and time.perf_counter() >= self.end:
            self.end = None
            FakeBatch.in_flight[self.shape] -= 1
        return self.end is None

    def norm_history(self, count):
        rows = []
        for s in self.solves:
            ks = np.arange(s.steps - count + 1, s.steps + 1)
            rows.append(np.exp(-2.0 * s.dt * np.array([s.mu(k) for k in ks])))
        return np.array(rows)

    def close(self):
        pass


class FakeContext:
    def mem_info(self):
        return 1 << 40, 1 << 40


@pytest.fixture
def fake_device(monkeypatch):
    FakeBatch.in_flight, FakeBatch.most_in_flight, FakeBatch.launches = {}, {}, []
    made = []

    def make_solve(n, x, k, kind, dt, beta, static=None, precision="c64", ctx=None):
        s = FakeSolve(n, kind, dt, precision)
        made.append(s)
        return s

    monkeypatch.setattr(serve, "make_solve", make_solve)
    monkeypatch.setattr(serve._lib, "Batch", FakeBatch)
    monkeypatch.setattr(serve, "_footprint", lambda req, precision: 1)
    return made


def imag(i, n, arrival=0.0):
    return Request(id=i, arrival=arrival, n=n, kind="imaginary", precision="c128", beta=100.0, dt=5e-4,
                   tol=1e-6)


def test_large_real_run_ends_on_its_step_count_in_pieces(fake_device):
    """A 256^3 complex64 real-time run of 133 steps: ticks of 50, 50 and 33
    steps, each sent as launches within the budget (alone, each tick's pieces
    are queued at once)."""
    req = Request(id=0, arrival=0.0, n=256, kind="real", precision="c64", beta=100.0, steps=133)
    outcomes, _ = Server(ctx=FakeContext(), check_every=50).run([req])
    assert outcomes[0].steps == 133 and fake_device[0].steps == 133
    pieces = [steps for _, steps, _ in FakeBatch.launches]
    assert sum(pieces) == 133 and max(pieces) <= 16


def test_real_run_ends_on_its_step_count(fake_device):
    req = Request(id=0, arrival=0.0, n=64, kind="real", precision="c64", beta=100.0, steps=120)
    outcomes, _ = Server(ctx=FakeContext(), check_every=50).run([req])
    assert len(outcomes) == 1 and outcomes[0].steps == 120
    assert [steps for _, steps, _ in FakeBatch.launches] == [50, 50, 20]
    assert fake_device[0].steps == 120 and fake_device[0].closed


def test_small_requests_do_not_wait_for_a_large_one(fake_device):
    reqs = [imag(0, 256)] + [imag(i, 64) for i in range(1, 4)] + [imag(4, 96, arrival=0.005)]
    outcomes, stats = Server(ctx=FakeContext(), check_every=50).run(reqs)
    assert sorted(o.id for o in outcomes) == list(range(5))
    lat = {o.id: o.latency for o in outcomes}
    big_tick = 50 * SECONDS_PER_STEP[256]
    # the small ones converge in a few of their own ticks, well inside the
    # large request's run, instead of one large tick per small tick
    steps = {o.id: o.steps for o in outcomes}
    assert all(lat[i] < lat[0] for i in (1, 2, 3, 4))
    assert all(lat[i] < 0.5 * lat[0] for i in (1, 2, 3))
    assert lat[0] >= (steps[0] // 50) * big_tick * 0.9
    # equal-shape members share one group; never two groups of a shape at once
    assert max(FakeBatch.most_in_flight.values()) == 1
    # the large shape's 50-step ticks go to the device in equal pieces that fit
    # the launch budget (5 steps of a 256^3 complex128 solve), the small
    # shapes' in one launch
    assert {steps for shape, steps, _ in FakeBatch.launches if shape[0] == 256} == {5}
    assert {steps for shape, steps, _ in FakeBatch.launches if shape[0] == 64} == {50}
    assert steps[0] % 50 == 0
    assert any(shape[0] == 64 and members == 3 for shape, _, members in FakeBatch.launches)
    assert all(o.converged for o in outcomes)
    assert all(s.closed for s in fake_device)
    assert stats["requests"] == 5


def test_late_member_joins_its_shapes_next_tick(fake_device):
    reqs = [imag(0, 96), imag(1, 96, arrival=0.004)]
    outcomes, _ = Server(ctx=FakeContext(), check_every=50).run(reqs)
    assert sorted(o.id for o in outcomes) == [0, 1]
    assert FakeBatch.most_in_flight[(96, "c128", "imaginary")] == 1
    # once the second request is admitted, both advance in one group
    assert any(members == 2 for _, _, members in FakeBatch.launches)
