"""Batched solves, hipGraph replay, the parameter sweep and the solve server.

A batch member must produce the same field as the same solve stepped alone
(bit-identical: same kernels, same per-member reductions), over several graph
replays with eager work and synchronisation between them.
"""

import numpy as np
import pytest

from oracle import gpe_oracle as orc

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


def members(kind, precision, n=64, count=3):
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(n, 16.0)
    out = []
    for i in range(count):
        s = make_solve(n, x, k, kind, 5e-4 * (1 + 0.1 * i), 100.0 * (i + 1),
                       static=f"x**2 + y**2 + {1 + i}*z**2", precision=precision)
        out.append(s)
    return out


@pytest.mark.parametrize("kind", ["imaginary", "real"])
@pytest.mark.parametrize("precision", ["c64", "c128"])
@pytest.mark.parametrize("graph", [False, True])
def test_batch_members_match_solo_runs(kind, precision, graph, require_gpu):
    from gpe_simulator_amd import _lib
    batched, alone = members(kind, precision), members(kind, precision)
    for i, (a, b) in enumerate(zip(batched, alone)):
        psi = orc.synthetic_state(64, 16.0, seed=i, beta=100.0 * (i + 1))
        a.upload(psi)
        b.upload(psi)
    batch = _lib.Batch(batched)
    for replay in range(3):
        batch.step(20, graph=graph)
        for s in alone:
            s.step(20)
        # eager work between replays: the fields are read back and compared
        for row, (a, b) in enumerate(zip(batched, alone)):
            np.testing.assert_array_equal(a.to_numpy(), b.to_numpy())
            if kind == "imaginary":
                np.testing.assert_array_equal(a.norm_history(20), b.norm_history(20))
                np.testing.assert_array_equal(batch.norm_history(20)[row], b.norm_history(20))
    batch.close()


def test_fresh_member_history_while_another_batch_runs(require_gpu):
    """Buffers a solve allocates on its first batched step (norm history,
    partials) are cleared on the solve's own stream, so a long batch on
    another stream cannot delay the clear past the new member's writes."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(128, 20.0)
    busy = [make_solve(128, x, k, "real", 1e-4, 100.0, static="x**2 + y**2 + z**2", precision="c64")
            for _ in range(4)]
    for s in busy:
        s.gaussian(1.0)
    long_batch = _lib.Batch(busy)
    x, k = orc.get_grids(64, 16.0)
    for trial in range(3):
        long_batch.launch(400, graph=True)
        fresh = members("imaginary", "c64", count=2)
        alone = members("imaginary", "c64", count=2)
        for a, b in zip(fresh, alone):
            a.gaussian(1.0)
            b.gaussian(1.0)
        batch = _lib.Batch(fresh)
        batch.launch(50, graph=True)
        batch.wait()
        long_batch.wait()
        for a, b in zip(fresh, alone):
            b.step(50)
            np.testing.assert_array_equal(a.norm_history(50), b.norm_history(50))
        batch.close()
    long_batch.close()


def test_batch_rejects_mixed_shapes(require_gpu):
    from gpe_simulator_amd import _lib
    a = members("imaginary", "c64", n=64, count=1)[0]
    b = members("imaginary", "c64", n=96, count=1)[0]
    with pytest.raises(AssertionError):
        _lib.Batch([a, b])


def test_sweep_batched_matches_sequential(require_gpu):
    from gpe_simulator_amd.sweep import config_sweep, solve_sequential, sweep_batched
    reqs = config_sweep(count=4, seed=3, n=64, length=20.0, tol_osc=1e-9)
    batched = sweep_batched(reqs, check_every=50)
    for req, res in zip(reqs, batched):
        assert res.converged
        seq = solve_sequential(req, check_every=50)
        assert seq.converged
        assert res.mu == pytest.approx(seq.mu, rel=1e-8)
        assert res.energy == pytest.approx(seq.energy, rel=1e-8)
        # the norm-decay estimate agrees with the FP64 functional to O(dtau)
        assert res.mu_estimate == pytest.approx(res.mu, rel=1e-2)


def test_server_results_same_alone_and_in_mix(require_gpu):
    from gpe_simulator_amd.serve import Server, request_stream
    reqs = request_stream(8, rate_per_s=50.0, seed=7, sides=(64, 96), side_p=(0.5, 0.5),
                          tol_osc=5e-9, real_steps=200)
    mixed, stats = Server(check_every=50).run(reqs)
    assert stats["requests"] == len(reqs)
    assert stats["max_active"] >= 2
    by_id = {o.id: o for o in mixed}
    for req in reqs:
        alone, _ = Server(check_every=50).run([req])
        a, m = alone[0], by_id[req.id]
        assert m.converged and a.converged
        assert m.steps == a.steps
        assert m.mu == pytest.approx(a.mu, rel=1e-12)
        assert m.energy == pytest.approx(a.energy, rel=1e-12)


@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_stirrers_of_different_speeds_match_solo_runs(precision, require_gpu):
    """Members with their own time-dependent programs (beams at different
    angular speeds, one member without a beam) batched: each is
    bit-identical to the same run alone, for chunked multi-step calls."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    from gpe_simulator_amd.units import stirrer_expression
    n, length, dt, steps = 64, 16.0, 5e-4, 300
    x, k = orc.get_grids(n, length)
    beams = [stirrer_expression(5.0, 3.0, w, 2.0) for w in (0.2, 0.6, 1.1)] + [None]
    psi0 = orc.synthetic_state(n, length, seed=4, beta=400.0)

    def member(beam, i):
        s = make_solve(n, x, k, "real", dt * (1 + 0.1 * i), 400.0, static="x**2 + y**2 + z**2", dynamic=beam,
                       precision=precision)
        s.upload(psi0)
        return s

    batched = [member(b, i) for i, b in enumerate(beams)]
    alone = [member(b, i) for i, b in enumerate(beams)]
    times = np.stack([midpoint_times(0, 2 * steps, dt * (1 + 0.1 * i)) for i in range(len(beams))])
    batch = _lib.Batch(batched)
    with pytest.raises(AssertionError):
        batch.step(10)                       # time-dependent members need their times
    batch.step(steps, t_mid=times[:, :steps])
    batch.step(steps, t_mid=np.ascontiguousarray(times[:, steps:]))
    for i, (a, b) in enumerate(zip(batched, alone)):
        b.step(steps, times[i, :steps])
        b.step(steps, times[i, steps:])
        np.testing.assert_array_equal(a.to_numpy(), b.to_numpy())
    lz = [s.angular_momentum() for s in batched]
    assert len({round(v, 12) for v in lz}) == len(lz)      # the speeds imprint different Lz
    batch.close()


def test_stirring_sweep_matches_sequential(require_gpu):
    from gpe_simulator_amd.sweep import stir_batched, stir_sequential, stirrer_runs
    runs = stirrer_runs([0.3, 0.7, 1.2], n=64, length=16.0, g=200.0, steps=200)
    psi0 = orc.synthetic_state(64, 16.0, seed=6, beta=400.0)
    batched = stir_batched(runs, psi0, precision="c64")
    for run, got in zip(runs, batched):
        alone = stir_sequential(run, psi0, precision="c64")
        assert got == alone


def test_sweep_over_mixed_sides(require_gpu):
    """Requests of two grid sides relax in separate batches of one sweep,
    each to the same result as alone."""
    from gpe_simulator_amd.sweep import config_sweep, solve_sequential, sweep_batched
    reqs = config_sweep(count=2, seed=5, n=64, length=20.0) + config_sweep(count=2, seed=6, n=96, length=20.0)
    for req, res in zip(reqs, sweep_batched(reqs)):
        assert res.converged
        assert res.mu == pytest.approx(solve_sequential(req).mu, rel=1e-8)


def test_server_warm_up_tunes_every_shape(require_gpu):
    """Server.warm tunes each (side, precision, kind) once on its context;
    later solves of those shapes start from the tuned variants."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.serve import Server
    ctx = _lib.Context(0)
    server = Server(ctx=ctx)
    shapes = [(64, "c64", "imaginary"), (64, "c128", "imaginary"), (96, "c64", "real")]
    seconds = server.warm(shapes)
    assert seconds > 0
    for n, prec, kind in shapes:
        key = (n, _lib.C64 if prec == "c64" else _lib.C128, _lib.IMAG if kind == "imaginary" else _lib.REAL, 1)
        assert ctx.tunings[key] is not None
        assert ctx.default_variants(*key[:3]) == ctx.tunings[key].choice
    again = server.warm(shapes)
    assert again < seconds
    ctx.close()


def test_server_real_run_stops_on_its_step_count(require_gpu):
    """A real-time request whose step count is not a multiple of the tick
    (120 steps, ticks of 50) takes a short last tick: it reports 120 steps
    and the diagnostics of the state after exactly 120 steps."""
    from gpe_simulator_amd.engine import make_solve
    from gpe_simulator_amd.grids import get_grids
    from gpe_simulator_amd.serve import Request, Server
    req = Request(id=0, arrival=0.0, n=64, kind="real", precision="c128", beta=300.0, dt=5e-4, steps=120)
    outcomes, _ = Server(check_every=50).run([req])
    assert outcomes[0].steps == 120
    x, k = get_grids(req.n, req.length)
    solve = make_solve(req.n, x, k, "real", req.dt, req.beta, static=req.static, precision="c128")
    solve.gaussian(req.sigma)
    solve.step(120)
    diag = solve.diagnostics(120 * req.dt, spectrum=False)
    solve.close()
    assert outcomes[0].mu == pytest.approx(diag["mu"], rel=1e-12)
    assert outcomes[0].energy == pytest.approx(diag["E"], rel=1e-12)
    with pytest.raises(ValueError):
        Server(check_every=50).run([Request(id=1, arrival=0.0, n=64, kind="real", precision="c64",
                                            beta=300.0, steps=0)])


def test_config_sweep_64_solves_at_96(require_gpu):
    """The batched parameter sweep at its full size: 64 imaginary-time
    ground states at 96^3 on [-10,10)^3 with seed-drawn g, omega_z and start
    widths (seed 0), relaxed to |dmu| < 2e-9 (reference units): every member
    converges, and a fixed sample of members matches the same request solved
    alone in complex128 to 1e-8 in mu and E."""
    from gpe_simulator_amd.sweep import config_sweep, solve_sequential, sweep_batched
    reqs = config_sweep(count=64, seed=0)
    assert len(reqs) == 64 and {r.n for r in reqs} == {96}
    results = sweep_batched(reqs)
    assert all(r.converged for r in results)
    for i in (0, 21, 42, 63):
        seq = solve_sequential(reqs[i])
        assert seq.converged
        assert results[i].mu == pytest.approx(seq.mu, rel=1e-8)
        assert results[i].energy == pytest.approx(seq.energy, rel=1e-8)


def test_serving_stream_every_side_kind_and_precision(require_gpu):
    """The serving stream at its full request mix: a seeded Poisson stream of
    12 requests holding every side of {64, 96, 128, 192, 256}, ground states
    and 2000-step real-time runs, complex64 and complex128.  Requests overlap
    on the device, and every result is the same as the request served alone
    (same step count; mu and E to 1e-12)."""
    from gpe_simulator_amd.serve import Server, request_stream
    reqs = request_stream(12, rate_per_s=8.0, seed=2)
    assert {r.n for r in reqs} == {64, 96, 128, 192, 256}
    assert {(r.kind, r.precision) for r in reqs} == {(k, p) for k in ("imaginary", "real") for p in ("c64", "c128")}
    assert {r.steps for r in reqs if r.kind == "real"} == {2000}
    mixed, stats = Server(check_every=50).run(reqs)
    assert stats["requests"] == len(reqs)
    assert stats["max_active"] >= 3
    assert stats["p50_ms"] <= stats["p99_ms"]
    by_id = {o.id: o for o in mixed}
    for req in reqs:
        alone, _ = Server(check_every=50).run([req])
        a, m = alone[0], by_id[req.id]
        assert m.converged and a.converged, req
        assert m.steps == a.steps
        if req.kind == "real":
            assert m.steps == req.steps
        assert m.mu == pytest.approx(a.mu, rel=1e-12)
        assert m.energy == pytest.approx(a.energy, rel=1e-12)
