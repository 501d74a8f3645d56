"""Pass tile variants and the autotuner on the device.

Every compiled variant of every fast side computes the same step: each one,
set for all pass kinds (diagnostic passes included), is checked against the
FP64 oracle to the same bars as the default shapes (complex128: relative L2
<= 1e-11 per step, mu/E to 1e-11; complex64: relative L2 <= 2e-5; nk counts
bit-exact).
"""

import numpy as np
import pytest

from oracle import gpe_oracle as orc

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]

BEAM = "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)"
SIDES = [32, 48, 64, 96, 128, 192, 256]


@pytest.fixture(autouse=True, scope="module")
def threaded_oracle():
    orc.use_threads(8)
    yield
    orc.use_threads(None)


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def variant_count(n, precision):
    from gpe_simulator_amd import _lib
    return len(_lib.pass_variants(n, _lib.C64 if precision == "c64" else _lib.C128))


def plans_for(n, precision):
    """Every step plan of (side, precision), the default plan last."""
    from gpe_simulator_amd import _lib
    prec = _lib.C64 if precision == "c64" else _lib.C128
    plans = [_lib.PLAN_FOUR_SWEEP, _lib.PLAN_SEPARABLE]
    if n in _lib.PLANE_SIDES[prec]:
        plans.append(_lib.PLAN_PLANE)
    default = _lib.default_plan(prec, _lib.REAL, n=n)
    return [p for p in plans if p != default] + [default]


@pytest.mark.parametrize("precision", ["c128", "c64"])
@pytest.mark.parametrize("n", SIDES)
def test_every_variant_matches_the_oracle(n, precision, require_gpu):
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    length, dt, beta, steps = 24.0, 5e-4, 400.0, 2 if n < 256 else 1
    real = orc.Problem(n, length, "real", dt, beta, static="x**2 + y**2 + 2*z**2", dynamic=BEAM)
    imag = orc.Problem(n, length, "imaginary", dt, beta, static="x**2 + y**2 + 2*z**2")
    psi0 = orc.synthetic_state(n, length, seed=n, beta=beta)
    exp_real = real.run(psi0, steps)
    exp_imag = imag.run(psi0, steps)
    t = 0.37
    e_ref = real.energies(psi0, t)
    spectrum = orc.nk_extraction(np.fft.fftn(psi0), real.x, real.k)
    dx = real.x[1] - real.x[0]
    tol = 1e-11 if precision == "c128" else 2e-5
    for v in range(variant_count(n, precision)):
        print(f"side {n} {precision} variant {v}", flush=True)
        s = make_solve(n, real.x, real.k, "real", dt, beta, static=real.static_source, dynamic=BEAM,
                       precision=precision)
        s.set_variants([v] * len(_lib.PASS_KINDS))
        for plan in plans_for(n, precision):     # every step plan, default last
            s.set_plan(plan)
            s.upload(psi0)
            s.step(steps, midpoint_times(0, steps, dt))
            assert rel_l2(s.to_numpy(), exp_real) < tol, f"variant {v} plan {plan}"
        # diagnostic passes on the same variant, on the initial state
        s.upload(psi0)
        d = s.diagnostics(t)
        for key in ("norm", "E_kin", "E_pot", "E_int", "mu"):
            assert d[key] == pytest.approx(e_ref[key], rel=1e-11 if precision == "c128" else 2e-6), key
        np.testing.assert_array_equal(d["counts"], spectrum["counts"])
        np.testing.assert_allclose(d["occupation"] * dx ** 3 / n ** 3, spectrum["Nk"],
                                   rtol=1e-10 if precision == "c128" else 2e-5, atol=1e-14)
        s.close()
        s = make_solve(n, imag.x, imag.k, "imaginary", dt, beta, static=imag.static_source, precision=precision)
        s.set_variants([v] * len(_lib.PASS_KINDS))
        for plan in plans_for(n, precision):
            s.set_plan(plan)
            s.upload(psi0)
            s.step(steps)
            assert rel_l2(s.to_numpy(), exp_imag) < tol, f"variant {v} plan {plan}"
        s.close()


def test_autotune_sets_context_defaults(require_gpu):
    from gpe_simulator_amd import _lib, tune
    from gpe_simulator_amd.engine import make_solve
    ctx = _lib.Context(0)
    result = tune.autotune(96, _lib.C64, _lib.IMAG, ctx=ctx)
    print(result.report(), flush=True)
    count = variant_count(96, "c64")
    assert all(0 <= v < count for v in result.choice)
    assert set(result.timings) == set(range(count))
    assert all(0 <= v < count for v in result.event_choice)
    # the whole-step stage starts from the per-pass choice and keeps the faster
    # step (5% covers the noise between the two separate timings)
    assert 0 < result.step_us <= result.event_step_us * 1.05
    ctx.set_variants(96, _lib.C64, _lib.IMAG, result.choice)
    x, k = orc.get_grids(96, 20.0)
    with_ctx = make_solve(96, x, k, "imaginary", 5e-4, 100.0, static="x**2 + y**2 + z**2", precision="c64",
                          ctx=ctx)
    assert with_ctx.variants() == result.choice
    with_ctx.close()
    ctx.close()


def test_cutoff_plane_solves_tune_x_kin_alone(require_gpu):
    """A solve with the momentum cutoff on the plane plan tunes x_kin, in
    scratch solves with a cutoff on the plane plan, and no other kind: the
    plane passes have one shape per side."""
    from gpe_simulator_amd import _lib, tune
    from gpe_simulator_amd.engine import make_solve
    ctx = _lib.Context(0)
    x, k = orc.get_grids(64, 16.0)
    s = make_solve(64, x, k, "real", 1e-3, 100.0, static="x**2 + y**2 + z**2", kcut=0.6 * np.abs(k).max(),
                   precision="c64", ctx=ctx)
    assert s.plan == _lib.PLAN_PLANE
    result = ctx.tunings[(64, _lib.C64, _lib.REAL, 1, tune.PLAN_PLANE_CUT)]
    print(result.report(), flush=True)
    assert result.plan == tune.PLAN_PLANE_CUT
    x_kin = _lib.PASS_KINDS.index("x_kin")
    assert all(v == 0 for i, v in enumerate(result.choice) if i != x_kin)
    assert set(result.timings[0]) == {"x_kin"}
    assert s.variants()[x_kin] == result.choice[x_kin]
    s.close()
    ctx.close()


def test_batch_members_must_share_variants(require_gpu):
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(64, 16.0)
    a = make_solve(64, x, k, "imaginary", 5e-4, 100.0, static="x**2 + y**2 + z**2", precision="c64")
    b = make_solve(64, x, k, "imaginary", 5e-4, 100.0, static="x**2 + y**2 + z**2", precision="c64")
    a.set_variants([0] * len(_lib.PASS_KINDS))
    b.set_variants([1] * len(_lib.PASS_KINDS))
    batch = _lib.Batch([a, b])
    with pytest.raises(AssertionError):
        batch.step(2)
    b.set_variants([0] * len(_lib.PASS_KINDS))
    batch.step(2)
    batch.close()


def test_variant_out_of_range_is_rejected(require_gpu):
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(64, 16.0)
    s = make_solve(64, x, k, "real", 5e-4, 100.0, precision="c64")
    with pytest.raises(AssertionError):
        s.set_variants([99] * len(_lib.PASS_KINDS))
    with pytest.raises(AssertionError):
        s.set_variants([0] * 3)


@pytest.mark.parametrize("precision,tol", [("c128", 1e-12), ("c64", 2e-5)])
@pytest.mark.parametrize("n", [96, 256])
def test_reference_structured_step_under_every_variant(n, precision, tol, require_gpu):
    """The plain pass kinds of gpe_step_unfused (half-step alone, inverse FFT
    along z, FFT and inverse FFT along x) under every compiled variant of the
    side, including the two-round exchange, against the fused step."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    length, dt, steps = 16.0, 5e-4, 3
    x, k = orc.get_grids(n, length)
    psi0 = orc.synthetic_state(n, length, seed=2, beta=100.0)
    tm = midpoint_times(0, steps, dt)
    fused = make_solve(n, x, k, "real", dt, 100.0, static="x**2 + y**2 + z**2", dynamic=BEAM,
                       precision=precision)
    fused.set_variants([0] * len(_lib.PASS_KINDS))
    fused.upload(psi0)
    fused.step(steps, tm)
    expected = fused.to_numpy()
    fused.close()
    for v in range(variant_count(n, precision)):
        solve = make_solve(n, x, k, "real", dt, 100.0, static="x**2 + y**2 + z**2", dynamic=BEAM,
                           precision=precision)
        solve.set_variants([v] * len(_lib.PASS_KINDS))
        solve.upload(psi0)
        solve.step_unfused(steps, tm)
        assert rel_l2(solve.to_numpy(), expected) < tol, (n, precision, v)
        solve.close()
