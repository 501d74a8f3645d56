"""GPU parity of libgpe_hip against the FP64 oracle on identical seeded inputs.

Tolerances: complex128 runs match the oracle to relative L2 <= 1e-11 per step
(rounding of separable kinetic factors and twiddles); complex64 runs to
relative L2 <= 2e-5 after a few steps.  Shell counts are bit-exact.
"""

import numpy as np
import pytest

from oracle import gpe_oracle as orc

pytestmark = pytest.mark.gpu


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def random_state(n, length, seed=42):
    x, _ = orc.get_grids(n, length)
    rng = np.random.default_rng(seed)
    psi = rng.normal(size=(n,) * 3) + 1j * rng.normal(size=(n,) * 3)
    return orc.normalise(psi, x)


def make(n, length, kind, dt, beta, static=None, dynamic=None, kcut=None, precision="c128"):
    from gpe_simulator_amd.engine import make_solve
    problem = orc.Problem(n, length, kind, dt, beta, static=static, dynamic=dynamic, kcut=kcut)
    solve = make_solve(n, problem.x, problem.k, kind, dt, beta, static=static, dynamic=dynamic, kcut=kcut,
                       precision=precision)
    return problem, solve


def run_both(problem, solve, psi0, steps, first=0):
    from gpe_simulator_amd.engine import midpoint_times
    solve.upload(psi0)
    tm = midpoint_times(first, steps, problem.dt) if problem.dynamic_source else None
    solve.step(steps, tm)
    expected = problem.run(psi0, steps, first=first)
    return solve.to_numpy(), expected


CASES = [
    # (kind, static, dynamic, kcut)
    ("real", "x**2 + y**2 + z**2", None, None),
    ("real", "x/2 + y**2/10 + abs(z)/5", "sin(t)*x + cos(pi)*z", None),
    ("real", "0.002*(x*x + y*y + z*z)", "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)", 2.5),
    ("real", "sqrt(x**2 + y**2 + 1) * z / 10", "cos(t)*x*z/20", None),
    ("imaginary", "x**2 + y**2 + 16*z**2", None, None),
]


@pytest.mark.parametrize("n", [8, 18, 64])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_one_step_c128(n, case, require_gpu):
    kind, static, dynamic, kcut = CASES[case]
    length = 8.0
    problem, solve = make(n, length, kind, 1e-3, 50.0, static, dynamic, kcut)
    psi0 = random_state(n, length)
    got, expected = run_both(problem, solve, psi0, 1, first=3)
    assert rel_l2(got, expected) < 1e-11


@pytest.mark.parametrize("n", [64, 96])
@pytest.mark.parametrize("case", range(len(CASES)))
def test_fused_steps_c128(n, case, require_gpu):
    kind, static, dynamic, kcut = CASES[case]
    length = 12.0
    problem, solve = make(n, length, kind, 2e-3, 30.0, static, dynamic, kcut)
    psi0 = orc.synthetic_state(n, length, seed=1, beta=30.0)
    got, expected = run_both(problem, solve, psi0, 25)
    assert rel_l2(got, expected) < 1e-10


@pytest.mark.parametrize("n", [128, 192, 256])
def test_fast_sides_one_step_c128(n, require_gpu):
    length = 20.0
    problem, solve = make(n, length, "real", 5e-4, 100.0, "x**2 + y**2 + z**2",
                          "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)")
    psi0 = orc.synthetic_state(n, length, seed=2, beta=100.0)
    got, expected = run_both(problem, solve, psi0, 2)
    assert rel_l2(got, expected) < 1e-11


@pytest.mark.parametrize("n", [18, 64, 96, 256])
def test_c64_steps(n, require_gpu):
    length = 16.0
    problem, solve = make(n, length, "real", 1e-3, 200.0, "x**2 + y**2 + z**2",
                          "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)", precision="c64")
    psi0 = orc.synthetic_state(n, length, seed=3, beta=200.0)
    got, expected = run_both(problem, solve, psi0, 4)
    assert rel_l2(got, expected) < 2e-5


@pytest.mark.parametrize("n", [8, 18, 64, 96])
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_diagnostics(n, precision, require_gpu):
    length = 10.0
    problem, solve = make(n, length, "real", 1e-3, 80.0, "x**2 + 0.5*y**2 + z**2",
                          "cos(t)*y/10", precision=precision)
    psi0 = random_state(n, length, seed=7)
    solve.upload(psi0)
    t = 0.37
    got = solve.diagnostics(t)
    exp = problem.energies(psi0, t)
    tol = 1e-11 if precision == "c128" else 2e-6
    for key in ("norm", "E_kin", "E_pot", "E_int", "E", "mu", "parseval"):
        assert got[key] == pytest.approx(exp[key], rel=tol), key
    spectrum = orc.nk_extraction(np.fft.fftn(psi0), problem.x, problem.k)
    np.testing.assert_array_equal(got["counts"], spectrum["counts"])
    dx = problem.x[1] - problem.x[0]
    np.testing.assert_allclose(got["occupation"] * dx ** 3 / n ** 3, spectrum["Nk"], rtol=tol * 10,
                               atol=tol * 1e-3)
    view = orc.spectrum_view(np.fft.fftn(psi0), problem.x, problem.k)
    np.testing.assert_array_equal(got["view_counts"], view["counts"])
    np.testing.assert_allclose(got["view_occupation"] * dx ** 3 / n ** 3, view["Nk"], rtol=tol * 10,
                               atol=tol * 1e-3)


@pytest.mark.parametrize("n", [7, 128, 192, 256])
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_spectrum_view_and_shells_on_large_and_odd_sides(n, precision, require_gpu):
    """Shell and equal-width view counts bit-exact, sums to FP64 (complex128)
    or float (complex64) accuracy, on the fast sides and an odd one."""
    length = 16.0
    problem, solve = make(n, length, "real", 1e-3, 10.0, precision=precision)
    psi0 = random_state(n, length, seed=11)
    solve.upload(psi0)
    got = solve.diagnostics(0.0)
    solve.close()
    psi_k = np.fft.fftn(psi0)
    dx = problem.x[1] - problem.x[0]
    tol = 1e-11 if precision == "c128" else 2e-6
    shells = orc.nk_extraction(psi_k, problem.x, problem.k)
    view = orc.spectrum_view(psi_k, problem.x, problem.k)
    np.testing.assert_array_equal(got["counts"], shells["counts"])
    np.testing.assert_array_equal(got["view_counts"], view["counts"])
    assert got["view_counts"].sum() == n ** 3
    for occ, ref in ((got["occupation"], shells["Nk"]), (got["view_occupation"], view["Nk"])):
        np.testing.assert_allclose(occ * dx ** 3 / n ** 3, ref, rtol=tol * 10, atol=tol * 1e-3 * ref.max())


@pytest.mark.parametrize("n", [18, 64])
def test_imaginary_normalisation_and_history(n, require_gpu):
    length = 12.0
    problem, solve = make(n, length, "imaginary", 1e-3, 100.0, "x**2 + y**2 + z**2")
    psi0 = orc.gaussian(problem.x, 1.0)
    got, expected = run_both(problem, solve, psi0, 30)
    assert rel_l2(got, expected) < 1e-10
    assert solve.norm() == pytest.approx(1.0, abs=1e-12)
    history = solve.norm_history(30)
    assert np.all(history < 1.0) and np.all(history > 0.9)


def test_gaussian_initial_state(require_gpu):
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(64, 20.0)
    solve = make_solve(64, x, k, "imaginary", 1e-3, 0.0)
    solve.gaussian(1.3)
    np.testing.assert_allclose(solve.to_numpy(), orc.gaussian(x, 1.3), rtol=1e-13, atol=1e-15)
    assert solve.norm() == pytest.approx(1.0, abs=1e-13)


@pytest.mark.parametrize("n", [18, 64])
@pytest.mark.parametrize("precision,static,tol", [
    # half-step phases up to a few radians: FP64 Cody-Waite reduction by pi/2
    ("c128", "40*(x*x + y*y + z*z)", 1e-11),
    ("c64", "40*(x*x + y*y + z*z)", 2e-5),
    # phases of ~1e4 rad (complex128; a float phase that large carries ~1e-3
    # rad of rounding by itself)
    ("c128", "2e7 + 40*(x*x + y*y)", 1e-10),
    # phases of ~1e10 rad: q = rint(x * 2/pi) far beyond the int range; the
    # phase value itself carries ~2e-6 rad of FP64 rounding at this size
    ("c128", "2e13 + 40*(x*x + y*y)", 1e-5),
])
def test_large_phases(n, precision, static, tol, require_gpu):
    length = 12.0
    problem, solve = make(n, length, "real", 1e-3, 50.0, static, precision=precision)
    psi0 = random_state(n, length, seed=3)
    got, expected = run_both(problem, solve, psi0, 3)
    assert rel_l2(got, expected) < tol


def test_c64_phase_of_a_million_radians(require_gpu):
    """complex64 phases of ~1e6 rad take the FP64 reduction: the field stays
    finite and the rotation keeps the norm.  (The float phase itself carries
    ~0.06 rad of rounding, so the values are not compared.)"""
    problem, solve = make(18, 12.0, "real", 1e-3, 0.0, "2e9 + 40*(x*x + y*y)", precision="c64")
    solve.upload(random_state(18, 12.0, seed=4))
    n0 = solve.norm()
    solve.step(3)
    assert np.isfinite(solve.to_numpy()).all()
    assert solve.norm() == pytest.approx(n0, rel=1e-5)


@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_phase_beyond_limit_gives_nan(precision, require_gpu):
    """Half-step phases of 2^40 rad (1.1e12) and more have no meaningful
    value in either precision: the field turns NaN instead of carrying a
    silent number (here 2.5e12 rad)."""
    problem, solve = make(18, 12.0, "real", 1e-3, 0.0, "5e15", precision=precision)
    solve.upload(random_state(18, 12.0, seed=5))
    solve.step(1)
    assert np.isnan(solve.to_numpy()).all()


@pytest.mark.parametrize("n", [64, 96])
@pytest.mark.parametrize("precision,tol", [("c128", 1e-10), ("c64", 3e-5)])
def test_complex_static_volume_on_fast_sides(n, precision, tol, require_gpu):
    """V_static = box - i*absorber as uploaded volumes (GPE_TGT_VOL and
    GPE_TGT_VOL_IM) on a fast transform, with the trap as a program and the
    moving beam: 6 real-time steps, so the merged two-half-step pass runs
    with the absorber's decay (gpe/potentials.py:31-37 assembly)."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import midpoint_times
    length, dt, beta = 16.0, 1e-3, 100.0
    x, _ = orc.get_grids(n, length)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    box = np.where(np.abs(Z) > 6.0, 50.0, 0.0)
    absorber = 0.8 * np.clip(r - 5.0, 0.0, None) ** 2
    v_static = box - 1j * absorber
    static, beam = "x**2 + y**2 + z**2", "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)"
    problem = orc.Problem(n, length, "real", dt, beta, static=static, dynamic=beam, v_static=v_static)
    solve = make(n, length, "real", dt, beta, static, beam, precision=precision)[1]
    solve.volume_upload(_lib.TGT_VOL, v_static.real)
    solve.volume_upload(_lib.TGT_VOL_IM, v_static.imag)
    psi0 = orc.synthetic_state(n, length, seed=8, beta=beta)
    solve.upload(psi0)
    solve.step(6, midpoint_times(0, 6, dt))
    expected = problem.run(psi0, 6)
    assert rel_l2(solve.to_numpy(), expected) < tol
    assert solve.norm() == pytest.approx(orc.get_N(expected, x), rel=tol * 10)
    assert orc.get_N(expected, x) < orc.get_N(psi0, x) * (1 - 1e-5)     # the absorber removed norm


@pytest.mark.parametrize("n", [18, 64, 128, 256])
@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_cutoff_mask_matches_the_oracle_on_the_sphere(n, precision, require_gpu):
    """The zeroed modes of the masked kinetic step are exactly the oracle's
    (gpe/kernels.py:24-45: zero where kx*kx + ky*ky + kz*kz > kcut*kcut).

    A delta at the origin has a flat spectrum.  With beta = 0 and no potential
    the half-steps multiply by exactly 1, so after one step the spectrum is the
    kinetic phase (modulus 1) on kept modes and 0 on cut ones.  The cutoff is
    |k| of the axis mode 9 on a box of 13.7: the axis modes tie (k^2 ==
    kcut^2, kept by the strict >), and of the other modes with integer
    i^2 + j^2 + k^2 = 81 some round above the cutoff and some below.  Summing
    in another order, contracting to FMAs or comparing with >= moves some of
    them; the input check below shows that the reversed order does."""
    length = 13.7
    kcut = abs(orc.get_grids(n, length)[1][9])
    problem, solve = make(n, length, "real", 1e-3, 0.0, kcut=kcut, precision=precision)
    k = problem.k
    k2 = orc.k_squared(k)
    kcut2 = kcut * kcut
    cut = k2 > kcut2
    reversed_order = (k[None, None, :] ** 2 + k[None, :, None] ** 2) + k[:, None, None] ** 2
    assert np.count_nonzero(k2 == kcut2) >= 3
    assert np.count_nonzero((reversed_order > kcut2) != cut) > 0
    assert 0 < np.count_nonzero(cut) < cut.size
    psi0 = np.zeros((n,) * 3, dtype=np.complex128)
    psi0[0, 0, 0] = 1.0
    solve.upload(psi0)
    solve.step(1)
    spectrum = np.abs(np.fft.fftn(solve.to_numpy()))
    solve.close()
    assert np.all(np.abs(spectrum[~cut] - 1) < 1e-4), precision
    assert np.all(spectrum[cut] < 1e-4), precision


def test_scalar_only_diagnostics_keep_the_last_spectrum(require_gpu):
    """Diagnostics without the spectrum return the same scalars and leave the
    view of the last spectrum call in place."""
    from gpe_simulator_amd import _lib
    length, n = 12.0, 64
    problem, solve = make(n, length, "real", 1e-3, 30.0, "x**2 + y**2 + z**2")
    with pytest.raises(_lib.GpeError, match="no gpe_diagnostics call"):
        out = np.zeros(_lib.VIEW_BINS, dtype=np.int64)
        _lib.check(_lib.lib().gpe_spectrum_view(solve.handle, _lib.i64ptr(out), None, _lib.VIEW_BINS))
    psi0 = random_state(n, length, seed=5)
    solve.upload(psi0)
    first = solve.diagnostics(0.0)
    solve.step(3)
    scalars = solve.diagnostics(0.0, spectrum=False)
    again = solve.diagnostics(0.0)
    for key in ("norm", "E_kin", "E_pot", "E_int", "mu"):
        assert scalars[key] == again[key], key
    out = np.zeros(_lib.VIEW_BINS, dtype=np.int64)
    occ = np.zeros(_lib.VIEW_BINS)
    _lib.check(_lib.lib().gpe_spectrum_view(solve.handle, _lib.i64ptr(out), _lib.dptr(occ), _lib.VIEW_BINS))
    np.testing.assert_array_equal(occ, again["view_occupation"])
    assert not np.array_equal(first["view_occupation"], again["view_occupation"])
    solve.close()


@pytest.mark.parametrize("n", [18, 64, 96, 256])
@pytest.mark.parametrize("precision,tol_fused,tol_oracle", [("c128", 1e-12, 1e-11), ("c64", 2e-5, 3e-5)])
def test_reference_structured_step_matches_fused_and_oracle(n, precision, tol_fused, tol_oracle, require_gpu):
    """gpe_step_unfused (one launch per reference kernel and axis) against the
    fused plan and the oracle: trap, moving beam (per-step time slots),
    momentum cutoff and an absorber volume (the complex static potential)."""
    from gpe_simulator_amd.engine import midpoint_times
    from gpe_simulator_amd import _lib
    length, steps, dt = 14.0, 6, 1e-3
    static, beam = "x**2 + y**2 + z**2", "10*exp(-((x-2*cos(1.0*t))**2+(y-2*sin(1.0*t))**2)/2)"
    kcut = 0.6 * np.pi * n / length
    problem, unfused = make(n, length, "real", dt, 40.0, static, beam, kcut, precision=precision)
    _, fused = make(n, length, "real", dt, 40.0, static, beam, kcut, precision=precision)
    x = problem.x
    absorber = 3.0 * np.maximum(np.abs(x) - 0.3 * length, 0)
    volume = -(absorber[:, None, None] + absorber[None, :, None] + absorber[None, None, :])
    for solve in (unfused, fused):
        solve.volume_upload(_lib.TGT_VOL_IM, volume.ravel())
    problem.v_static = problem.v_static + 1j * volume
    psi0 = random_state(n, length, seed=9)
    tm = midpoint_times(0, steps, dt)
    for solve in (unfused, fused):
        solve.upload(psi0)
    unfused.step_unfused(steps, tm)
    fused.step(steps, tm)
    got, ref = unfused.to_numpy(), fused.to_numpy()
    unfused.close()
    fused.close()
    assert rel_l2(got, ref) < tol_fused
    expected = problem.run(psi0, steps)
    assert rel_l2(got, expected) < tol_oracle


def test_reference_structured_step_is_real_time_only(require_gpu):
    from gpe_simulator_amd import _lib
    _, solve = make(16, 8.0, "imaginary", 1e-3, 10.0)
    with pytest.raises(_lib.GpeInvalid, match="real time only"):
        solve.step_unfused(2)
    solve.close()


def step_counts(plan, kind, steps, cut=False):
    """Launches per pass kind of `steps` steps of one call (`cut`: a solve
    with the momentum cutoff, real time)."""
    from gpe_simulator_amd import _lib
    if plan == _lib.PLAN_PLANE and cut:
        return {"zy_first": 1, "zy_mid": steps - 1, "x_kin": steps, "zy_last": 1}
    if plan == _lib.PLAN_PLANE:
        if kind == "real":
            return {"zyk_first": 1, "zyk_mid": steps - 1, "xk": steps - 1, "xk_last": 1}
        return {"zyk_first": steps, "xk_last": steps}
    if kind == "real":
        return {"zk_first": 1, "zk_mid": steps - 1, "yk": steps, "xk": steps - 1, "xk_last": 1}
    return {"zk_first": steps, "yk": steps, "xk_last": steps}


@pytest.mark.parametrize("plan_name", ["separable", "plane", "plane_cut"])
def test_stamped_step_matches_the_plain_step_and_reports_every_launch(plan_name, require_gpu):
    """gpe_step_stamped advances the state exactly as gpe_step and reports
    one positive span per pass launch (time-dependent tables break the chain
    of spans and are not charged to any pass), under the separable and the
    plane plan, and the plane plan of a solve with the momentum cutoff."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import midpoint_times
    n, length, dt, steps = 64, 12.0, 1e-3, 7
    beam = "10*exp(-((x-2*cos(1.0*t))**2+(y-2*sin(1.0*t))**2)/2)"
    cut = plan_name == "plane_cut"
    kcut = 0.6 * np.pi * n / length if cut else None
    _, stamped = make(n, length, "real", dt, 30.0, "x**2 + y**2 + z**2", beam, kcut)
    _, plain = make(n, length, "real", dt, 30.0, "x**2 + y**2 + z**2", beam, kcut)
    plan = _lib.PLAN_SEPARABLE if plan_name == "separable" else _lib.PLAN_PLANE
    psi0 = random_state(n, length, seed=3)
    tm = midpoint_times(0, steps, dt)
    for solve in (stamped, plain):
        solve.set_plan(plan)
        solve.upload(psi0)
    spans = stamped.step_stamped(steps, tm)
    plain.step(steps, tm)
    np.testing.assert_array_equal(stamped.to_numpy(), plain.to_numpy())
    counts = {k: v[1] for k, v in spans.items() if v[1]}
    assert counts == step_counts(plan, "real", steps, cut)
    for kind, (us, count) in spans.items():
        if count:
            assert 0 < us / count < 5e3, (kind, us, count)
    stamped.close()
    plain.close()


@pytest.mark.parametrize("plane", [False, True])
@pytest.mark.parametrize("kind", ["real", "imaginary", "real_cut"])
def test_kernel_timed_step_matches_the_plain_step(kind, plane, require_gpu):
    """gpe_step_kernel_times advances the state exactly as gpe_step and
    reports one positive duration per pass launch, each no longer than the
    stream interval around the whole call, under the separable and the plane
    plan (real_cut: the momentum cutoff, four-sweep or plane plan)."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import midpoint_times
    n, length, dt, steps = 64, 12.0, 1e-3, 6
    static = "x**2 + y**2 + z**2"
    cut = kind == "real_cut"
    kind = "real" if cut else kind
    kcut = 0.6 * np.pi * n / length if cut else None
    beam = "10*exp(-((x-2*cos(1.0*t))**2+(y-2*sin(1.0*t))**2)/2)" if kind == "real" else None
    _, timed = make(n, length, kind, dt, 30.0, static, beam, kcut)
    _, plain = make(n, length, kind, dt, 30.0, static, beam, kcut)
    plan = _lib.PLAN_PLANE if plane else (_lib.PLAN_FOUR_SWEEP if cut else _lib.PLAN_SEPARABLE)
    psi0 = random_state(n, length, seed=5)
    tm = midpoint_times(0, steps, dt) if beam else None
    for solve in (timed, plain):
        solve.set_plan(plan)
        solve.upload(psi0)
    timed.mark(0)
    spans = timed.step_kernel_times(steps, tm)
    timed.mark(1)
    whole_us = timed.marked_ms() * 1e3
    plain.step(steps, tm)
    np.testing.assert_array_equal(timed.to_numpy(), plain.to_numpy())
    counts = {k: v[1] for k, v in spans.items() if v[1]}
    if cut and not plane:
        assert counts == {"z_first": 1, "y_fwd": steps, "x_kin": steps, "y_inv": steps, "z_mid": steps - 1,
                          "z_last": 1}
    else:
        assert counts == step_counts(plan, kind, steps, cut)
    total = 0.0
    for name, (us, count) in spans.items():
        if count:
            assert 0 < us / count < 5e3, (name, us, count)
            total += us
    assert total <= whole_us
    timed.close()
    plain.close()
