"""Long-run parity against the FP64 oracle and precision checks on the device.

Bars (stated per test): complex128 psi within relative L2 1e-10 of the oracle
after 1000 steps, mu and E within 1e-11 relative; complex64 + FP64 reductions
mu within 1e-6 relative of the FP64 result.
"""

import numpy as np
import pytest

from oracle import gpe_oracle as orc

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(600)]


@pytest.fixture(autouse=True, scope="module")
def threaded_oracle():
    orc.use_threads(8)
    yield
    orc.use_threads(None)

BEAM = "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)"


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def threaded_run(p, psi, steps, workers=16):
    """Problem.run of the oracle at large sizes: the same per-point functions
    (real_half_step, the kinetic factor exp(-i k^2 dt)) applied slab by slab
    on worker threads, and the oracle's threaded 3D FFTs.  Real time, no
    momentum cutoff."""
    from concurrent.futures import ThreadPoolExecutor
    assert p.kind == "real" and p.kcut is None
    psi = np.array(psi, dtype=np.complex128, copy=True)
    kin = np.exp(-1j * orc.k_squared(p.k) * p.dt)
    slabs = [slice(i, min(i + 8, p.n)) for i in range(0, p.n, 8)]
    with ThreadPoolExecutor(workers) as pool:
        def each(fn):
            list(pool.map(fn, slabs))
        for i in range(steps):
            potential = p.U
            if p.dynamic_source:
                potential = potential + p.W(i * p.dt + p.dt / 2)
            half = lambda s: orc.real_half_step(psi[s], p.v_static[s], potential[s], p.beta, p.dt)  # noqa: E731
            each(half)
            psi_k = orc._fftn(psi)
            each(lambda s: psi_k[s].__imul__(kin[s]))
            psi[...] = orc._ifftn(psi_k)
            each(half)
    return psi


def engine(problem, precision, kind=None):
    from gpe_simulator_amd.engine import make_solve
    return make_solve(problem.n, problem.x, problem.k, kind or problem.kind, problem.dt, problem.beta,
                      static=problem.static_source, dynamic=problem.dynamic_source, precision=precision)


def test_imaginary_ground_state_2000_steps_c128(require_gpu):
    """The CPU-runnable configuration in reference units: 64^3 on [-10,10)^3,
    V = x^2+y^2+z^2, beta = 200, Gaussian start, dtau = 5e-4, renormalised
    every step, 2000 steps."""
    p = orc.Problem(64, 20.0, "imaginary", 5e-4, 200.0, static="x**2 + y**2 + z**2")
    psi0 = orc.gaussian(p.x, 1.0)
    expected = p.run(psi0, 2000)
    solve = engine(p, "c128")
    solve.gaussian(1.0)
    solve.step(2000)
    got = solve.to_numpy()
    assert rel_l2(got, expected) < 1e-10
    e_ref, e_dev = p.energies(expected), solve.diagnostics(0.0, spectrum=False)
    assert e_dev["mu"] == pytest.approx(e_ref["mu"], rel=1e-11)
    assert e_dev["E"] == pytest.approx(e_ref["E"], rel=1e-11)
    assert e_dev["norm"] == pytest.approx(1.0, abs=1e-13)


def test_stirring_1000_steps_c128(require_gpu):
    """Real-time stirring at 64^3 with the moving Gaussian beam, 1000 steps."""
    from gpe_simulator_amd.engine import midpoint_times
    p = orc.Problem(64, 16.0, "real", 5e-4, 400.0, static="x**2 + y**2 + z**2", dynamic=BEAM)
    psi0 = orc.synthetic_state(64, 16.0, seed=11, beta=400.0)
    expected = p.run(psi0, 1000)
    solve = engine(p, "c128")
    solve.upload(psi0)
    solve.step(1000, midpoint_times(0, 1000, p.dt))
    got = solve.to_numpy()
    assert rel_l2(got, expected) < 1e-10
    t = 1000 * p.dt
    e_ref, e_dev = p.energies(expected, t), solve.diagnostics(t, spectrum=False)
    assert e_dev["mu"] == pytest.approx(e_ref["mu"], rel=1e-11)
    assert e_dev["E"] == pytest.approx(e_ref["E"], rel=1e-11)
    # real time conserves the norm (no absorber)
    assert e_dev["norm"] == pytest.approx(orc.get_N(psi0, p.x), rel=1e-12)


def test_c64_mu_against_fp64_at_128(require_gpu):
    """complex64 field + FP64 reductions: mu within 1e-6 of the FP64 oracle
    after 100 stirring steps at 128^3."""
    from gpe_simulator_amd.engine import midpoint_times
    p = orc.Problem(128, 24.0, "real", 5e-4, 1000.0, static="x**2 + y**2 + 16*z**2", dynamic=BEAM)
    psi0 = orc.synthetic_state(128, 24.0, seed=2, beta=1000.0,
                               potential=p.U)
    expected = p.run(psi0, 100)
    solve = engine(p, "c64")
    solve.upload(psi0)
    solve.step(100, midpoint_times(0, 100, p.dt))
    t = 100 * p.dt
    mu_ref = p.energies(expected, t)["mu"]
    diag = solve.diagnostics(t, spectrum=False)
    assert abs(diag["mu"] - mu_ref) / abs(mu_ref) < 1e-6
    assert rel_l2(solve.to_numpy(), expected) < 1e-4


def test_headline_256_c64_vs_c128(require_gpu):
    """256^3 stirring run: c64 and c128 engines agree on mu to 1e-6 and the
    c128 norm is conserved to 1e-12 over 200 steps; c128 is parity-tested
    against the oracle above."""
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    x, k = orc.get_grids(256, 40.0)
    psi0 = orc.synthetic_state(256, 40.0, seed=0, beta=4000.0)
    out = {}
    for precision in ("c64", "c128"):
        solve = make_solve(256, x, k, "real", 2.5e-4, 4000.0, static="x**2 + y**2 + z**2", dynamic=BEAM,
                           precision=precision)
        solve.upload(psi0)
        n0 = solve.norm()
        solve.step(200, midpoint_times(0, 200, 2.5e-4))
        out[precision] = (solve.diagnostics(200 * 2.5e-4, spectrum=False), n0)
        solve.close()
    d64, d128 = out["c64"][0], out["c128"][0]
    assert abs(d64["mu"] - d128["mu"]) / abs(d128["mu"]) < 1e-6
    assert d128["norm"] == pytest.approx(out["c128"][1], rel=1e-12)
    assert d64["norm"] == pytest.approx(out["c64"][1], rel=1e-6)


def test_time_reversibility_256_c128(require_gpu):
    """Full-size property: S(-dt) S(dt) psi = psi for the real potential."""
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(256, 40.0)
    psi0 = orc.synthetic_state(256, 40.0, seed=5, beta=4000.0)
    fwd = make_solve(256, x, k, "real", 2.5e-4, 4000.0, static="x**2 + y**2 + z**2", dynamic=BEAM)
    fwd.upload(psi0)
    fwd.step(1, np.array([0.3 + 1.25e-4]))
    mid = fwd.to_numpy()
    fwd.close()
    # For a real potential conj o S(dt) o conj = S(-dt) with the same midpoint
    # time, and S(-dt) S(dt) is the identity.
    back = make_solve(256, x, k, "real", 2.5e-4, 4000.0, static="x**2 + y**2 + z**2", dynamic=BEAM)
    back.upload(np.conj(mid))
    back.step(1, np.array([0.3 + 1.25e-4]))
    np.testing.assert_allclose(np.conj(back.to_numpy()), psi0, rtol=0, atol=1e-12 * np.abs(psi0).max())


def test_quench_128_c64_against_c128(require_gpu):
    """The quench configuration in reference units: 128^3 on [-12,12)^3,
    V = x^2 + y^2 + 16 z^2 (omega = (1,1,4)), beta = 1000 (g = 500) relaxed in
    complex128 to |dmu| < 2e-10, then beta raised by 10% and 10000 real-time
    steps of dt = 5e-5 in complex128 and in complex64 from the same state.

    complex128 conserves the norm to 1e-11 and the energy to 1e-7 relative
    (Strang splitting is not exactly energy-conserving). complex64 stays
    within 1e-6 of it in mu and in the energy per particle E/N, and within
    1e-4 in the relative L2 distance of the density. Its norm drifts
    (measured -1.2e-6 over these 10000 steps), so the bar on N is 5e-6. A
    Fourier mode turns by k^2 dt per step, which is 1e-6..1e-4 rad for the
    modes that hold most of the norm. A rotation that small moves the larger
    float component by less than half an ulp, so round-to-nearest keeps that
    component while the other moves. The value then follows the tangent
    instead of the circle, and the rounding error has a systematic sign.
    """
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    from gpe_simulator_amd.sweep import mu_from_norms
    n, length, static = 128, 24.0, "x**2 + y**2 + 16.0*z**2"
    x, k = orc.get_grids(n, length)
    ground = make_solve(n, x, k, "imaginary", 5e-4, 1000.0, static=static, precision="c128")
    ground.gaussian(1.0)
    for chunk in range(400):
        ground.step(50)
        mu = mu_from_norms(ground.norm_history(50), 5e-4)
        if abs(mu[-1] - mu[-2]) < 2e-10:
            break
    else:
        pytest.fail("ground state did not converge")
    psi0 = ground.to_numpy()
    ground.close()
    steps, dt = 10000, 5e-5
    tm = midpoint_times(0, steps, dt)
    out = {}
    for precision in ("c128", "c64"):
        s = make_solve(n, x, k, "real", dt, 1100.0, static=static, precision=precision)
        s.upload(psi0)
        d0 = s.diagnostics(0.0, spectrum=False)
        s.step(steps, tm)
        out[precision] = (d0, s.diagnostics(steps * dt, spectrum=False), s.to_numpy())
        s.close()
    (a0, a1, psi128), (b0, b1, psi64) = out["c128"], out["c64"]
    assert a1["norm"] == pytest.approx(a0["norm"], rel=1e-11)
    assert a1["E"] == pytest.approx(a0["E"], rel=1e-7)
    assert b1["norm"] == pytest.approx(a1["norm"], rel=5e-6)
    assert b1["E"] / b1["norm"] == pytest.approx(a1["E"] / a1["norm"], rel=1e-6)
    assert b1["mu"] == pytest.approx(a1["mu"], rel=1e-6)
    rho128, rho64 = np.abs(psi128) ** 2, np.abs(psi64) ** 2
    assert np.linalg.norm(rho64 - rho128) / np.linalg.norm(rho128) < 1e-4


def test_quench_trap_1000_steps_c128_at_128(require_gpu):
    """Config 2's grid and trap at full size: 128^3 on [-12,12)^3,
    V = x^2 + y^2 + 16 z^2, beta = 1100 (the quenched g = 550), dt = 5e-5,
    1000 real-time steps from a seeded Thomas-Fermi start: complex128 psi
    within relative L2 1e-10 of the FP64 oracle, mu and E within 1e-11."""
    p = orc.Problem(128, 24.0, "real", 5e-5, 1100.0, static="x**2 + y**2 + 16*z**2")
    psi0 = orc.synthetic_state(128, 24.0, seed=21, beta=1100.0, potential=p.U)
    expected = threaded_run(p, psi0, 1000)
    solve = engine(p, "c128")
    solve.upload(psi0)
    solve.step(1000)
    assert rel_l2(solve.to_numpy(), expected) < 1e-10
    e_ref, e_dev = p.energies(expected), solve.diagnostics(1000 * p.dt, spectrum=False)
    assert e_dev["mu"] == pytest.approx(e_ref["mu"], rel=1e-11)
    assert e_dev["E"] == pytest.approx(e_ref["E"], rel=1e-11)
    assert e_dev["norm"] == pytest.approx(orc.get_N(psi0, p.x), rel=1e-12)
    solve.close()


def test_stirring_50_steps_c128_at_256(require_gpu):
    """The headline workload's own parameters at full size: 256^3 on
    [-20,20)^3, V = r^2, beta = 4000, the moving beam, dt = 2.5e-4, 50 steps
    from the seeded Thomas-Fermi start: complex128 psi within relative L2
    1e-10 of the FP64 oracle, mu and E within 1e-11, and the complex64 run
    of the same steps within 1e-6 of the oracle's mu."""
    from gpe_simulator_amd.engine import midpoint_times
    p = orc.Problem(256, 40.0, "real", 2.5e-4, 4000.0, static="x**2 + y**2 + z**2", dynamic=BEAM)
    psi0 = orc.synthetic_state(256, 40.0, seed=0, beta=4000.0)
    expected = threaded_run(p, psi0, 50)
    t = 50 * p.dt
    e_ref = p.energies(expected, t)
    for precision in ("c128", "c64"):
        solve = engine(p, precision)
        solve.upload(psi0)
        solve.step(50, midpoint_times(0, 50, p.dt))
        e_dev = solve.diagnostics(t, spectrum=False)
        if precision == "c128":
            assert rel_l2(solve.to_numpy(), expected) < 1e-10
            assert e_dev["mu"] == pytest.approx(e_ref["mu"], rel=1e-11)
            assert e_dev["E"] == pytest.approx(e_ref["E"], rel=1e-11)
        else:
            assert e_dev["mu"] == pytest.approx(e_ref["mu"], rel=1e-6)
            assert rel_l2(solve.to_numpy(), expected) < 1e-4
        solve.close()


@pytest.mark.parametrize("precision", ["c128", "c64"])
def test_fast_512_matches_the_oracle(precision, require_gpu):
    """The 512 transform (16 threads per pencil, 32 values each) under every
    compiled variant and both step plans: one real-time step with the
    moving beam and one imaginary-time step against the FP64 oracle
    (complex128: relative L2 1e-11; complex64: 2e-5)."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    n, length, dt, beta = 512, 48.0, 5e-4, 400.0
    real = orc.Problem(n, length, "real", dt, beta, static="x**2 + y**2 + 2*z**2", dynamic=BEAM)
    psi0 = orc.synthetic_state(n, length, seed=9, beta=beta)
    expected = threaded_run(real, psi0, 1)
    imag = orc.Problem(n, length, "imaginary", dt, beta, static="x**2 + y**2 + 2*z**2")
    expected_imag = imag.step(psi0.copy())
    tol = 1e-11 if precision == "c128" else 2e-5
    nvar = len(_lib.pass_variants(n, _lib.C64 if precision == "c64" else _lib.C128))
    assert nvar >= 2
    for v in range(nvar):
        print(f"512 {precision} variant {v}", flush=True)
        s = make_solve(n, real.x, real.k, "real", dt, beta, static=real.static_source, dynamic=BEAM,
                       precision=precision)
        s.set_variants([v] * len(_lib.PASS_KINDS))
        for plan in (_lib.PLAN_FOUR_SWEEP, _lib.PLAN_SEPARABLE):
            s.set_plan(plan)
            s.upload(psi0)
            s.step(1, midpoint_times(0, 1, dt))
            assert rel_l2(s.to_numpy(), expected) < tol, (v, plan)
        s.close()
        s = make_solve(n, imag.x, imag.k, "imaginary", dt, beta, static=imag.static_source, precision=precision)
        s.set_variants([v] * len(_lib.PASS_KINDS))
        s.upload(psi0)
        s.step(1)
        assert rel_l2(s.to_numpy(), expected_imag) < tol, v
        s.close()


@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_free_gaussian_at_1024(precision, require_gpu):
    """The 1024 transform on the reference's largest benchmarked grid: a
    Gaussian (sigma = 1.5) on [-32, 32)^3 with no potential and no
    interaction spreads as the closed form
    psi = (pi s0)^(-3/4) (s0/s)^(3/2) exp(-r^2 / (2 s)), s = s0 + 2 i t
    (i dpsi/dt = -lap psi; the spectral step is exact for it), checked on the
    x = 0 plane after 40 steps of dt = 1e-3, with the norm kept."""
    from gpe_simulator_amd.engine import make_solve
    n, length, dt, steps, sigma = 1024, 64.0, 1e-3, 40, 1.5
    x, k = orc.get_grids(n, length)
    s = make_solve(n, x, k, "real", dt, 0.0, precision=precision)
    s.gaussian(sigma)
    n0 = s.norm()
    s.step(steps)
    plane = s.download((n // 2) * n * n, n * n).reshape(n, n)   # psi[x = 0, :, :]
    t = steps * dt
    s0 = sigma ** 2
    st = s0 + 2j * t
    Y, Z = x[:, None], x[None, :]
    exact = (np.pi * s0) ** -0.75 * (s0 / st) ** 1.5 * np.exp(-(Y * Y + Z * Z) / (2 * st))
    tol = 1e-10 if precision == "c128" else 1e-5
    assert np.max(np.abs(plane - exact)) < tol * np.abs(exact).max()
    assert s.norm() == pytest.approx(n0, rel=1e-12 if precision == "c128" else 1e-5)
    s.close()
