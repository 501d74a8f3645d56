"""Pins of the FP64 oracle (CPU).

Each check is a closed form, an invariant or a brute-force computation that
the reference's own tests use (tests/test_propagators.py:61-158,
tests/test_diagnostics.py:31-81 of the reference) or that the mathematics
fixes; a plausible slip in the oracle (a missing 1/2 on dt, k^2 with a 1/2, a
wrong sign in the nonlinear phase, a midpoint time off by dt/2, a shifted
shell) fails at least one of them.
"""

import math

import numpy as np
import pytest

from oracle import gpe_oracle as orc

TWO_PI = 2 * np.pi


def plane_wave(x, mode=1):
    n = len(x)
    wave = np.exp(1j * mode * x)[:, None, None]
    return np.broadcast_to(wave, (n,) * 3).copy() / TWO_PI ** 1.5


@pytest.mark.parametrize("static", [False, True])
@pytest.mark.parametrize("dynamic", [False, True])
@pytest.mark.parametrize("cutoff", [False, True])
def test_real_phase_closed_form(static, dynamic, cutoff):
    # tests/test_propagators.py:61-84: N=8, L=2pi, a=10, 3 steps of dt=0.01.
    beta = orc.get_beta(150000, 10)
    p = orc.Problem(8, TWO_PI, "real", 0.01, beta, static="1/2" if static else None,
                    dynamic="t/4" if dynamic else None, kcut=1.5 if cutoff else None)
    psi0 = plane_wave(p.x)
    result = p.run(psi0, 3)
    T = 3 * 0.01
    phase = (1 + beta / TWO_PI ** 3) * T + (0.5 * T if static else 0) + (T ** 2 / 8 if dynamic else 0)
    np.testing.assert_allclose(result, psi0 * np.exp(-1j * phase), rtol=1e-11, atol=1e-12)
    assert orc.get_N(result, p.x) == pytest.approx(1, abs=1e-11)


@pytest.mark.parametrize("kind,dynamic", [("imaginary", False), ("real", False), ("real", True)])
def test_expression_half_step_at_each_point(kind, dynamic):
    # tests/test_propagators.py:86-112: odd N=7, psi = 0.2+0.1i, beta = 0.
    x, _ = orc.get_grids(7, TWO_PI)
    X, Y, Z = x[:, None, None], x[None, :, None], x[None, None, :]
    potential = X / 2 + Y ** 2 / 10 + np.abs(Z) / 5
    t = 0.37
    U = orc.eval_expression("x/2 + y**2/10 + abs(z)/5", x)
    if dynamic:
        U = U + orc.eval_expression("sin(t)*x + cos(pi)*z", x, t)
        potential = potential + np.sin(t) * X - Z
    psi = np.full((7,) * 3, 0.2 + 0.1j)
    dt = 0.01
    if kind == "imaginary":
        got = orc.imag_half_step(psi, np.zeros((7,) * 3), U, 0.0, dt)
        expected = (0.2 + 0.1j) * np.exp(-potential * dt / 2)
    else:
        got = orc.real_half_step(psi, np.zeros((7,) * 3, complex), U, 0.0, dt)
        expected = (0.2 + 0.1j) * np.exp(-1j * potential * dt / 2)
    np.testing.assert_allclose(got, expected, rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("static", [False, True])
def test_imaginary_decay_and_normalisation(static):
    # tests/test_propagators.py:114-129
    p = orc.Problem(8, TWO_PI, "imaginary", 0.01, 0.0, static="1/2" if static else None)
    low, high = plane_wave(p.x, 1), plane_wave(p.x, 2)
    result = p.run(low + high, 3)
    T = 0.03
    expected = low * np.exp(-T) + high * np.exp(-4 * T)
    dx = p.x[1] - p.x[0]
    expected /= np.sqrt(np.sum(np.abs(expected) ** 2) * dx ** 3)
    np.testing.assert_allclose(result, expected, rtol=1e-11, atol=1e-12)


def test_absorber_norm_loss():
    # tests/test_propagators.py:131-144
    gamma = 0.4
    v = np.full((8,) * 3, -1j * gamma)
    p = orc.Problem(8, TWO_PI, "real", 0.01, 0.0, v_static=v)
    psi0 = plane_wave(p.x)
    result = p.run(psi0, 3)
    T = 0.03
    np.testing.assert_allclose(result, psi0 * np.exp((-gamma - 1j) * T), rtol=1e-11, atol=1e-12)
    assert orc.get_N(result, p.x) == pytest.approx(np.exp(-2 * gamma * T), abs=1e-11)


def test_cutoff_removes_high_mode():
    # tests/test_propagators.py:146-158
    p = orc.Problem(8, TWO_PI, "real", 0.01, 0.0, kcut=1.5)
    low, high = plane_wave(p.x, 1), plane_wave(p.x, 2)
    result = p.run((low + high) / np.sqrt(2), 3)
    np.testing.assert_allclose(result, low * np.exp(-1j * 0.03) / np.sqrt(2), rtol=1e-11, atol=1e-12)
    assert orc.get_N(result, p.x) == pytest.approx(0.5, abs=1e-11)


def test_norm_and_kinetic_energy_of_plane_wave():
    # tests/test_diagnostics.py:31-39: mode (2, -3, 1) has k^2 = 14.
    x, k = orc.get_grids(16, TWO_PI)
    phase = 2 * x[:, None, None] - 3 * x[None, :, None] + x[None, None, :]
    psi = np.exp(1j * phase) / TWO_PI ** 1.5
    assert orc.get_N(psi, x) == pytest.approx(1, abs=1e-12)
    assert orc.get_E(np.fft.fftn(psi), x, k) == pytest.approx(14, abs=1e-11)


@pytest.mark.parametrize("n", [2, 7, 8, 18])
def test_shells_counts_and_parseval(n):
    # tests/test_diagnostics.py:52-72, with brute-force shell indices.
    x, k = orc.get_grids(n, TWO_PI)
    rng = np.random.default_rng(42)
    psi = rng.normal(size=(n,) * 3) + 1j * rng.normal(size=(n,) * 3)
    spectrum = orc.nk_extraction(np.fft.fftn(psi), x, k)
    dk = abs(k[1] - k[0])
    counts = {}
    for a in k:
        for b in k:
            for c in k:
                s = int(math.floor(math.sqrt(a * a + b * b + c * c) / dk + 0.5))
                counts[s] = counts.get(s, 0) + 1
    brute = np.zeros(len(spectrum["k"]), dtype=np.int64)
    for s, c in counts.items():
        brute[s] = c
    np.testing.assert_array_equal(spectrum["counts"], brute)
    assert spectrum["counts"].sum() == n ** 3
    assert spectrum["Nk"].sum() == pytest.approx(orc.get_N(psi, x), rel=1e-12)
    assert spectrum["edges"][0] == 0
    np.testing.assert_array_equal(spectrum["complete_shells"], spectrum["edges"][1:] <= n * dk / 2)


def test_zero_mode_spectrum():
    # tests/test_diagnostics.py:75-81
    x, k = orc.get_grids(8, TWO_PI)
    psi = np.ones((8,) * 3, dtype=np.complex128) / TWO_PI ** 1.5
    spectrum = orc.nk_extraction(np.fft.fftn(psi), x, k)
    assert spectrum["Nk"][0] == pytest.approx(1)
    np.testing.assert_allclose(spectrum["Nk"][1:], 0, atol=1e-25)
    assert spectrum["counts"][0] == 1


@pytest.mark.parametrize("n", [2, 7, 8, 18, 64, 96, 128, 192, 256])
@pytest.mark.parametrize("length", [TWO_PI, 20.0, 24.0, 40.0, 70.0])
def test_integer_shell_identity(n, length):
    """floor(|k|/dk + 1/2) from FP64 axes equals m + (s > m^2 + m), m = isqrt(s)."""
    x, k = orc.get_grids(n, length)
    dk = abs(float(k[1] - k[0]))
    p = np.rint(k / dk).astype(np.int64)
    assert np.all(np.abs(p * dk - k) < 1e-9 * max(1.0, dk))
    P, Q, R = np.meshgrid(p, p, p, indexing="ij")
    fp64 = np.floor(np.sqrt(orc.k_squared(k)) / dk + 0.5).astype(np.int64)
    np.testing.assert_array_equal(orc.shell_index_integer(P, Q, R), fp64)


def test_fft_convention_brute_force():
    """numpy.fft.fftn is the unnormalised forward DFT e^{-2 pi i n k / N}."""
    rng = np.random.default_rng(1)
    n = 4
    a = rng.normal(size=(n,) * 3) + 1j * rng.normal(size=(n,) * 3)
    idx = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(idx, idx) / n)
    brute = np.einsum("ia,jb,kc,abc->ijk", W, W, W, a)
    np.testing.assert_allclose(np.fft.fftn(a), brute, rtol=1e-12, atol=1e-12)


def test_time_reversibility():
    """S(-dt) S(dt) psi = psi for a real potential without cutoff."""
    n, length, beta = 16, 10.0, 40.0
    fwd = orc.Problem(n, length, "real", 2e-3, beta, static="x**2 + y**2/2", dynamic="cos(t)*x")
    bwd = orc.Problem(n, length, "real", -2e-3, beta, static="x**2 + y**2/2", dynamic="cos(t)*x")
    psi0 = orc.synthetic_state(n, length, seed=4, beta=beta)
    t = 0.3
    one = fwd.step(psi0.copy(), t)
    # the forward step used t_mid = t + dt/2; the backward step from t + dt uses
    # (t + dt) + (-dt)/2, the same midpoint
    back = bwd.step(one.copy(), t + 2e-3)
    np.testing.assert_allclose(back, psi0, rtol=1e-12, atol=1e-13)


def test_stage_clock_uses_index_times_dt():
    """launch_stage passes t_i = i*dt; the oracle's run uses the same times
    (tests/test_workflow.py:262 pins [0, 0.125] for a second stage)."""
    seen = []
    p = orc.Problem(4, 8.0, "real", 0.125, 0.0, dynamic="t*y")
    original = p.W

    def spy(t):
        seen.append(t)
        return original(t)

    p.W = spy
    p.run(np.ones((4,) * 3, complex), 2)
    assert seen == [0 * 0.125 + 0.125 / 2, 1 * 0.125 + 0.125 / 2]


def test_harmonic_ground_state_energy_and_mu():
    """beta = 0, V = r^2: pi^{-3/4} exp(-r^2/2) has E = mu = 3 (reference units)."""
    p = orc.Problem(64, 20.0, "imaginary", 1e-3, 0.0, static="x**2 + y**2 + z**2")
    X, Y, Z = p.x[:, None, None], p.x[None, :, None], p.x[None, None, :]
    psi = (np.pi ** -0.75 * np.exp(-(X * X + Y * Y + Z * Z) / 2)).astype(complex)
    e = p.energies(psi)
    assert e["norm"] == pytest.approx(1, rel=1e-12)
    assert e["E"] == pytest.approx(3, rel=1e-10)
    assert e["mu"] == pytest.approx(3, rel=1e-10)
    assert e["E_kin"] == pytest.approx(1.5, rel=1e-10)


def test_anisotropic_ground_state_energy():
    """omega = (1, 1, 4): V = x^2 + y^2 + 16 z^2, ground state energy 1 + 1 + 4."""
    p = orc.Problem(64, 16.0, "imaginary", 1e-3, 0.0, static="x**2 + y**2 + 16*z**2")
    X, Y, Z = p.x[:, None, None], p.x[None, :, None], p.x[None, None, :]
    psi = np.exp(-(X * X + Y * Y) / 2 - 2 * Z * Z).astype(complex)
    psi = orc.normalise(psi, p.x)
    assert p.energies(psi)["E"] == pytest.approx(6, rel=1e-9)


def test_plane_wave_mu_and_energy():
    """e^{ik.x}/L^{3/2} in V0: mu = k^2 + V0 + beta/L^3, E = k^2 + V0 + beta/(2 L^3)."""
    length, beta, v0 = TWO_PI, 7.0, 0.5
    p = orc.Problem(16, length, "real", 1e-3, beta, static="0.5")
    phase = 2 * p.x[:, None, None] + 0 * p.x[None, :, None] + p.x[None, None, :]
    psi = np.exp(1j * phase) / length ** 1.5
    assert psi.shape == (16,) * 3
    e = p.energies(psi)
    assert e["mu"] == pytest.approx(5 + v0 + beta / length ** 3, rel=1e-12)
    assert e["E"] == pytest.approx(5 + v0 + beta / (2 * length ** 3), rel=1e-12)


def test_virial_identity_at_converged_ground_state():
    """Harmonic trap: 2 E_kin - 2 E_pot + 3 E_int = 0 at the ground state; the
    imaginary-time fixed point satisfies it up to O(dtau)."""

    def residual(dt, psi, steps):
        p = orc.Problem(32, 12.0, "imaginary", dt, 60.0, static="x**2 + y**2 + z**2")
        psi = p.run(psi, steps)
        e = p.energies(psi)
        return abs(2 * e["E_kin"] - 2 * e["E_pot"] + 3 * e["E_int"]) / e["E"], psi

    coarse, psi = residual(2e-3, orc.gaussian(orc.get_grids(32, 12.0)[0], 1.0), 1500)
    fine, _ = residual(5e-4, psi, 2000)
    assert coarse < 5e-3
    assert fine < 0.35 * coarse


def test_imaginary_time_lowers_energy_monotonically():
    p = orc.Problem(16, 10.0, "imaginary", 5e-3, 20.0, static="x**2 + y**2 + z**2")
    psi = orc.gaussian(p.x, 1.6)
    energies = []
    for _ in range(6):
        psi = p.run(psi, 10)
        energies.append(p.energies(psi)["E"])
    assert all(b < a for a, b in zip(energies, energies[1:]))


def decimal_view_bin(s, h, nview=orc.VIEW_BINS):
    """floor(nview sqrt(s) / (sqrt(3) h)) to 50 digits, clamped to the last bin.
    Where the value is an integer (3 h^2 b^2 == nview^2 s) the decimal digits
    cannot settle the floor, so the exact equality decides."""
    from decimal import Decimal, getcontext
    getcontext().prec = 50
    value = Decimal(nview) * Decimal(s).sqrt() / (Decimal(3).sqrt() * h)
    nearest = int(value.to_integral_value())
    if abs(value - nearest) < Decimal("1e-40"):
        assert 3 * h * h * nearest * nearest == nview * nview * s
        b = nearest
    else:
        b = int(value)          # floor of a positive value
    return min(b, nview - 1)


@pytest.mark.parametrize("n", [2, 3, 7, 8, 18, 64, 96])
def test_spectrum_view_bin_every_square(n):
    """Every |m|^2 the grid holds, against the defining formula."""
    h = n // 2
    s = np.arange(3 * h * h + 1)
    got = orc.view_bin_from_square(s, n)
    expected = np.array([decimal_view_bin(int(v), h) for v in s])
    np.testing.assert_array_equal(got, expected)


@pytest.mark.parametrize("n", [128, 192, 256, 1024])
def test_spectrum_view_bin_at_every_boundary(n):
    """Around each bin boundary |m|^2 = 3 h^2 b^2 / nview^2 (and at random
    squares) at large sides, against the defining formula."""
    h, nview = n // 2, orc.VIEW_BINS
    rng = np.random.default_rng(n)
    squares = set(rng.integers(0, 3 * h * h + 1, size=300).tolist()) | {0, 3 * h * h}
    for b in range(1, nview + 1):
        edge = 3 * h * h * b * b // (nview * nview)
        squares |= {v for v in range(edge - 2, edge + 3) if 0 <= v <= 3 * h * h}
    s = np.array(sorted(squares))
    np.testing.assert_array_equal(orc.view_bin_from_square(s, n),
                                  [decimal_view_bin(int(v), h) for v in s])


@pytest.mark.parametrize("n", [7, 8, 18])
def test_spectrum_view_totals_and_limits(n):
    x, k = orc.get_grids(n, 20.0)
    rng = np.random.default_rng(3)
    psi = rng.normal(size=(n,) * 3) + 1j * rng.normal(size=(n,) * 3)
    view = orc.spectrum_view(np.fft.fftn(psi), x, k)
    shells = orc.nk_extraction(np.fft.fftn(psi), x, k)
    assert view["counts"].sum() == n ** 3
    assert view["Nk"].sum() == pytest.approx(orc.get_N(psi, x), rel=1e-12)
    assert view["Nk"].sum() == pytest.approx(shells["Nk"].sum(), rel=1e-12)
    dk = abs(k[1] - k[0])
    assert view["k_max"] == pytest.approx(math.sqrt(3) * (n // 2) * dk, rel=1e-15)
    assert view["edges"][-1] == pytest.approx(view["k_max"], rel=1e-15)
    h = n // 2
    assert orc.spectrum_view_bin(0, 0, 0, n) == 0
    assert orc.spectrum_view_bin(h, h, -h, n) == orc.VIEW_BINS - 1
    # a plane wave lands in the bin of its |m|
    mode = (1, 2, -2)
    wave = np.exp(1j * (mode[0] * x[:, None, None] + mode[1] * x[None, :, None] + mode[2] * x[None, None, :])
                  * TWO_PI / 20.0)
    view = orc.spectrum_view(np.fft.fftn(wave), x, k)
    expected = orc.spectrum_view_bin(*mode, n)
    assert view["occupation"][expected] == pytest.approx(np.sum(np.abs(np.fft.fftn(wave)) ** 2), rel=1e-12)
    assert view["occupation"].sum() - view["occupation"][expected] < 1e-12 * view["occupation"][expected]


# ----------------------------------------------- initial state, beta, VM --
def test_gaussian_unit_width_is_the_oscillator_ground_state():
    """sigma = 1 on [-10, 10)^3, 64^3: the normalised discrete Gaussian equals
    the continuum ground state pi^(-3/4) exp(-r^2/2) (the grid sum of a
    Gaussian this deep inside the box is its integral to rounding)."""
    x, _ = orc.get_grids(64, 20.0)
    psi = orc.gaussian(x, 1.0)
    X, Y, Z = x[:, None, None], x[None, :, None], x[None, None, :]
    exact = np.pi ** -0.75 * np.exp(-(X * X + Y * Y + Z * Z) / 2)
    assert psi.dtype == np.complex128
    assert np.max(np.abs(psi - exact)) < 1e-13 * np.max(exact)
    assert orc.get_N(psi, x) == pytest.approx(1.0, abs=1e-13)


@pytest.mark.parametrize("sigma", [0.8, 1.3])
def test_gaussian_peak_and_profile(sigma):
    """Peak (pi sigma^2)^(-3/4) at the origin (a grid point) and the axis
    profile exp(-x^2 / (2 sigma^2)): a sigma^2 in place of 2 sigma^2, or a
    missing normalisation, fails one of them."""
    x, _ = orc.get_grids(64, 20.0)
    psi = orc.gaussian(x, sigma)
    c = 32
    assert x[c] == 0.0
    assert psi[c, c, c].real == pytest.approx((np.pi * sigma ** 2) ** -0.75, rel=1e-12)
    for i in (c + 3, c - 5, c + 8):
        ratio = (psi[i, c, c] / psi[c, c, c]).real
        assert ratio == pytest.approx(math.exp(-x[i] ** 2 / (2 * sigma ** 2)), rel=1e-12)
        assert psi[c, i, c] == psi[i, c, c] == psi[c, c, i]
    assert np.all(psi.imag == 0)


@pytest.mark.parametrize("beta, a", [(200.0, 1.0025308467604033), (1000.0, 5.012654233802016),
                                     (1100.0, 5.513919657182218), (4000.0, 20.050616935208065)])
def test_beta_of_the_workload_table(beta, a):
    """beta = 8 pi N_atom a a0 / 1 um at N_atom = 150000 reproduces the
    table of scattering lengths the workloads are set with (beta = 2g for
    g = 100, 500, 550, 2000), and the constant 1.3299673896737356e-3 per
    atom and unit of a."""
    assert orc.get_beta(150000, a) == pytest.approx(beta, rel=1e-13)
    assert orc.get_beta(1.0, 1.0) == pytest.approx(1.3299673896737356e-3, rel=1e-15)
    assert orc.get_beta(300000, a) == pytest.approx(2 * beta, rel=1e-13)


@pytest.mark.parametrize("func", ["sin", "cos", "tan", "exp", "sqrt", "tanh", "abs", "log"])
def test_expression_functions_match_math(func):
    """Every function of the grammar (config_editor/expressions.py:8) maps to
    the C function of the same value (abs -> fabs): checked against Python's
    math at points of the axis, on an argument that keeps sqrt and log real."""
    x, _ = orc.get_grids(8, 4.0)
    arg = "x + 3" if func in ("sqrt", "log") else "x"
    got = orc.eval_expression(f"{func}({arg})", x)
    ref = {"abs": abs}.get(func) or getattr(math, func)
    for i in range(8):
        val = x[i] + 3 if func in ("sqrt", "log") else x[i]
        assert got[i, 0, 0] == pytest.approx(ref(val), rel=1e-15, abs=1e-300)
        assert got[i, 5, 2] == got[i, 0, 0]          # depends on x only


def test_expression_operators_constants_and_axes():
    """Binary operators, ** as pow, unary minus, pi, t and the axis roles
    (x first index, z last) against direct evaluation."""
    x, _ = orc.get_grids(6, 3.0)
    t = 0.37
    got = orc.eval_expression("-(x**3) + 2**y / (z - 7) * pi - t*cos(pi*t)", x, t=t)
    for i, j, k in [(0, 0, 0), (1, 4, 2), (5, 2, 3)]:
        ref = -(x[i] ** 3) + 2 ** x[j] / (x[k] - 7) * math.pi - t * math.cos(math.pi * t)
        assert got[i, j, k] == pytest.approx(ref, rel=1e-14)
    assert orc.eval_expression("3", x).shape == (6, 6, 6)
    assert np.all(orc.eval_expression("-2.5", x) == -2.5)


# ----------------------------------------------------------- static volumes --
def test_box_potential_at_hand_evaluated_points():
    """Walls of a cylinder of radius 5 and length 6 (thickness 1, depth 40) on
    a 32^3 grid of spacing 0.5: depth on the tube (rho = 5 at (5,0,z) and
    (3,4,z)) and on the end sheets (|z| = 3), max(tube, sheet) elsewhere."""
    x, _ = orc.get_grids(32, 16.0)
    v = orc.box_potential(x, 5.0, 6.0, 1.0, 40.0)
    idx = {float(val): i for i, val in enumerate(x)}
    at = lambda px, py, pz: v[idx[px], idx[py], idx[pz]]      # noqa: E731
    assert at(5.0, 0.0, 0.0) == pytest.approx(40.0, rel=1e-15)
    assert at(3.0, 4.0, 0.5) == pytest.approx(40.0, rel=1e-15)
    assert at(0.0, 0.0, 3.0) == pytest.approx(40.0, rel=1e-15)
    assert at(0.0, 0.0, -3.0) == pytest.approx(40.0, rel=1e-15)
    # centre: tube exp(-25/2) < sheet exp(-9/2)
    assert at(0.0, 0.0, 0.0) == pytest.approx(40.0 * math.exp(-4.5), rel=1e-14)
    # rho = 1.5, z = 0: tube exp(-3.5^2/2) vs sheet exp(-9/2)
    assert at(1.5, 0.0, 0.0) == pytest.approx(40.0 * max(math.exp(-3.5 ** 2 / 2), math.exp(-4.5)), rel=1e-14)
    # outside the tube and past the sheet: both walls decay
    r = math.hypot(7.0, 6.0)
    assert at(7.0, 6.0, 7.0) == pytest.approx(
        40.0 * max(math.exp(-(r - 5) ** 2 / 2), math.exp(-(7 - 3) ** 2 / 2)), rel=1e-14)
    assert np.all(v <= 40.0) and np.all(v > 0)


def test_absorber_ramps_at_faces_edges_and_corners():
    """Ramps of width 2 and amplitude 3 on [-8, 8)^3 (spacing 0.5): zero
    inside, 3*2 on the first plane of each axis, 3*1.5 on the last (x = 7.5
    lies 0.5 inside the face), and the sum over the faces a point is near."""
    x, _ = orc.get_grids(32, 16.0)
    a = orc.absorption_potential(2.0, 3.0, x, 16.0)
    c = 16                                   # x = 0
    assert a[c, c, c] == 0.0
    assert a[0, c, c] == a[c, 0, c] == a[c, c, 0] == pytest.approx(6.0)
    assert a[31, c, c] == a[c, 31, c] == a[c, c, 31] == pytest.approx(4.5)
    assert a[1, c, c] == pytest.approx(3.0 * 1.5)          # x = -7.5
    assert a[4, c, c] == 0.0                                # x = -6: the ramp ends
    assert a[0, 0, c] == pytest.approx(12.0)               # edge: two faces
    assert a[0, 0, 0] == pytest.approx(18.0)               # corner: three faces
    assert a[0, 31, 0] == pytest.approx(6.0 + 4.5 + 6.0)
    assert np.all(a >= 0)


def test_total_static_potential_has_a_negative_imaginary_absorber():
    re = np.full((2, 2, 2), 1.5)
    im = np.full((2, 2, 2), 0.25)
    v = orc.total_static_potential(re, im)
    assert v.dtype == np.complex128
    assert np.all(v.real == 1.5) and np.all(v.imag == -0.25)


# ------------------------------------------------------------ monitor images --
def test_monitor_images_sums_and_modes():
    """TOF projections and the column density each integrate to the norm
    (Parseval); a single plane-wave mode lands at its fftshifted indices; the
    kz = 0 slice of a mode with kz != 0 is empty; the mid plane holds the
    field's own density and phase."""
    n, length = 16, 2 * np.pi
    x, k = orc.get_grids(n, length)
    dx, dk = x[1] - x[0], abs(k[1] - k[0])
    X, Y, Z = x[:, None, None], x[None, :, None], x[None, None, :]
    psi = np.exp(1j * (2 * X - 3 * Y + 1 * Z)) / length ** 1.5
    img = orc.monitor_images(psi, x, k)
    norm = orc.get_N(psi, x)
    assert np.sum(img["TOF_xy"]) * dk * dk == pytest.approx(norm, rel=1e-12)
    assert np.sum(img["TOF_yz"]) * dk * dk == pytest.approx(norm, rel=1e-12)
    assert np.sum(img["density_xy"]) * dx * dx == pytest.approx(norm, rel=1e-12)
    ks = np.fft.fftshift(k)
    i, j = int(np.argmin(np.abs(ks - 2))), int(np.argmin(np.abs(ks + 3)))
    peak = np.unravel_index(np.argmax(img["TOF_xy"]), img["TOF_xy"].shape)
    assert peak == (i, j)
    assert img["TOF_xy"][i, j] == pytest.approx(norm / dk ** 2, rel=1e-12)
    assert np.max(img["TOF_slice_xy"]) < 1e-25
    assert np.max(img["TOF_slice_yz"]) < 1e-25       # kx = 2 != 0
    np.testing.assert_allclose(img["density_slice_xy"], np.abs(psi[:, :, n // 2]) ** 2, rtol=1e-14)
    np.testing.assert_allclose(img["phase_xy"], np.angle(psi[:, :, n // 2]), rtol=1e-14, atol=1e-15)
    # a centred Gaussian: the projections are symmetric under k -> -k
    g = orc.gaussian(x, 0.7)
    gi = orc.monitor_images(g, x, k)
    tof = gi["TOF_xy"][1:, 1:]                        # drop the unpaired Nyquist row and column
    np.testing.assert_allclose(tof, tof[::-1, ::-1], rtol=1e-10, atol=1e-16)
    plane = np.abs(g[:, :, n // 2]) ** 2
    dark = plane <= plane.max() * 1e-12
    assert dark.any() and not dark.all()
    assert np.array_equal(np.isnan(gi["phase_xy"]), dark)
