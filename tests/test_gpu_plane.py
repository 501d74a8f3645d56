"""The plane plan (one pass per x plane for the z and y axes) against the
separable plan, the four-sweep plan and the FP64 oracle.

A plane pass applies the same operations to every value, in the same order,
as the passes it replaces (the LDS only moves values): ZK_* then YK without a
momentum cutoff; with one, Y_INV, the Z kind of the four-sweep plan and
Y_FWD around the 3D kinetic pass X_KIN.  So a run under the plane plan
matches the same run under the separable (four-sweep) plan to rounding in
every branch of the local half-step: real time with a moving beam, an
absorber, a static volume, imaginary time with renormalisation.  Batches of
plane-plan solves are bit-identical to the same solves run alone.
Reference step: gpe/make_propagator.py:38-152; cutoff: gpe/kernels.py:24-45.
"""

import numpy as np
import pytest

from oracle import gpe_oracle as orc

pytestmark = pytest.mark.gpu

STATIC = "x**2 + y**2 + 4*z**2"
BEAM = "10*exp(-((x-3*cos(1.5*t))**2+(y-3*sin(1.5*t))**2)/2)"
PLANE_SHAPES = [(64, "c64"), (96, "c64"), (128, "c64"), (64, "c128")]
# plane plan against the separable plan after a few steps (relative L2)
TOL = {"c64": 1e-5, "c128": 1e-13}


def rel_l2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def run(n, precision, kind, plan, steps, volume=None, dynamic=BEAM, seed=5, length=16.0, dt=1e-3, beta=150.0,
        kcut=None):
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    x, k = orc.get_grids(n, length)
    solve = make_solve(n, x, k, kind, dt, beta, static=STATIC, dynamic=dynamic if kind == "real" else None,
                       kcut=kcut, precision=precision)
    solve.set_plan(plan)
    assert solve.plan == plan
    if volume is not None:
        solve.volume_upload(_lib.TGT_VOL, volume.real)
        if np.any(volume.imag):
            solve.volume_upload(_lib.TGT_VOL_IM, volume.imag)
    psi0 = orc.synthetic_state(n, length, seed=seed, beta=beta)
    solve.upload(psi0)
    if kind == "real":
        solve.step(steps, midpoint_times(0, steps, dt))
    else:
        solve.step(steps)
    out = solve.to_numpy()
    diag = solve.diagnostics(steps * dt, spectrum=False)
    solve.close()
    return psi0, out, diag


@pytest.mark.parametrize("n,precision", PLANE_SHAPES)
@pytest.mark.parametrize("kind", ["real", "imaginary"])
def test_plane_plan_matches_the_separable_plan(n, precision, kind, require_gpu):
    from gpe_simulator_amd import _lib
    _, sep, d1 = run(n, precision, kind, _lib.PLAN_SEPARABLE, 7)
    _, pln, d2 = run(n, precision, kind, _lib.PLAN_PLANE, 7)
    err = rel_l2(pln, sep)
    print(f"plane vs separable n={n} {precision} {kind}: rel L2 {err:.3e}, "
          f"mu {abs(d2['mu'] / d1['mu'] - 1):.3e}, bit-identical {np.array_equal(sep, pln)}")
    assert err < TOL[precision]
    assert d2["mu"] == pytest.approx(d1["mu"], rel=TOL[precision])
    assert d2["norm"] == pytest.approx(d1["norm"], rel=TOL[precision])


@pytest.mark.parametrize("n,precision", [(64, "c64"), (128, "c64"), (64, "c128")])
def test_plane_plan_with_absorber_and_volume(n, precision, require_gpu):
    """Static box - i*absorber volume: the VOL and VIM branches of the merged
    and single half-steps."""
    from gpe_simulator_amd import _lib
    length = 16.0
    x, _ = orc.get_grids(n, length)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    volume = np.where(np.abs(Z) > 6.0, 50.0, 0.0) - 1j * 0.8 * np.clip(r - 5.0, 0.0, None) ** 2
    _, sep, _ = run(n, precision, "real", _lib.PLAN_SEPARABLE, 5, volume=volume)
    _, pln, _ = run(n, precision, "real", _lib.PLAN_PLANE, 5, volume=volume)
    print(f"absorber n={n} {precision}: rel L2 {rel_l2(pln, sep):.3e}")
    assert rel_l2(pln, sep) < TOL[precision]
    # a real-only volume (VOL without VIM) in imaginary time
    _, sep, _ = run(n, precision, "imaginary", _lib.PLAN_SEPARABLE, 4, volume=volume.real + 0j)
    _, pln, _ = run(n, precision, "imaginary", _lib.PLAN_PLANE, 4, volume=volume.real + 0j)
    print(f"volume, imaginary n={n} {precision}: rel L2 {rel_l2(pln, sep):.3e}")
    assert rel_l2(pln, sep) < TOL[precision]


@pytest.mark.parametrize("n,precision,tol", [(64, "c128", 1e-11), (64, "c64", 3e-5), (96, "c64", 3e-5),
                                             (128, "c64", 3e-5)])
@pytest.mark.parametrize("kind", ["real", "imaginary"])
def test_plane_plan_matches_the_oracle(n, precision, tol, kind, require_gpu):
    from gpe_simulator_amd import _lib
    length, dt, beta, steps = 16.0, 1e-3, 150.0, 6
    psi0, got, _ = run(n, precision, kind, _lib.PLAN_PLANE, steps)
    expected = orc.Problem(n, length, kind, dt, beta, static=STATIC,
                           dynamic=BEAM if kind == "real" else None).run(psi0, steps)
    err = rel_l2(got, expected)
    print(f"plane vs oracle n={n} {precision} {kind}: rel L2 {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("n,precision", [(96, "c64"), (64, "c128")])
def test_plane_batch_members_match_solo_runs(n, precision, require_gpu):
    """Imaginary-time members of different trap strengths in one batch
    (one launch per pass, graph-captured) against each run alone."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    length, dt, steps = 16.0, 1e-3, 6
    x, k = orc.get_grids(n, length)
    solves, starts = [], []
    for j, (w, beta) in enumerate(((1.0, 100.0), (2.0, 300.0), (0.5, 50.0))):
        s = make_solve(n, x, k, "imaginary", dt, beta, static=f"{w}*(x**2 + y**2) + z**2", precision=precision)
        s.set_plan(_lib.PLAN_PLANE)
        psi0 = orc.synthetic_state(n, length, seed=20 + j, beta=beta)
        s.upload(psi0)
        solves.append(s)
        starts.append((psi0, w, beta))
    batch = _lib.Batch(solves)
    batch.step(steps, graph=True)
    got = [s.to_numpy() for s in solves]
    batch.close()
    for s in solves:
        s.close()
    for (psi0, w, beta), g in zip(starts, got):
        s = make_solve(n, x, k, "imaginary", dt, beta, static=f"{w}*(x**2 + y**2) + z**2", precision=precision)
        s.set_plan(_lib.PLAN_PLANE)
        s.upload(psi0)
        s.step(steps)
        assert np.array_equal(s.to_numpy(), g)
        s.close()


def test_plane_plan_is_refused_where_no_plane_fits(require_gpu):
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(256, 20.0)
    s = make_solve(256, x, k, "real", 1e-3, 100.0, static=STATIC, precision="c64")
    with pytest.raises(_lib.GpeInvalid):
        s.set_plan(_lib.PLAN_PLANE)
    assert s.plan == _lib.PLAN_SEPARABLE
    s.close()
    x, k = orc.get_grids(96, 20.0)
    s = make_solve(96, x, k, "real", 1e-3, 100.0, static=STATIC, precision="c128")
    with pytest.raises(_lib.GpeInvalid):
        s.set_plan(_lib.PLAN_PLANE)
    s.close()


# ---- the plane plan with a momentum cutoff (ZY kinds around X_KIN)

def kcut_of(n, length=16.0):
    """A cutoff inside the grid's k range (0.6 k_max): it removes about 70%
    of the modes, so a kinetic step without it differs at O(1)."""
    return 0.6 * np.pi * n / length


@pytest.mark.parametrize("n,precision", PLANE_SHAPES)
def test_cut_plane_plan_matches_the_four_sweep_plan(n, precision, require_gpu):
    """(The cutoff is real-time only: gpe/make_propagator.py:5-101.)"""
    from gpe_simulator_amd import _lib
    kind = "real"
    kcut = kcut_of(n)
    _, four, d1 = run(n, precision, kind, _lib.PLAN_FOUR_SWEEP, 7, kcut=kcut)
    _, pln, d2 = run(n, precision, kind, _lib.PLAN_PLANE, 7, kcut=kcut)
    err = rel_l2(pln, four)
    print(f"cut plane vs four-sweep n={n} {precision} {kind}: rel L2 {err:.3e}, "
          f"mu {abs(d2['mu'] / d1['mu'] - 1):.3e}, bit-identical {np.array_equal(four, pln)}")
    assert err < TOL[precision]
    assert d2["mu"] == pytest.approx(d1["mu"], rel=TOL[precision])
    assert d2["norm"] == pytest.approx(d1["norm"], rel=TOL[precision])
    # the cutoff acts: without it the same run ends elsewhere
    _, uncut, _ = run(n, precision, kind, _lib.PLAN_PLANE, 7)
    assert rel_l2(uncut, pln) > 100 * TOL[precision]


@pytest.mark.parametrize("n,precision,tol", [(32, "c128", 1e-11), (48, "c128", 1e-11), (64, "c128", 1e-11),
                                             (32, "c64", 3e-5), (48, "c64", 3e-5), (64, "c64", 3e-5),
                                             (96, "c64", 3e-5), (128, "c64", 3e-5)])
def test_cut_plane_plan_matches_the_oracle(n, precision, tol, require_gpu):
    """Every plane side, real time with the moving beam and the cutoff."""
    from gpe_simulator_amd import _lib
    length, dt, beta, steps = 16.0, 1e-3, 150.0, 5
    kcut = kcut_of(n, length)
    psi0, got, _ = run(n, precision, "real", _lib.PLAN_PLANE, steps, kcut=kcut)
    expected = orc.Problem(n, length, "real", dt, beta, static=STATIC, dynamic=BEAM, kcut=kcut).run(psi0, steps)
    err = rel_l2(got, expected)
    print(f"cut plane vs oracle n={n} {precision}: rel L2 {err:.3e}")
    assert err < tol


@pytest.mark.parametrize("n,precision", [(64, "c64"), (128, "c64"), (64, "c128")])
def test_cut_plane_plan_with_absorber_and_volume(n, precision, require_gpu):
    from gpe_simulator_amd import _lib
    length = 16.0
    x, _ = orc.get_grids(n, length)
    X, Y, Z = np.meshgrid(x, x, x, indexing="ij")
    r = np.sqrt(X * X + Y * Y + Z * Z)
    volume = np.where(np.abs(Z) > 6.0, 50.0, 0.0) - 1j * 0.8 * np.clip(r - 5.0, 0.0, None) ** 2
    kcut = kcut_of(n, length)
    _, four, _ = run(n, precision, "real", _lib.PLAN_FOUR_SWEEP, 5, volume=volume, kcut=kcut)
    _, pln, _ = run(n, precision, "real", _lib.PLAN_PLANE, 5, volume=volume, kcut=kcut)
    print(f"cut absorber n={n} {precision}: rel L2 {rel_l2(pln, four):.3e}")
    assert rel_l2(pln, four) < TOL[precision]


@pytest.mark.parametrize("n,precision", [(64, "c64"), (128, "c64"), (64, "c128")])
def test_cut_plane_batch_members_match_solo_runs(n, precision, require_gpu):
    """Real-time members of different trap strengths in one graph-captured
    batch against each run alone, bit for bit."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    length, dt, steps = 16.0, 1e-3, 6
    x, k = orc.get_grids(n, length)
    kcut = kcut_of(n, length)

    def member(w, beta):
        s = make_solve(n, x, k, "real", dt, beta, static=f"{w}*(x**2 + y**2) + z**2", kcut=kcut,
                       precision=precision)
        assert s.plan == _lib.PLAN_PLANE
        return s

    solves, starts = [], []
    for j, (w, beta) in enumerate(((1.0, 100.0), (2.0, 300.0), (0.5, 50.0))):
        s = member(w, beta)
        psi0 = orc.synthetic_state(n, length, seed=40 + j, beta=beta)
        s.upload(psi0)
        solves.append(s)
        starts.append((psi0, w, beta))
    batch = _lib.Batch(solves)
    batch.step(steps, graph=True)
    got = [s.to_numpy() for s in solves]
    batch.close()
    for s in solves:
        s.close()
    for (psi0, w, beta), g in zip(starts, got):
        s = member(w, beta)
        s.upload(psi0)
        s.step(steps)
        assert np.array_equal(s.to_numpy(), g)
        s.close()


def test_cut_plane_plan_rules(require_gpu):
    """Masked solves start on the plane plan where a plane fits and the side is
    a multiple of 64, else on the four-sweep plan; the separable plan is
    refused for them; a plane-plan batch does not mix masked and unmasked
    members."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(64, 16.0)
    cut = make_solve(64, x, k, "real", 1e-3, 100.0, static=STATIC, kcut=kcut_of(64), precision="c64")
    assert cut.plan == _lib.PLAN_PLANE == _lib.default_plan(_lib.C64, _lib.REAL, masked=True, n=64)
    with pytest.raises(_lib.GpeInvalid, match="not separable"):
        cut.set_plan(_lib.PLAN_SEPARABLE)
    assert cut.plan == _lib.PLAN_PLANE
    plain = make_solve(64, x, k, "real", 1e-3, 100.0, static=STATIC, precision="c64")
    assert plain.plan == _lib.PLAN_PLANE
    batch = _lib.Batch([cut, plain])
    with pytest.raises(_lib.GpeInvalid, match="same step plan"):
        batch.step(1, graph=False)
    batch.close()
    # on the four-sweep plan a batch may mix them (X_KIN reads each member's mask)
    cut.set_plan(_lib.PLAN_FOUR_SWEEP)
    plain.set_plan(_lib.PLAN_FOUR_SWEEP)
    batch = _lib.Batch([cut, plain])
    batch.step(1, graph=False)
    batch.close()
    cut.close()
    plain.close()
    # every plane side takes the plane plan with the cutoff, as the host rule says
    for n, precision in ((32, "c64"), (48, "c64"), (96, "c64"), (48, "c128"), (128, "c128")):
        x, k = orc.get_grids(n, 16.0)
        s = make_solve(n, x, k, "real", 1e-3, 100.0, static=STATIC, kcut=kcut_of(n), precision=precision)
        prec = _lib.C64 if precision == "c64" else _lib.C128
        assert s.plan == _lib.default_plan(prec, _lib.REAL, masked=True, n=n), (n, precision)
        s.close()
    x, k = orc.get_grids(256, 20.0)
    big = make_solve(256, x, k, "real", 1e-3, 100.0, static=STATIC, kcut=kcut_of(256, 20.0), precision="c64")
    assert big.plan == _lib.PLAN_FOUR_SWEEP == _lib.default_plan(_lib.C64, _lib.REAL, masked=True, n=256)
    with pytest.raises(_lib.GpeInvalid):
        big.set_plan(_lib.PLAN_PLANE)
    big.close()


def test_cut_plane_long_run_matches_the_oracle(require_gpu):
    """200 complex128 steps at 64^3 on the cutoff plane plan (moving beam,
    cutoff at 0.6 k_max) against the FP64 oracle: the ZY passes accumulate
    no error beyond rounding over many steps."""
    from gpe_simulator_amd import _lib
    n, length, dt, beta, steps = 64, 16.0, 1e-3, 150.0, 200
    kcut = kcut_of(n, length)
    psi0, got, _ = run(n, "c128", "real", _lib.PLAN_PLANE, steps, kcut=kcut)
    expected = orc.Problem(n, length, "real", dt, beta, static=STATIC, dynamic=BEAM, kcut=kcut).run(psi0, steps)
    err = rel_l2(got, expected)
    print(f"cut plane vs oracle, {steps} steps: rel L2 {err:.3e}")
    assert err < 1e-10
