"""Time-dependent potentials compiled at run time (hiprtc) against the
on-device interpreter, and both against the FP64 oracle.

The compiled tables evaluate the same FP64 operations, in the same order and
without FMA contraction, with the same device math functions, so a run with
compiled programs must be bit-identical to the same run through the
interpreter.  Programs that differ only in their constants share one module.
"""

import numpy as np
import pytest

from oracle import gpe_oracle as orc

pytestmark = pytest.mark.gpu

BEAM = "10*exp(-((x-4*cos(1.0*t))**2+(y-4*sin(1.0*t))**2)/2)"
CASES = {
    "xy beam": (None, BEAM),
    "z drive": ("x**2 + y**2 + z**2", "3*sin(2.5*t)*z**2 + 0.5*tanh(t)*z"),
    "volume": ("x**2 + y**2 + z**2", "2*cos(0.7*t)*x*y*z/10 + 0.1*exp(-(x*x+y*y+z*z)/8)*sqrt(1+t*t)"),
    "mixed": ("x**2 + y**2 + z**2", BEAM + " + 0.5*z*z*cos(t) + 0.01*x*y*z*t"),
}


def run(n, length, static, dynamic, precision, jit, steps=12):
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve, midpoint_times
    x, k = orc.get_grids(n, length)
    psi0 = orc.synthetic_state(n, length, seed=3, beta=200.0)
    _lib.set_jit(jit)
    try:
        solve = make_solve(n, x, k, "real", 1e-3, 200.0, static=static, dynamic=dynamic, precision=precision)
    finally:
        _lib.set_jit(True)
    solve.upload(psi0)
    solve.step(steps, midpoint_times(0, steps, 1e-3))
    out = solve.to_numpy()
    diag = solve.diagnostics(steps * 1e-3, spectrum=False)
    solve.close()
    return psi0, out, diag


@pytest.mark.parametrize("precision", ["c64", "c128"])
@pytest.mark.parametrize("case", sorted(CASES))
def test_compiled_programs_match_the_interpreter_bit_for_bit(case, precision, require_gpu):
    from gpe_simulator_amd import _lib
    static, dynamic = CASES[case]
    before = _lib.jit_stats()
    _, compiled, d1 = run(64, 16.0, static, dynamic, precision, jit=True)
    after = _lib.jit_stats()
    _, interpreted, d2 = run(64, 16.0, static, dynamic, precision, jit=False)
    assert after["failures"] == before["failures"]
    assert after["launches"] > before["launches"], "the compiled tables were not used"
    assert np.array_equal(compiled, interpreted)
    assert d1["E"] == d2["E"] and d1["mu"] == d2["mu"]


def test_compiled_programs_match_the_oracle(require_gpu):
    static, dynamic = CASES["mixed"]
    n, length, steps = 32, 16.0, 12
    psi0, got, _ = run(n, length, static, dynamic, "c128", jit=True, steps=steps)
    problem = orc.Problem(n, length, "real", 1e-3, 200.0, static=static, dynamic=dynamic)
    expected = problem.run(psi0, steps)
    err = float(np.linalg.norm(got - expected) / np.linalg.norm(expected))
    assert err < 1e-11, err


def test_constants_share_one_module(require_gpu):
    """Stirrers of different speeds and strengths compile once."""
    from gpe_simulator_amd import _lib
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(32, 16.0)
    template = "{a}*exp(-((x-4*cos({w}*t))**2+(y-4*sin({w}*t))**2)/2) + 0.25*cos({w}*t)*x*y*z"
    make_solve(32, x, k, "real", 1e-3, 100.0, dynamic=template.format(a=7, w=0.31), precision="c64").close()
    before = _lib.jit_stats()
    for a, w in ((3, 0.5), (9, 1.7), (1.5, 2.25)):
        make_solve(32, x, k, "real", 1e-3, 100.0, dynamic=template.format(a=a, w=w), precision="c64").close()
    after = _lib.jit_stats()
    assert after["compiles"] == before["compiles"]
    assert after["failures"] == before["failures"]
