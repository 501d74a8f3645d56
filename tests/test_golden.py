"""Committed oracle fixtures (tests/golden/oracle_n12.npz, written by
tests/golden/make_fixtures.py).

CPU: the oracle still reproduces them, so its outputs cannot change unnoticed.
GPU: the engine reproduces them without the oracle in the loop: complex128
psi within relative L2 1e-11 after 3 steps, energies within 1e-11 relative,
shell counts bit-exact; complex64 psi within 2e-5.
"""

import importlib.util
import json
from pathlib import Path

import numpy as np
import pytest

GOLDEN = Path(__file__).resolve().parent / "golden"
FIXTURE = GOLDEN / "oracle_n12.npz"


def fixtures():
    with np.load(FIXTURE, allow_pickle=False) as f:
        data = {k: f[k] for k in f.files}
    data["cases"] = json.loads(str(data["cases"]))
    return data


def maker():
    spec = importlib.util.spec_from_file_location("make_fixtures", GOLDEN / "make_fixtures.py")
    module = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(module)
    return module


def rel_l2(a, b):
    return float(np.linalg.norm((a - b).ravel()) / np.linalg.norm(b.ravel()))


def test_oracle_reproduces_the_fixtures():
    data, fresh = fixtures(), maker().compute()
    assert data["cases"] == [list(c) for c in maker().CASES]
    np.testing.assert_array_equal(data["psi0"], fresh["psi0"])
    for i in range(len(data["cases"])):
        assert rel_l2(fresh[f"out{i}"], data[f"out{i}"]) < 1e-13, i
    np.testing.assert_allclose(fresh["energies"], data["energies"], rtol=1e-13)
    np.testing.assert_array_equal(fresh["nk_counts"], data["nk_counts"])
    np.testing.assert_allclose(fresh["nk_Nk"], data["nk_Nk"], rtol=1e-12, atol=1e-15)


def test_fixture_cases_are_distinct():
    """Each case moves the state, and differently from the others."""
    data = fixtures()
    outs = [data[f"out{i}"] for i in range(len(data["cases"]))]
    for i, a in enumerate(outs):
        assert rel_l2(a, data["psi0"]) > 1e-4, i
        for b in outs[:i]:
            assert rel_l2(a, b) > 1e-4


def engine(case, precision):
    from gpe_simulator_amd.engine import make_solve
    from gpe_simulator_amd.grids import get_grids
    kind, static, dynamic, kcut, dt, beta = case
    mk = maker()
    x, k = get_grids(mk.N, mk.LENGTH)
    return make_solve(mk.N, x, k, kind, dt, beta, static=static, dynamic=dynamic, kcut=kcut,
                      precision=precision)


@pytest.mark.gpu
@pytest.mark.parametrize("precision,tol", [("c128", 1e-11), ("c64", 2e-5)])
def test_engine_reproduces_the_fixtures(precision, tol, require_gpu):
    from gpe_simulator_amd.engine import midpoint_times
    data = fixtures()
    steps = maker().STEPS
    for i, case in enumerate(data["cases"]):
        solve = engine(case, precision)
        solve.upload(data["psi0"])
        solve.step(steps, midpoint_times(0, steps, case[4]) if case[2] else None)
        assert rel_l2(solve.to_numpy(), data[f"out{i}"]) < tol, (i, case)
        solve.close()


@pytest.mark.gpu
def test_engine_diagnostics_reproduce_the_fixtures(require_gpu):
    data, mk = fixtures(), maker()
    solve = engine(data["cases"][1], "c128")
    solve.upload(data["psi0"])
    got = solve.diagnostics(mk.DIAG_T)
    for key, value in zip(data["energy_keys"], data["energies"]):
        assert got[str(key)] == pytest.approx(value, rel=1e-11), key
    np.testing.assert_array_equal(got["counts"], data["nk_counts"])
    dx = mk.LENGTH / mk.N
    np.testing.assert_allclose(got["occupation"] * dx ** 3 / mk.N ** 3, data["nk_Nk"], rtol=1e-10, atol=1e-16)
    solve.close()
