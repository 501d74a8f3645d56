"""Imaginary-time norm finished inside the next step's first pass.

Within one call of several imaginary-time steps, each step but the last
leaves its FP64 norm partials to the first pass of the next step, which sums
them in every block exactly as the separate finisher would (same order, same
bits) and publishes norm, scale and history from its first block.  The last
step of a call keeps the finisher.  So one call of k steps must give the
same field and the same per-step norm history, bit for bit, as k calls of
one step (where every step uses the finisher).  Reference step:
gpe/make_propagator.py:103-152 (normalise after every imaginary step).
"""

import numpy as np
import pytest

from oracle import gpe_oracle as orc

pytestmark = pytest.mark.gpu

STATIC = "x**2 + y**2 + 2*z**2"


def solve_for(n, precision, plan, seed=4, length=16.0, beta=200.0):
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(n, length)
    s = make_solve(n, x, k, "imaginary", 1e-3, beta, static=STATIC, precision=precision)
    s.set_plan(plan)
    s.upload(orc.synthetic_state(n, length, seed=seed, beta=beta))
    return s


@pytest.mark.parametrize("n,precision,plan_name", [
    (64, "c64", "plane"), (64, "c128", "plane"), (128, "c64", "plane"), (96, "c64", "separable"),
    (64, "c64", "separable"), (128, "c128", "separable"), (64, "c64", "four_sweep")])
def test_folded_norm_matches_the_finisher(n, precision, plan_name, require_gpu):
    from gpe_simulator_amd import _lib
    plan = {"plane": _lib.PLAN_PLANE, "separable": _lib.PLAN_SEPARABLE, "four_sweep": _lib.PLAN_FOUR_SWEEP}[plan_name]
    steps = 6
    one_call = solve_for(n, precision, plan)
    one_call.step(steps)
    hist_fold = one_call.norm_history(steps)
    psi_fold = one_call.to_numpy()
    norm_fold = one_call.norm()
    single = solve_for(n, precision, plan)
    hist_single = []
    for _ in range(steps):
        single.step(1)
        hist_single.append(single.norm_history(1)[0])
    psi_single = single.to_numpy()
    assert np.array_equal(psi_fold, psi_single)
    assert np.array_equal(hist_fold, np.array(hist_single))
    assert norm_fold == single.norm()
    # the normalised start loses norm each step, and the history records it
    assert np.all(np.isfinite(hist_fold)) and np.all(hist_fold > 0)
    one_call.close()
    single.close()


def test_folded_norm_in_a_graph_captured_batch(require_gpu):
    """Batch members (one launch per pass, replayed from a graph) fold the
    norm per member and match single steps of each member alone."""
    from gpe_simulator_amd import _lib
    n, steps = 64, 5
    solves = [solve_for(n, "c64", _lib.PLAN_PLANE, seed=10 + j, beta=100.0 + 50 * j) for j in range(3)]
    batch = _lib.Batch(solves)
    batch.step(steps, graph=True)
    hist = batch.norm_history(steps)
    got = [s.to_numpy() for s in solves]
    batch.close()
    for s in solves:
        s.close()
    for j in range(3):
        s = solve_for(n, "c64", _lib.PLAN_PLANE, seed=10 + j, beta=100.0 + 50 * j)
        h = []
        for _ in range(steps):
            s.step(1)
            h.append(s.norm_history(1)[0])
        assert np.array_equal(s.to_numpy(), got[j])
        assert np.array_equal(np.array(h), hist[j])
        s.close()
