"""gpe_step_noalloc, the allocation-free step, against gpe_step.

Two calls in a row through the C ABI, the second on a smaller grid, each
checked bit for bit against gpe_step on an identical solve (same kernels,
same summation order): the field, the per-step norm history and the norm.
The device array that holds the norm partials keeps the previous call's
contents, which must not leak into the next call.
"""

import numpy as np
import pytest

from oracle import gpe_oracle as orc

pytestmark = [pytest.mark.gpu, pytest.mark.timeout(300)]


def pair(n, kind, precision, beta, seed):
    from gpe_simulator_amd.engine import make_solve
    x, k = orc.get_grids(n, 16.0)
    psi0 = orc.synthetic_state(n, 16.0, seed=seed, beta=beta)
    out = []
    for _ in range(2):
        s = make_solve(n, x, k, kind, 5e-4, beta, static="x**2 + y**2 + 2*z**2", precision=precision)
        s.upload(psi0)
        out.append(s)
    return out


@pytest.mark.parametrize("precision", ["c64", "c128"])
def test_noalloc_step_matches_gpe_step_on_two_calls(precision, require_gpu):
    from gpe_simulator_amd import _lib
    ctx = _lib.default_context()
    for n, steps, seed in ((96, 40, 1), (64, 25, 2)):          # the second call is smaller
        free_alloc, solve = pair(n, "imaginary", precision, 300.0, seed)
        solve.step(steps)                                      # the kernels of this shape are loaded now
        free_before, _ = ctx.mem_info()
        free_alloc.step_noalloc(steps)
        free_alloc.sync()
        free_after, _ = ctx.mem_info()
        np.testing.assert_array_equal(free_alloc.to_numpy(), solve.to_numpy())
        np.testing.assert_array_equal(free_alloc.norm_history(steps), solve.norm_history(steps))
        assert free_alloc.norm() == solve.norm()
        assert free_after == free_before                       # nothing was allocated
        free_alloc.close()
        solve.close()


def test_noalloc_step_real_time_and_limits(require_gpu):
    from gpe_simulator_amd import _lib
    free_alloc, solve = pair(64, "real", "c64", 300.0, 3)
    free_alloc.step_noalloc(30)
    solve.step(30)
    np.testing.assert_array_equal(free_alloc.to_numpy(), solve.to_numpy())
    free_alloc.close()
    solve.close()
    free_alloc, solve = pair(64, "imaginary", "c64", 300.0, 4)
    with pytest.raises(_lib.GpeInvalid, match="norm history"):
        free_alloc.step_noalloc(5000)                          # beyond the history it holds
    solve.step(2000)                                           # gpe_step grows its own history
    free_alloc.close()
    solve.close()
