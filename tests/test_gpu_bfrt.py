"""Bound-flipping ratio test on the GPU (sequentialDualSimplex.cpp:398-482): boxed variables
against the oracle and HiGHS, with both ratio-test rules."""
import numpy as np
import pytest

import _oracle as ora
import lpcheck
import mixed_lps
from parallel_ds_amd import default_params, solve_lp
from parallel_ds_amd import lp as lpmod
from parallel_ds_amd._native import DS_OPTIMAL

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ratio_test", [0, 1])
@pytest.mark.parametrize("m,n,density", [(50, 100, 0.1), (200, 400, 0.05), (1000, 2000, 0.01)])
def test_boxed_family_vs_oracle_and_highs(gpu, ratio_test, m, n, density):
    for s in range(3):
        L = lpmod.boxed_instance(m, n, density, s)
        r_o, x_o, b_o = ora.solve(L, tie_rule=1, ratio_test=ratio_test, cleanup=1)
        sol, basis, res = solve_lp(L, default_params(max_iterations=100000, ratio_test=ratio_test))
        hs, hz = mixed_lps.highs(L)
        assert res.status == r_o.status == DS_OPTIMAL and hs == 0
        assert abs(sol.Z - hz) <= 1e-9 * abs(hz), (s, sol.Z, hz)
        assert abs(sol.Z - r_o.Z) <= 1e-9 * abs(r_o.Z)
        l, u, _ = lpcheck.comp_bounds(L)
        x = sol.x
        assert np.all(x >= l - 1e-9) and np.all(x <= u + 1e-9)
        if m <= 200:
            pinf, dinf, _ = lpcheck.certificate(L, x, basis)
            assert pinf <= 1e-7 and dinf <= 1e-7


def test_boxed_mixed_small_lps_vs_oracle(gpu):
    """Small LPs with boxed, free and upper-bounded structurals and every row type: the GPU's
    status equals the oracle's (including 'dual unbounded' and the capped restart loops), and
    its objective equals the oracle's when optimal; this covers the flip loop that exhausts its
    candidates, which both sides restate from the reference."""
    seen = set()
    exhausted = 0
    for s in range(150):
        L = mixed_lps.mixed_lp(s, boxed=True)
        try:
            r_o, _, _ = ora.solve(L, max_iterations=500, tie_rule=1)
        except RuntimeError:
            continue
        sol, basis, res = solve_lp(L, default_params(max_iterations=500))
        assert res.status == r_o.status, (s, res.status, r_o.status)
        seen.add(res.status)
        exhausted += r_o.bfrt_exhausted > 0
        if res.status == DS_OPTIMAL:
            assert abs(sol.Z - r_o.Z) <= 1e-9 * max(1.0, abs(r_o.Z)), (s, sol.Z, r_o.Z)
    assert DS_OPTIMAL in seen and exhausted > 0
