"""Termination statuses on the GPU against the oracle: small LPs of every bound and row type,
including primal infeasible ('dual unbounded') and unbounded ones (the reference's phase 1
restarts without end on those; both sides report the iteration cap)."""
import numpy as np
import pytest

import _oracle as ora
import mixed_lps
from parallel_ds_amd import default_params, solve_lp
from parallel_ds_amd._native import DS_ITERATION_LIMIT, DS_OPTIMAL, DS_PRIMAL_INFEASIBLE

pytestmark = pytest.mark.gpu

MESSAGES = {DS_OPTIMAL: "optimal solution", DS_PRIMAL_INFEASIBLE: "dual unbounded"}


@pytest.mark.parametrize("free_upper", [False, True])
def test_statuses_match_oracle(gpu, free_upper):
    seen = set()
    for s in range(120):
        L = mixed_lps.mixed_lp(s, free_upper=free_upper)
        try:
            r_o, x_o, b_o = ora.solve(L, max_iterations=500, tie_rule=1)
        except RuntimeError:
            continue  # the oracle's dense LU met a singular kernel
        sol, basis, res = solve_lp(L, default_params(max_iterations=500))
        assert res.status == r_o.status, (s, res.status, r_o.status)
        seen.add(res.status)
        if res.status in MESSAGES:
            assert sol.message == MESSAGES[res.status]
        if res.status == DS_OPTIMAL:
            assert abs(sol.Z - r_o.Z) <= 1e-9 * max(1.0, abs(r_o.Z)), (s, sol.Z, r_o.Z)
    assert {DS_OPTIMAL, DS_PRIMAL_INFEASIBLE, DS_ITERATION_LIMIT} <= seen


def test_infeasible_config0_row(gpu):
    """The small configuration plus one row  a^T x <= -1  with a >= 0: primal infeasible. The
    oracle and the GPU both end 'dual unbounded' (status 1), and HiGHS calls it infeasible."""
    import scipy.sparse as sp
    from parallel_ds_amd import lp as lpmod
    L0 = lpmod.config0(4)
    A = sp.csc_matrix((L0.val, L0.row_idx, L0.col_ptr), shape=(L0.m, L0.n)).tolil()
    A.resize((L0.m + 1, L0.n))
    rng = np.random.default_rng(4)
    for j in rng.choice(L0.n, 30, replace=False):
        A[L0.m, j] = rng.uniform(0.1, 1.0)
    A = A.tocsc()
    A.sort_indices()
    b = np.concatenate([L0.upper_range, [-1.0]])
    L = lpmod.packing_lp(L0.m + 1, L0.n, A.indptr, A.indices, A.data, b, L0.costs)
    assert mixed_lps.highs(L)[0] == 1
    r_o, _, _ = ora.solve(L, max_iterations=20000)
    sol, basis, res = solve_lp(L, default_params(max_iterations=20000))
    assert r_o.status == res.status == DS_PRIMAL_INFEASIBLE
    assert sol.message == "dual unbounded" and not sol.solved
