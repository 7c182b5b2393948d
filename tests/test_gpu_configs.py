"""GPU tests at the benchmark configurations: the scaled single-LP recipe and the learned-warm-start
base LP against HiGHS fixtures, the full-size single LP through size-independent residuals, and
the Harris ratio test at the small and transportation configurations."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import _oracle as ora
import lpcheck
from parallel_ds_amd import DualSimplex, default_params, solve_lp
from parallel_ds_amd import lp as lpmod
from parallel_ds_amd._native import DS_OPTIMAL

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HI = json.load(open(os.path.join(G, "highs_objectives.json")))
CAP = default_params(max_iterations=200000)


def sparse_certificate(L, x, basis):
    """(primal infeasibility, dual infeasibility) of (x, basis) with sparse products and one
    dense basis solve (m <= a few thousand)."""
    A = sp.csc_matrix((L.val, L.row_idx, L.col_ptr), shape=(L.m, L.n))
    AF = sp.hstack([A, sp.eye(L.m)]).tocsc()
    l, u, c = lpcheck.comp_bounds(L)
    pinf = max(0.0, float(np.max(l - x)), float(np.max(x - u)), float(np.max(np.abs(AF @ x))))
    B = AF[:, basis].toarray()
    y = np.linalg.solve(B.T, c[basis])
    d = c - AF.T @ y
    nb = np.setdiff1d(np.arange(L.n + L.m), basis)
    at_l = np.isfinite(l[nb]) & (np.abs(x[nb] - l[nb]) <= 1e-9 * np.maximum(1, np.abs(np.nan_to_num(l[nb]))))
    at_u = np.isfinite(u[nb]) & (np.abs(x[nb] - u[nb]) <= 1e-9 * np.maximum(1, np.abs(np.nan_to_num(u[nb]))))
    dn = d[nb]
    viol = np.where(at_l & ~at_u, -dn, np.where(at_u & ~at_l, dn, np.where(at_l & at_u, 0.0, np.abs(dn))))
    return pinf, float(max(0.0, viol.max()))


@pytest.mark.parametrize("key,recipe", [("kcol_m2000_n4000_k10_seed1", (2000, 4000, 10, 1)),
                                        ("kcol_m2000_n5000_k5_seed4", (2000, 5000, 5, 4))])
def test_single_lp_m2000_vs_highs(gpu, key, recipe):
    """The single-LP recipe scaled to m = 2,000 (10 nonzeros per column, seed 1) and the
    learned-warm-start base LP (m = 2,000, n = 5,000, 5 per column, seed 4), solved from the
    slack basis with refactor every 64: optimal, objective within 1e-9 of HiGHS, the same
    (nondegenerate, hence unique) basis, feasibility within 1e-7."""
    g = HI[key]
    m, n, k, seed = recipe
    L = lpmod.gen_kcol(m, n, k, seed)
    prm = default_params(max_iterations=400000, refactor_freq=64)
    sol, basis, res = solve_lp(L, prm)
    print("%s: %d iterations (%d phase 1), %.1f ms, kernel %d, %d refactors"
          % (key, res.iterations, res.phase1_iterations, res.solve_ms, res.kernel_k, res.refactors))
    assert res.status == DS_OPTIMAL and sol.message == "optimal solution"
    assert abs(sol.Z - g["objective"]) <= 1e-9 * abs(g["objective"])
    assert g["nondegenerate"]
    assert sorted(basis.tolist()) == sorted(g["basis"])
    pinf, dinf = sparse_certificate(L, sol.x, basis)
    assert pinf <= 1e-7 and dinf <= 1e-7


def test_full_size_single_lp_factor_residuals(gpu):
    """The full single-LP configuration (m = 20,000, n = 40,000, 10 per column, refactor every
    64), run for 300 iterations from the slack basis: 300 = 4 x 64 + 44, so the factor in use is
    the last refactorization plus 44 eta columns. BTRAN, FTRAN and the entering-column FTRAN on
    that basis must satisfy the residual bound of a backward-stable solve against the sparse
    basis matrix, ||B x - b||_inf <= 1e-10 ||B||_inf ||x||_inf (and B^T for BTRAN)."""
    L = lpmod.config1(seed=1)
    prm = default_params(max_iterations=300, refactor_freq=64)
    A = sp.csc_matrix((L.val, L.row_idx, L.col_ptr), shape=(L.m, L.n))
    AF = sp.hstack([A, sp.eye(L.m)]).tocsc()
    with DualSimplex(L, prm) as s:
        s.initDualSimplex()
        s.presolve("minInfeas")
        s.solve()
        r = s.result()
        assert r.iterations == 300 and r.status == 3
        basis = s.basis()
        B = AF[:, basis]
        nB = abs(B).sum(axis=1).max()
        rng = np.random.default_rng(9)
        for p in (0, 17, 9999, L.m - 1):
            rho = s.btran(p)
            e = np.zeros(L.m)
            e[p] = 1.0
            assert np.abs(B.T @ rho - e).max() <= 1e-10 * nB * np.abs(rho).max()
        b = rng.standard_normal(L.m)
        x = s.ftran(b)
        assert np.abs(B @ x - b).max() <= 1e-10 * nB * np.abs(x).max()
        nonb = s.nonbasis()
        for j in (int(nonb[0]), int(nonb[123]), int(nonb[-1])):
            x = s.ftran_column(j)
            aj = AF[:, j].toarray().ravel()
            assert np.abs(B @ x - aj).max() <= 1e-10 * nB * np.abs(x).max()
        assert r.kernel_k > 0


@pytest.mark.parametrize("seed", range(5))
def test_config0_harris_vs_oracle(gpu, seed):
    """Harris two-pass ratio test (ratio_test = 1) on the small configuration: the status of the
    oracle in the same mode, the objective within 1e-9 of HiGHS and of the oracle, certified."""
    L = lpmod.config0(seed)
    g = HI["config0"][seed]
    r_o, x_o, b_o = ora.solve(L, ratio_test=1, tie_rule=1)
    sol, basis, res = solve_lp(L, default_params(max_iterations=50000, ratio_test=1))
    assert res.status == r_o.status == 0
    assert abs(sol.Z - g["objective"]) <= 1e-9 * abs(g["objective"])
    assert abs(sol.Z - r_o.Z) <= 1e-9 * abs(r_o.Z)
    pinf, dinf, _ = lpcheck.certificate(L, sol.x, basis)
    assert pinf <= 1e-7 and dinf <= 1e-7


def test_transport_full_size_harris_vs_highs(gpu):
    """The full transportation LP (m = 1,500, n = 500,000) with the Harris ratio test, the mode
    this degenerate configuration is meant to exercise: optimal at HiGHS' objective."""
    ref = HI["transport_500x1000_seed2"]["objective"]
    L = lpmod.config2(seed=2, S=500, D=1000)
    with DualSimplex(L, default_params(max_iterations=200000, ratio_test=1)) as s:
        s.initDualSimplex()
        assert s.presolve("minInfeas") == 0
        s.solve()
        res = s.result()
        sol = s.getSolution()
    print("transport harris: %d iterations, %.1f ms, perturbed %d" % (res.iterations, res.solve_ms, res.perturbed))
    assert res.status == DS_OPTIMAL and sol.solved
    assert abs(sol.Z - ref) <= 1e-9 * abs(ref)
