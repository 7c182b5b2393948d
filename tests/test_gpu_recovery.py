"""Recovery paths of the reference's loop forced from the host (sequentialDualSimplex.cpp:285-293,
:428-433; singular factorizations), and warm starts from a given basis, on the GPU. Every forced
run must end where the unforced run and HiGHS end."""
import json
import os

import numpy as np
import pytest
import scipy.sparse as sp

import _oracle as ora
import lpcheck
from parallel_ds_amd import DualSimplex, default_params, solve_lp
from parallel_ds_amd import lp as lpmod
from parallel_ds_amd._native import (DS_FORCE_DUAL_RESTART, DS_FORCE_SINGULAR, DS_FORCE_STABILITY, DS_OPTIMAL,
                                     DsError)

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HI = json.load(open(os.path.join(G, "highs_objectives.json")))


def solve(L, **kw):
    kw.setdefault("max_iterations", 50000)
    return solve_lp(L, default_params(**kw))


@pytest.mark.parametrize("event", [DS_FORCE_STABILITY, DS_FORCE_DUAL_RESTART, DS_FORCE_SINGULAR])
@pytest.mark.parametrize("seed", [0, 1])
def test_forced_recovery_reaches_the_optimum(gpu, event, seed):
    L = lpmod.config0(seed)
    ref = HI["config0"][seed]["objective"]
    sol0, b0, r0 = solve(L, refactor_freq=64)
    assert r0.status == DS_OPTIMAL
    at = r0.phase1_iterations + 5  # inside phase 2
    sol, basis, r = solve(L, refactor_freq=64, force_event=event, force_at=at)
    print("event %d at %d: %d -> %d iterations, refactors %d -> %d, restarts %d -> %d, repairs %d"
          % (event, at, r0.iterations, r.iterations, r0.refactors, r.refactors, r0.restarts, r.restarts, r.repairs))
    assert r.status == DS_OPTIMAL
    assert abs(sol.Z - ref) <= 1e-9 * abs(ref)
    pinf, dinf, _ = lpcheck.certificate(L, sol.x, basis)
    assert pinf <= 1e-7 and dinf <= 1e-7
    if event == DS_FORCE_STABILITY:
        assert r.refactors >= r0.refactors + 1
    elif event == DS_FORCE_DUAL_RESTART:
        assert r.restarts >= r0.restarts + 1
    else:
        assert r.repairs >= 1


@pytest.mark.parametrize("seed", [0, 3, 5])
def test_warm_start_from_the_optimal_basis(gpu, seed):
    """From HiGHS' (nondegenerate, unique) optimal basis the solve needs no pivot."""
    L = lpmod.config0(seed)
    g = HI["config0"][seed]
    if not g["nondegenerate"]:
        pytest.skip("degenerate instance")
    sol, basis, r = solve_lp(L, default_params(max_iterations=1000), start=np.array(g["basis"], np.int32))
    assert r.status == DS_OPTIMAL and r.iterations <= 2 and r.repairs == 0
    assert abs(sol.Z - g["objective"]) <= 1e-9 * abs(g["objective"])
    assert sorted(basis.tolist()) == g["basis"]


@pytest.mark.parametrize("seed", [0, 3])
def test_warm_start_matches_oracle_solve_from(gpu, seed):
    """From a nonsingular basis of 40 random structurals (and slacks on the other rows), the GPU
    and the oracle's solve-from reach the same status and objective (HiGHS')."""
    import scipy.linalg as sla
    L = lpmod.config0(seed)
    ref = HI["config0"][seed]["objective"]
    rng = np.random.default_rng(seed)
    A = L.dense()
    struct = rng.choice(L.n, 40, replace=False)
    _, _, piv = sla.qr(A[:, struct].T, pivoting=True)
    rows = piv[:40]
    start = np.sort(np.concatenate([struct, L.n + np.setdiff1d(np.arange(L.m), rows)])).astype(np.int32)
    r_o, _, _ = ora.solve_from(L, start, tie_rule=1)
    sol, basis, r = solve_lp(L, default_params(max_iterations=50000), start=start)
    assert r.status == r_o.status == DS_OPTIMAL and r.repairs == 0
    assert abs(sol.Z - ref) <= 1e-9 * abs(ref) and abs(r_o.Z - ref) <= 1e-9 * abs(ref)


def _dependent_lp(m, n, k, dep_col, seed):
    rng = np.random.default_rng(seed)
    A = (rng.random((m, n)) < 0.05) * rng.uniform(0.1, 1.0, (m, n))
    A[np.arange(m), np.arange(m) % n] += rng.uniform(0.5, 1.0, m)
    if dep_col is not None:
        A[:, dep_col] = A[:, 0] + 0.7 * A[:, 1]
    C = sp.csc_matrix(A)
    C.sort_indices()
    L = lpmod.packing_lp(m, n, C.indptr.astype(np.int32), C.indices.astype(np.int32), C.data.copy(),
                         A @ np.ones(n) + 1.0, -np.ones(n))
    basis = np.concatenate([np.arange(k), n + np.arange(k, m)]).astype(np.int32)
    return L, A, basis


@pytest.mark.parametrize("m,k,dep_col", [(320, 300, 2), (1200, 900, 700)])
def test_warm_start_repairs_a_singular_start(gpu, m, k, dep_col):
    """A starting basis with one structural an exact combination of two others: the hook-level
    ds_set_basis rejects it (DS_ESINGULAR); ds_init_basis replaces the dependent column by a slack
    (repairs >= 1) and the solve ends at HiGHS' optimum. At k = 900 the dependency sits in panel
    10 of the factorization and the tournament runs merge rounds."""
    import mixed_lps
    L, A, basis = _dependent_lp(m, 400 if m < 1000 else 1000, k, dep_col, 5)
    with DualSimplex(L, default_params(max_iterations=100000)) as s:
        s.initDualSimplex()
        with pytest.raises(DsError) as e:
            s.set_basis(basis)
        assert e.value.code == -3
    hs, hz = mixed_lps.highs(L)
    sol, b, r = solve_lp(L, default_params(max_iterations=100000), start=basis)
    assert r.repairs >= 1
    assert r.status == DS_OPTIMAL and hs == 0
    assert abs(sol.Z - hz) <= 1e-9 * abs(hz)


def test_multiround_factor_without_dependency(gpu):
    """k = 900 structurals (merge rounds of the tournament in the first panels): BTRAN and FTRAN
    on that basis against dense solves."""
    L, A, basis = _dependent_lp(1200, 1000, 900, None, 6)
    B = np.hstack([A, np.eye(1200)])[:, basis]
    rng = np.random.default_rng(1)
    with DualSimplex(L, default_params()) as s:
        s.initDualSimplex()
        s.set_basis(basis)
        b = rng.standard_normal(1200)
        x = s.ftran(b)
        ref = np.linalg.solve(B, b)
        assert np.abs(x - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())
        for p in (0, 650, 1199):
            rho = s.btran(p)
            e = np.zeros(1200)
            e[p] = 1.0
            ref = np.linalg.solve(B.T, e)
            assert np.abs(rho - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())


@pytest.mark.parametrize("cands", [2, 4, 8])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_multiple_candidate_rows_reach_the_optimum(gpu, cands, seed):
    """Several leaving-row candidates per major iteration (the reference's parallel dual
    simplex, on one GPU): a different pivot sequence, the same optimum (HiGHS) and status."""
    L = lpmod.config0(seed)
    ref = HI["config0"][seed]["objective"]
    sol0, b0, r0 = solve(L)
    sol, basis, r = solve(L, candidates=cands)
    print("candidates %d: %d iterations (sequential %d)" % (cands, r.iterations, r0.iterations))
    assert r.status == DS_OPTIMAL
    assert abs(sol.Z - ref) <= 1e-9 * abs(ref)
    pinf, dinf, _ = lpcheck.certificate(L, sol.x, basis)
    assert pinf <= 1e-7 and dinf <= 1e-7


@pytest.mark.parametrize("cands", [4])
def test_multiple_candidates_on_boxed_and_scaled_single_lp(gpu, cands):
    import mixed_lps
    for s in range(2):
        L = lpmod.boxed_instance(200, 400, 0.05, s)
        hs, hz = mixed_lps.highs(L)
        sol, basis, r = solve(L, candidates=cands)
        assert r.status == DS_OPTIMAL and abs(sol.Z - hz) <= 1e-9 * abs(hz)
    g = HI["kcol_m2000_n4000_k10_seed1"]
    L = lpmod.gen_kcol(2000, 4000, 10, 1)
    sol, basis, r = solve(L, candidates=cands, refactor_freq=64, max_iterations=400000)
    print("scaled single LP, %d candidates: %d iterations, %.0f ms" % (cands, r.iterations, r.solve_ms))
    assert r.status == DS_OPTIMAL and abs(sol.Z - g["objective"]) <= 1e-9 * abs(g["objective"])
    assert sorted(basis.tolist()) == sorted(g["basis"])
