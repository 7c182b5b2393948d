"""GPU parity tests: libpds.so kernels through the C-ABI against the CPU oracle."""
import numpy as np
import pytest

import _oracle as ora
import lpcheck
from parallel_ds_amd import DualSimplex, default_params, solve_lp
from parallel_ds_amd import lp as lpmod
from parallel_ds_amd._native import DS_OPTIMAL, lib, ptr

pytestmark = pytest.mark.gpu

INF = np.inf
CAP = default_params(max_iterations=50000)


def analytic_lps():
    """The four LPs of the reference's tests/simple_tests_DS.cpp with their stated optima."""
    T = [
        (3, 5, [1, .5, .5, 1, 1, 1, -.5, .5], [0, 3, 5, 8], [1, 2, 4, 2, 3, 0, 2, 4], [2, -1, 3, -2, 1],
         [1.5, 2, .5], [1.5, 2, .5], [0] * 5, [0] * 3, -4.5),
        (3, 5, [1, 1, -1, 5, 1, 3, -5, 1, 4], [0, 3, 6, 9], [0, 1, 4, 0, 2, 4, 0, 3, 4], [-14, 5, -2, 1, -8],
         [5, 41, 15], [5, 41, 15], [0] * 5, [0] * 3, -72.0),
        (2, 3, [2, 1, 1, 3, 8, 2], [0, 3, 6], [0, 1, 2, 0, 1, 2], [3, 2, 3], [-INF, 8], [2, INF], [0, 0, 1],
         [3, 4], 4.5),
        (3, 5, [1, 1, 1, 1, 3, 1, 1, 1], [0, 3, 6, 8], [0, 1, 2, 0, 1, 3, 0, 4], [-2, -3, 0, 0, 0],
         [5, 9, 4], [5, 9, 4], [0] * 5, [0] * 3, -12.0),
    ]
    out = []
    import scipy.sparse as sp
    for (m, n, vals, rp, ci, c, lr, ur, lb, rt, z) in T:
        A = sp.csr_matrix((vals, ci, rp), shape=(m, n)).tocsc()
        A.sort_indices()
        L = lpmod.LP(m, n, A.indptr.astype(np.int32), A.indices.astype(np.int32), A.data.astype(float),
                     np.array(c, float), np.array(lb, float), np.full(n, INF), np.full(n, 4, np.int32),
                     np.array(lr, float), np.array(ur, float), np.array(rt, np.int32))
        out.append((L, z))
    return out


@pytest.mark.parametrize("use_mfma", [0, 1])
def test_trailing_update_tile(gpu, use_mfma):
    rng = np.random.default_rng(7)
    for nb in (64, 13):
        C = rng.standard_normal((64, 64))
        Pm = rng.standard_normal((64, 64))
        B = rng.standard_normal((nb, 64))
        ref = C - Pm[:, :nb] @ B
        out = np.ascontiguousarray(C.copy())
        rc = lib().ds_debug_gemm_tile(ptr(out), ptr(np.ascontiguousarray(Pm)), ptr(np.ascontiguousarray(B)), nb, use_mfma)
        assert rc == 0
        np.testing.assert_allclose(out, ref, rtol=0, atol=1e-12 * nb)


@pytest.mark.parametrize("case", range(4))
def test_analytic_lps(gpu, case):
    L, z = analytic_lps()[case]
    sol, basis, res = solve_lp(L, CAP)
    assert res.status == DS_OPTIMAL
    assert sol.message == "optimal solution"
    assert abs(sol.Z - z) <= 1e-9 * max(1.0, abs(z))


@pytest.mark.parametrize("seed", range(5))
def test_config0_vs_oracle(gpu, seed):
    L = lpmod.config0(seed)
    r_o, x_o, b_o = ora.solve(L)
    sol, basis, res = solve_lp(L, CAP)
    assert res.status == r_o.status == 0
    assert abs(sol.Z - r_o.Z) <= 1e-9 * abs(r_o.Z)
    pinf, dinf, z = lpcheck.certificate(L, sol.x, basis)
    assert pinf <= 1e-7 and dinf <= 1e-7
    if lpcheck.nondegenerate(L, x_o, b_o):
        assert sorted(basis.tolist()) == sorted(b_o.tolist())


def test_pivot_row_bitwise(gpu):
    L = lpmod.config0(3)
    with DualSimplex(L, CAP) as s:
        s.initDualSimplex()
        rng = np.random.default_rng(1)
        rho = rng.standard_normal(L.m)
        a = s.pivot_row(rho)
        nb = s.nonbasis()
    # oracle: Matrix::dotUpdate UPDATE_T on [A | I] over the nonbasic columns
    import scipy.sparse as sp
    AF = sp.hstack([sp.csc_matrix((L.val, L.row_idx, L.col_ptr), shape=(L.m, L.n)), sp.eye(L.m)]).tocsr()
    AF.sort_indices()
    ref = ora.dot_update((AF.data, AF.indptr, AF.indices, L.m, L.n + L.m), rho, np.zeros(L.n + L.m), 1.0, 0.0,
                         nb, 1, 0, L.n)
    assert np.array_equal(a.view(np.int64), ref.view(np.int64))


def test_btran_ftran_on_optimal_basis(gpu):
    L = lpmod.config0(1)
    r_o, x_o, b_o = ora.solve(L)
    AF = lpcheck.full_matrix(L)
    B = AF[:, b_o]
    with DualSimplex(L, CAP) as s:
        s.initDualSimplex()
        s.set_basis(b_o)
        rng = np.random.default_rng(2)
        for p in (0, 7, L.m - 1):
            rho = s.btran(p)
            e = np.zeros(L.m)
            e[p] = 1
            ref = np.linalg.solve(B.T, e)
            assert np.abs(rho - ref).max() <= 1e-10 * max(1, np.abs(ref).max())
        b = rng.standard_normal(L.m)
        x = s.ftran(b)
        ref = np.linalg.solve(B, b)
        assert np.abs(x - ref).max() <= 1e-10 * max(1, np.abs(ref).max())
        for j in (0, 5, L.n + 3):
            x = s.ftran_column(j)
            ref = np.linalg.solve(B, AF[:, j])
            assert np.abs(x - ref).max() <= 1e-10 * max(1, np.abs(ref).max())


@pytest.mark.parametrize("S,D,seed", [(10, 20, 2), (50, 100, 2)])
def test_transport_vs_oracle(gpu, S, D, seed):
    L = lpmod.gen_transport(S, D, seed)
    r_o, x_o, b_o = ora.solve(L)
    sol, basis, res = solve_lp(L, CAP)
    assert res.status == r_o.status == 0
    assert res.perturbed == r_o.perturbed
    assert abs(sol.Z - r_o.Z) <= 1e-9 * abs(r_o.Z)


@pytest.mark.parametrize("seed", [0, 1])
def test_btran_ftran_large_kernel(gpu, seed):
    """A basis whose kernel exceeds one tournament chunk (k = 341 > 256): the panels take the
    single-round tournament of eight waves (up to 512 rows per panel)."""
    L = lpmod.gen_bernoulli(600, 1200, 0.02, 40 + seed)
    r_o, x_o, b_o = ora.solve(L)
    assert r_o.status == 0
    AF = lpcheck.full_matrix(L)
    B = AF[:, b_o]
    assert (b_o < L.n).sum() > 256
    with DualSimplex(L, CAP) as s:
        s.initDualSimplex()
        s.set_basis(b_o)
        rng = np.random.default_rng(seed)
        b = rng.standard_normal(L.m)
        x = s.ftran(b)
        ref = np.linalg.solve(B, b)
        assert np.abs(x - ref).max() <= 1e-9 * max(1, np.abs(ref).max())
        for p in (3, 301, L.m - 2):
            rho = s.btran(p)
            e = np.zeros(L.m)
            e[p] = 1
            ref = np.linalg.solve(B.T, e)
            assert np.abs(rho - ref).max() <= 1e-9 * max(1, np.abs(ref).max())


def test_transport_full_size_vs_highs(gpu):
    """The balanced transportation LP at full size (500 x 1000: m = 1,500, n = 500,000, nnz = 10^6),
    solved from the slack basis: optimal, objective equal to HiGHS's to 1e-9 relative. The oracle
    is too slow for a test at this size (about 10 s); the bench times it."""
    import json
    import os
    gold = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "highs_objectives.json")))
    ref = gold["transport_500x1000_seed2"]["objective"]
    L = lpmod.config2(seed=2, S=500, D=1000)
    assert (L.m, L.n, L.nnz) == (1500, 500000, 1000000)
    with DualSimplex(L, CAP) as s:
        s.initDualSimplex()
        assert s.presolve("minInfeas") == 0
        s.solve()
        res = s.result()
        sol = s.getSolution()
    assert res.status == DS_OPTIMAL and sol.solved
    assert abs(sol.Z - ref) <= 1e-9 * abs(ref)


def test_time_pivot_row_ext(gpu):
    """The SpMV timing entry: positive warm and cold per-launch times, algorithmic bytes of the slack basis."""
    L = lpmod.config0(2)
    with DualSimplex(L, CAP) as s:
        s.initDualSimplex()
        w, c, by = s.time_pivot_row_ext(np.ones(L.m), reps=5, flush_bytes=16 << 20)
        w2, c2, _ = s.time_pivot_row_ext(np.ones(L.m), reps=3, flush_bytes=0)
    assert w > 0 and c > 0 and w2 > 0 and c2 == 0
    assert by == 12.0 * L.nnz + 4.0 * (L.n + 1) + 8.0 * L.n + 8.0 * L.m


@pytest.mark.parametrize("k,dep", [(300, None), (300, (2, 0, 1)), (900, None), (900, (200, 150, 190))])
def test_refactor_flags_a_numerically_singular_kernel(gpu, k, dep):
    """A k x k basis kernel (k = 300: one tournament round per panel; k = 900: panels with more
    than 512 rows left merge candidate lists over several rounds) with, in the dependent cases,
    one column an exact combination of two earlier ones (k = 900: column 200, in panel 3 of the
    merge path).  Rounding leaves its pivot near 1e-16 of the row's scale, not at zero; the
    refactorization must still report it singular (DS_ESINGULAR).  Without the dependency the
    same basis factors, and FTRAN and BTRAN match dense solves."""
    from parallel_ds_amd._native import DsError
    rng = np.random.default_rng(5)
    m, n = k + 20, k + 100
    A = (rng.random((m, n)) < 0.05) * rng.uniform(0.1, 1.0, (m, n))
    A[np.arange(m), np.arange(m) % n] += rng.uniform(0.5, 1.0, m)
    if dep is not None:
        c, c0, c1 = dep
        A[:, c] = A[:, c0] + 0.7 * A[:, c1]
    import scipy.sparse as sp
    C = sp.csc_matrix(A)
    C.sort_indices()
    L = lpmod.packing_lp(m, n, C.indptr.astype(np.int32), C.indices.astype(np.int32), C.data.copy(),
                         A @ np.ones(n) + 1.0, -np.ones(n))
    basis = np.concatenate([np.arange(k), n + np.arange(k, m)]).astype(np.int32)
    with DualSimplex(L, CAP) as s:
        s.initDualSimplex()
        if dep is not None:
            with pytest.raises(DsError) as e:
                s.set_basis(basis)
            assert e.value.code == -3
        else:
            s.set_basis(basis)
            B = np.hstack([A, np.eye(m)])[:, basis]
            b = rng.standard_normal(m)
            x = s.ftran(b)
            ref = np.linalg.solve(B, b)
            assert np.abs(x - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())
            for p in (0, k // 2, m - 1):
                rho = s.btran(p)
                e = np.zeros(m)
                e[p] = 1
                ref = np.linalg.solve(B.T, e)
                assert np.abs(rho - ref).max() <= 1e-9 * max(1.0, np.abs(ref).max())
