"""Small seeded LPs of every bound and row type (test-side), and their scipy HiGHS solution.

Structurals are Free, Upper (x <= 0), Lower (x >= 0) or Boxed (0 <= x <= u); rows are equalities,
<= or >= (Problem's range types, common/types.hpp:48-56).  The mix yields optimal, primal
infeasible and unbounded LPs, which pin the termination statuses."""
import numpy as np
import scipy.sparse as sp

from parallel_ds_amd import lp as lpmod

INF = np.inf


def mixed_lp(seed, boxed=False, m_max=5, n_max=6, free_upper=True):
    rng = np.random.default_rng(seed)
    m = int(rng.integers(1, m_max)) + 1
    n = int(rng.integers(1, n_max)) + 1
    A = rng.uniform(-1, 1, (m, n)) * (rng.random((m, n)) < 0.7)
    A[np.abs(A) < 1e-3] = 0.0
    c = rng.uniform(-1, 1, n)
    types = ([1, 3] if free_upper else []) + ([4, 2] if boxed else [4])
    bt = rng.choice(types, n).astype(np.int32)
    lb = np.where((bt == 4) | (bt == 2), 0.0, -INF)
    ub = np.where(bt == 3, 0.0, np.where(bt == 2, rng.uniform(0.2, 2.0, n), INF))
    rt = rng.choice([0, 3, 4], m).astype(np.int32)
    b = rng.uniform(-1, 1, m)
    lr = np.where(rt == 3, -INF, b)
    ur = np.where(rt == 4, INF, b)
    C = sp.csc_matrix(A)
    C.sort_indices()
    return lpmod.LP(m, n, C.indptr.astype(np.int32), C.indices.astype(np.int32), C.data.astype(float), c, lb, ub,
                    bt, lr, ur, rt, "mixed-%d" % seed)


def highs(L):
    """(status, objective) from scipy's HiGHS: 0 optimal, 1 infeasible, 3 unbounded.  Row range types:
    0 equality, 3 <= upper_range, 4 >= lower_range, 2 both (a ranged row)."""
    from scipy.optimize import linprog
    A = sp.csc_matrix((L.val, L.row_idx, L.col_ptr), shape=(L.m, L.n)).toarray()
    eq = L.range_type == 0
    le = (L.range_type == 3) | (L.range_type == 2)
    ge = (L.range_type == 4) | (L.range_type == 2)
    A_ub = np.vstack([A[le], -A[ge]])
    b_ub = np.concatenate([L.upper_range[le], -L.lower_range[ge]])
    bounds = [(None if not np.isfinite(a) else a, None if not np.isfinite(b) else b)
              for a, b in zip(L.lower_bound, L.upper_bound)]
    r = linprog(L.costs, A_ub=A_ub if len(b_ub) else None, b_ub=b_ub if len(b_ub) else None,
                A_eq=A[eq] if eq.any() else None, b_eq=L.upper_range[eq] if eq.any() else None,
                bounds=bounds, method="highs")
    st = {0: 0, 2: 1, 3: 3}.get(r.status, -1)
    return st, (float(r.fun) if r.status == 0 else None)
