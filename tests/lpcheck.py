"""Feasibility / optimality certificates computed from a basis with numpy (test-side)."""
import numpy as np


def full_matrix(lp):
    """Dense [A | I] of the computational form."""
    A = lp.dense()
    return np.hstack([A, np.eye(lp.m)])


def comp_bounds(lp):
    l = np.concatenate([lp.lower_bound, -lp.upper_range])
    u = np.concatenate([lp.upper_bound, -lp.lower_range])
    c = np.concatenate([lp.costs, np.zeros(lp.m)])
    return l, u, c


def certificate(lp, x, basis):
    """(primal infeasibility, dual infeasibility, objective) of (x, basis) in computational form."""
    AF = full_matrix(lp)
    l, u, c = comp_bounds(lp)
    pinf = max(0.0, float(np.max(l - x)), float(np.max(x - u)), float(np.max(np.abs(AF @ x))))
    B = AF[:, basis]
    y = np.linalg.solve(B.T, c[basis])
    d = c - AF.T @ y
    nb = np.setdiff1d(np.arange(lp.n + lp.m), basis)
    dinf = 0.0
    for j in nb:
        at_l = np.isfinite(l[j]) and abs(x[j] - l[j]) <= 1e-9 * max(1, abs(l[j]))
        at_u = np.isfinite(u[j]) and abs(x[j] - u[j]) <= 1e-9 * max(1, abs(u[j]))
        if at_l and at_u:
            continue
        if at_l:
            dinf = max(dinf, -d[j])
        elif at_u:
            dinf = max(dinf, d[j])
        else:
            dinf = max(dinf, abs(d[j]))
    return pinf, dinf, float(c @ x)


def nondegenerate(lp, x, basis, tol=1e-9):
    """True when every basic variable is strictly inside its bounds."""
    l, u, _ = comp_bounds(lp)
    xb = x[basis]
    return bool(np.all((xb - l[basis] > tol) & (u[basis] - xb > tol)))
