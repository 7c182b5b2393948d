"""The GPU path applies the PFI eta file through the inverse MF of the unit triangular
eta-pivot matrix (maintained one row per basis change).  This checks that algebra against the
oracle's sequential PFIsolve (linalg.cpp:12-46), including repeated pivot positions."""
import numpy as np
import pytest

import _oracle as ora


def build_mf(etas, piv):
    t = len(piv)
    MF = np.zeros((t, t))
    last = np.full(t, -1)
    for tt in range(t):
        prev = [j for j in range(tt) if piv[j] == piv[tt]]
        lt = max(prev) if prev else -1
        last[tt] = lt
        ecol = etas[:tt, piv[tt]]
        for j in range(tt):
            lo = max(j, lt)
            MF[tt, j] = float(np.dot(ecol[lo:tt], MF[lo:tt, j]))
        MF[tt, tt] = 1.0
    return MF, last


def ftran_mf(etas, piv, x0):
    MF, last = build_mf(etas, piv)
    xm = np.array([x0[piv[k]] if last[k] < 0 else 0.0 for k in range(len(piv))])
    s = MF @ xm
    out = x0.copy()
    for i in range(len(x0)):
        xi = x0[i]
        for k in range(len(piv)):
            if piv[k] == i:
                xi = s[k] * etas[k][i]
            elif s[k] != 0:
                xi = xi + s[k] * etas[k][i]
        out[i] = xi
    return out


def btran_mf(etas, piv, p, m):
    MF, last = build_mf(etas, piv)
    t = len(piv)
    at_p = [k for k in range(t) if piv[k] == p]
    kmax = max(at_p) if at_p else -1
    b = np.array([etas[k][p] if k >= kmax else 0.0 for k in range(t)])
    u = MF.T @ b
    v = np.zeros(m)
    v[p] = 1.0 if kmax < 0 else u[min(at_p)]
    for P in set(int(q) for q in piv):
        if P != p:
            v[P] = u[min(k for k in range(t) if piv[k] == P)]
    return v


@pytest.mark.parametrize("trial", range(40))
def test_mf_form_equals_sequential_pfi(trial):
    rng = np.random.default_rng(trial)
    m = 11
    t = int(rng.integers(1, 16))
    piv = rng.integers(0, m, t).astype(np.int32)
    etas = rng.standard_normal((t, m))
    x0 = rng.standard_normal(m)
    np.testing.assert_allclose(ftran_mf(etas, piv, x0), ora.pfi_solve(etas, piv, x0, False), rtol=1e-11, atol=1e-11)
    p = int(rng.integers(0, m))
    e = np.zeros(m)
    e[p] = 1.0
    np.testing.assert_allclose(btran_mf(etas, piv, p, m), ora.pfi_solve(etas, piv, e, True), rtol=1e-11, atol=1e-11)
