"""CPU checks that pin the warm-start checker (oracle/warm_ref.py): the MLP gradient
against central finite differences, one Adam step against hand-computed values, the
features of a 2 x 3 LP worked out by hand, and the top-m selection."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import warm_ref as W  # noqa: E402


def small_net(rng, width, kpad, nf):
    _, total = W.layout(width, kpad)
    p = rng.normal(0.0, 0.5, total)
    lay, _ = W.layout(width, kpad)
    o, s = lay["W1"]
    W1 = p[o:o + s[0] * s[1]].reshape(s)
    W1[:, nf:] = 0.0
    return p


@pytest.mark.parametrize("weighted", [False, True])
def test_loss_grad_matches_central_differences(weighted):
    rng = np.random.default_rng(3)
    width, kpad, nf, S = 6, 4, 3, 9
    p = small_net(rng, width, kpad, nf)
    F = rng.normal(0, 1, (S, nf))
    y = (rng.random(S) < 0.4).astype(float)
    w = rng.uniform(0.5, 2.0, S) if weighted else None
    mu, sd = rng.normal(0, 0.1, nf), rng.uniform(0.8, 1.2, nf)
    L, g = W.loss_grad(p, F, y, w, mu, sd, width, kpad)
    assert abs(L - W.loss(p, F, y, w, mu, sd, width, kpad)) < 1e-14
    h = 1e-6
    fd = np.zeros_like(p)
    for i in range(len(p)):
        e = np.zeros_like(p)
        e[i] = h
        fd[i] = (W.loss(p + e, F, y, w, mu, sd, width, kpad) - W.loss(p - e, F, y, w, mu, sd, width, kpad)) / (2 * h)
    assert np.max(np.abs(fd - g)) < 1e-7 * max(1.0, np.max(np.abs(g)))
    # padded input columns get no gradient
    lay, _ = W.layout(width, kpad)
    o, s = lay["W1"]
    assert np.all(g[o:o + s[0] * s[1]].reshape(s)[:, nf:] == 0.0)


def test_device_precision_gradient_reduces_to_fp64_without_rounding():
    rng = np.random.default_rng(4)
    width, kpad, nf, S = 8, 4, 3, 11
    p = small_net(rng, width, kpad, nf)
    F = rng.normal(0, 1, (S, nf))
    y = (rng.random(S) < 0.5).astype(float)
    mu, sd = np.zeros(nf), np.ones(nf)
    L, g = W.loss_grad(p, F.astype(np.float32).astype(np.float64), y, None, mu, sd, width, kpad)
    L2, g2, _ = W.loss_grad_device(p, F, y, None, mu, sd, lambda v: np.asarray(v, np.float64), width, kpad)
    assert abs(L - L2) < 1e-14 and np.max(np.abs(g - g2)) < 1e-14
    L3, g3, _ = W.loss_grad_device(p, F, y, None, mu, sd, W.bf16, width, kpad)
    assert 0 < np.max(np.abs(g3 - g)) < 0.05 * np.max(np.abs(g))  # rounding changes it, a little


def test_bf16_rounding_ties_to_even():
    # 1 + 2^-8 lies halfway between BF16 neighbours 1 and 1 + 2^-7: ties to the even (1)
    assert W.bf16(1.0 + 2.0 ** -8) == 1.0
    assert W.bf16(1.0 + 3 * 2.0 ** -8) == 1.0 + 2.0 ** -6
    assert W.bf16(1.0 + 2.0 ** -8 + 2.0 ** -20) == 1.0 + 2.0 ** -7
    assert W.bf16(-3.0) == -3.0


def test_adam_two_steps_by_hand():
    # step 1: g = 0.5 -> m = 0.05, v = 2.5e-4, bias-corrected 0.5 and 0.25
    p, m1, m2 = W.adam(np.array([1.0]), np.array([0.5]), np.zeros(1), np.zeros(1), 1, 0.1)
    assert abs(m1[0] - 0.05) < 1e-15 and abs(m2[0] - 2.5e-4) < 1e-18
    assert abs(p[0] - 0.900000002) < 1e-12
    # step 2: g = -0.5 -> m = -0.005, v = 4.9975e-4; m_hat = -0.005/0.19, v_hat = 0.25
    p, m1, m2 = W.adam(p, np.array([-0.5]), m1, m2, 2, 0.1)
    assert abs(m1[0] + 0.005) < 1e-15 and abs(m2[0] - 4.9975e-4) < 1e-17
    assert abs(p[0] - (0.900000002 + 0.1 * (0.005 / 0.19) / (0.5 + 1e-8))) < 1e-12
    assert abs(p[0] - 0.9052631599) < 1e-9


def test_features_of_a_hand_worked_lp():
    # A = [[1, 0, 0.5], [0.5, 1, 0]], b = [2, 4], c = [-1, -2, -3]
    col_ptr = [0, 2, 3, 4]
    row_idx = [0, 1, 1, 0]
    val = [1.0, 0.5, 1.0, 0.5]
    F = W.features(2, 3, col_ptr, row_idx, val, [2.0, 4.0], [-1.0, -2.0, -3.0])
    r5 = np.sqrt(1.25)
    expect = np.array([
        [0, -0.5, 1.5, r5, 2, 1, 1 / 3, 2 / 3, 1 / 3, 0.625, 0.8, 1.5, 0, 2 / 3, 0.75, 5 / 3],
        [0, -1, 1, 1, 1, 1, 1, 4 / 3, 4 / 3, 0.25, 4, -1 / 3, 1 / 3, 1 / 3, 0.375, 2 / 3],
        [0, -1.5, 0.5, 0.5, 1, 0.5, 3, 4 / 3, 2, 0.25, 6, -2 / 3, 2 / 3, 0, 0.75, 5 / 3],
        [1, 0, 1.5, r5, 2, 1, 2 / 3, 0.75, 5 / 3, 0.625, 0.5, 0.5, 0, 0, 0, 0],
        [1, 0, 1.5, r5, 2, 1, 4 / 3, 0.375, 2 / 3, 0.3125, 0, 0, 0, 0, 0, 0],
    ])
    expect = np.hstack([expect, np.zeros((5, 1))])  # column 16, the family prior, is filled only at prediction
    assert np.allclose(F, expect, atol=1e-12), np.argwhere(~np.isclose(F, expect, atol=1e-12))


def test_select_top_m_breaks_ties_by_index():
    assert list(W.select_top_m([0.1, 0.5, 0.5, -1.0, 0.3], 3)) == [1, 2, 4]
    assert list(W.rank(np.array([2.0, 1.0, 2.0, 0.5]))) == [2, 1, 3, 0]


def random_packing(rng, m, n, dens=0.3):
    A = (rng.random((m, n)) < dens) * rng.uniform(0.1, 1.0, (m, n))
    for j in range(n):
        if not A[:, j].any():
            A[rng.integers(m), j] = 0.5
    return A


def test_factorization_restatement_inverts_nonsingular_bases():
    rng = np.random.default_rng(21)
    done = 0
    for _ in range(200):
        m, n = int(rng.integers(2, 25)), int(rng.integers(2, 40))
        A = random_packing(rng, m, n)
        basis = rng.permutation(n + m)[:m]
        Bm = np.hstack([A, np.eye(m)])[:, basis]
        if abs(np.linalg.det(Bm)) < 1e-6:
            continue
        b2, Binv, nrep = W.gj_inverse(A, basis, n, 1e-7)
        assert nrep == 0 and list(b2) == list(basis)
        assert np.abs(Binv - np.linalg.inv(Bm)).max() < 1e-9 * max(1.0, np.abs(np.linalg.inv(Bm)).max())
        done += 1
    assert done > 50


def test_slack_substitution_repairs_dependent_columns():
    rng = np.random.default_rng(22)
    for trial in range(100):
        m, n = int(rng.integers(3, 20)), int(rng.integers(4, 30))
        A = random_packing(rng, m, n)
        # make some structural columns copies or sums of others
        k = int(rng.integers(1, 3))
        for t in range(k):
            i, j = rng.choice(n, 2, replace=False)
            A[:, i] = A[:, j] * (1.0 if t == 0 else 2.0)
        start = rng.permutation(n + m)[:m]
        Bm = np.hstack([A, np.eye(m)])[:, start]
        rank = np.linalg.matrix_rank(Bm)
        b2, Binv, nrep = W.gj_inverse(A, start, n, 1e-7)
        B2 = np.hstack([A, np.eye(m)])[:, b2]
        assert len(set(b2)) == m
        assert np.abs(Binv @ B2 - np.eye(m)).max() < 1e-8
        # at least the rank deficiency; a substituted slack can make a later
        # structural column dependent in turn
        assert m - rank <= nrep <= int(np.sum(np.asarray(start) < n)), (trial, nrep, m - rank)
        # only structural columns are ever replaced, each by a slack
        for i in range(m):
            if b2[i] != start[i]:
                assert start[i] < n <= b2[i]
        if nrep:
            with pytest.raises(ValueError):
                W.gj_inverse(A, start, n, 0.0)


def test_family_prior_by_hand():
    """Three training bases over four variables: variable 0 basic in all, 1 in two,
    2 in one, 3 in none."""
    Y = np.array([[1, 1, 0, 0], [1, 0, 1, 0], [1, 1, 0, 0]], np.float64)
    np.testing.assert_allclose(W.family_prior(Y, loo=False), [1.0, 2 / 3, 1 / 3, 0.0])
    # leave one out: each training row sees the other two bases only
    np.testing.assert_allclose(W.family_prior(Y, loo=True),
                               [[1.0, 0.5, 0.5, 0.0], [1.0, 1.0, 0.0, 0.0], [1.0, 0.5, 0.5, 0.0]])
