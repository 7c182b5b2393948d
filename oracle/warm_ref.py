"""CPU checker (test infrastructure) for the learned warm start.

The reference has no learned component (its solver always starts from the
slack basis, src/cpu/sequentialDualSimplex.cpp:230-250), so this module is not
a restatement of reference code: it states, in plain numpy FP64, what the
device kernels of the warm-start extension compute, so that tests can compare
them:

* ``forward`` / ``loss_grad`` / ``loss_grad_device``: the per-variable MLP  x -> relu(W1 x + b1) ->
  relu(W2 . + b2) -> relu(W3 . + b3) -> w4 . + b4  on standardised inputs,
  the mean weighted BCE-with-logits loss and its gradient by backpropagation
  (pinned by central finite differences in tests/test_warm_oracle.py), and
  the same with BF16 rounding where the device stores values (pinned by its
  reduction to loss_grad without rounding);
* ``adam``: one Adam step (Kingma & Ba, 2015, Algorithm 1), pinned by a
  hand-computed step;
* ``features``: the 16 per-variable features of a packing LP (pinned by a
  hand-computed 2 x 3 LP);
* ``select_top_m``: the starting basis chosen from per-variable scores.

Only tests/, the smoke check and bench.py's CPU baseline may import it.
"""
import numpy as np

WIDTH = 256
KPAD = 32
NF = 17


def layout(width=WIDTH, kpad=KPAD):
    out, o = {}, 0
    for name, shape in (("W1", (width, kpad)), ("b1", (width,)), ("W2", (width, width)), ("b2", (width,)),
                        ("W3", (width, width)), ("b3", (width,)), ("w4", (width,)), ("b4", (1,))):
        out[name] = (o, shape)
        o += int(np.prod(shape))
    return out, o


def unpack(p, width=WIDTH, kpad=KPAD):
    lay, _ = layout(width, kpad)
    return {k: np.asarray(p[o:o + int(np.prod(s))], np.float64).reshape(s) for k, (o, s) in lay.items()}


def standardise(F, mu, sd, kpad=KPAD):
    F = np.asarray(F, np.float64)
    X = np.zeros((F.shape[0], kpad))
    X[:, :F.shape[1]] = (F - np.asarray(mu, np.float64)) / np.asarray(sd, np.float64)
    return X


def forward(p, F, mu, sd, width=WIDTH, kpad=KPAD):
    """Logits and the hidden activations of every row of F."""
    P = unpack(p, width, kpad)
    X = standardise(F, mu, sd, kpad)
    H1 = np.maximum(X @ P["W1"].T + P["b1"], 0.0)
    H2 = np.maximum(H1 @ P["W2"].T + P["b2"], 0.0)
    H3 = np.maximum(H2 @ P["W3"].T + P["b3"], 0.0)
    z = H3 @ P["w4"] + P["b4"][0]
    return z, (X, H1, H2, H3)


def bce(z, y, w):
    return w * (np.maximum(z, 0.0) - z * y + np.log1p(np.exp(-np.abs(z))))


def loss(p, F, y, w, mu, sd, width=WIDTH, kpad=KPAD):
    z, _ = forward(p, F, mu, sd, width, kpad)
    w = np.ones(len(y)) if w is None else np.asarray(w, np.float64)
    return float(np.mean(bce(z, np.asarray(y, np.float64), w)))


def loss_grad(p, F, y, w, mu, sd, width=WIDTH, kpad=KPAD):
    """Mean weighted BCE-with-logits loss and its gradient (flat, FP64)."""
    P = unpack(p, width, kpad)
    y = np.asarray(y, np.float64)
    S = len(y)
    w = np.ones(S) if w is None else np.asarray(w, np.float64)
    z, (X, H1, H2, H3) = forward(p, F, mu, sd, width, kpad)
    L = float(np.mean(bce(z, y, w)))
    dz = w * (1.0 / (1.0 + np.exp(-z)) - y) / S
    g = {}
    g["w4"] = H3.T @ dz
    g["b4"] = np.array([dz.sum()])
    G3 = np.outer(dz, P["w4"]) * (H3 > 0)
    g["W3"] = G3.T @ H2
    g["b3"] = G3.sum(0)
    G2 = (G3 @ P["W3"]) * (H2 > 0)
    g["W2"] = G2.T @ H1
    g["b2"] = G2.sum(0)
    G1 = (G2 @ P["W2"]) * (H1 > 0)
    g["W1"] = G1.T @ X
    g["b1"] = G1.sum(0)
    lay, total = layout(width, kpad)
    flat = np.zeros(total)
    for k, (o, s) in lay.items():
        flat[o:o + int(np.prod(s))] = g[k].ravel()
    return L, flat


def bf16(x):
    """Round to the nearest BF16 (ties to even); FP64 in, FP64 out."""
    a = np.ascontiguousarray(np.asarray(x, np.float64).astype(np.float32)).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32).astype(np.float64)


def loss_grad_device(p, F, y, w, mu, sd, rnd=bf16, width=WIDTH, kpad=KPAD):
    """loss_grad with the device's storage precision: standardised inputs, the
    three weight matrices, every hidden activation and every back-propagated
    pre-activation gradient rounded by `rnd` (BF16 on the device) where the
    device stores them; sums in FP64.  With rnd = identity it is loss_grad."""
    P = unpack(p, width, kpad)
    y = np.asarray(y, np.float64)
    S = len(y)
    w = np.ones(S) if w is None else np.asarray(w, np.float64)
    F32 = np.asarray(F, np.float32)
    X = np.zeros((S, kpad))
    X[:, :F32.shape[1]] = rnd((F32 - np.asarray(mu, np.float32)) / np.asarray(sd, np.float32))
    W1, W2, W3 = rnd(P["W1"]), rnd(P["W2"]), rnd(P["W3"])
    H1 = rnd(np.maximum(X @ W1.T + P["b1"], 0.0))
    H2 = rnd(np.maximum(H1 @ W2.T + P["b2"], 0.0))
    H3 = rnd(np.maximum(H2 @ W3.T + P["b3"], 0.0))
    z = H3 @ P["w4"] + P["b4"][0]
    L = float(np.mean(bce(z, y, w)))
    dz = w * (1.0 / (1.0 + np.exp(-z)) - y) / S
    g = {"w4": H3.T @ dz, "b4": np.array([dz.sum()])}
    G3 = rnd(np.outer(dz, P["w4"]) * (H3 > 0))
    g["W3"], g["b3"] = G3.T @ H2, G3.sum(0)
    G2 = rnd((G3 @ W3) * (H2 > 0))
    g["W2"], g["b2"] = G2.T @ H1, G2.sum(0)
    G1 = rnd((G2 @ W2) * (H1 > 0))
    g["W1"], g["b1"] = G1.T @ X, G1.sum(0)
    lay, total = layout(width, kpad)
    flat = np.zeros(total)
    for k, (o, s) in lay.items():
        flat[o:o + int(np.prod(s))] = g[k].ravel()
    return L, flat, z


def adam(p, g, m1, m2, t, lr, b1=0.9, b2=0.999, eps=1e-8):
    """One Adam step at step number t (1-based); returns (p, m1, m2)."""
    m1 = b1 * m1 + (1 - b1) * g
    m2 = b2 * m2 + (1 - b2) * g * g
    p = p - lr * (m1 / (1 - b1 ** t)) / (np.sqrt(m2 / (1 - b2 ** t)) + eps)
    return p, m1, m2


def features(m, n, col_ptr, row_idx, val, b, c):
    """(n+m) x 17 features of the packing LP min c^T x, A x <= b, x >= 0
    (structurals first, then the slacks), in FP64.  Column 16, the family prior,
    is 0 here: only a trained network's prediction step fills it (family_prior)."""
    col_ptr = np.asarray(col_ptr)
    row_idx = np.asarray(row_idx)
    val = np.asarray(val, np.float64)
    b = np.asarray(b, np.float64)
    c = np.asarray(c, np.float64)
    A = np.zeros((m, n))
    for j in range(n):
        for e in range(col_ptr[j], col_ptr[j + 1]):
            A[row_idx[e], j] = val[e]
    nzA = A != 0
    Cm = np.abs(c).mean() if np.abs(c).mean() > 0 else 1.0
    Bm = b.mean() if b.mean() > 0 else 1.0
    colsum = A.sum(0)
    rowsum = A.sum(1)
    rz = nzA.sum(1)
    ratio = np.where(colsum > 0, np.abs(c) / np.where(colsum > 0, colsum, 1.0), 0.0)
    y = np.array([ratio[nzA[i]].mean() if rz[i] > 0 else 0.0 for i in range(m)])
    t = rowsum / b
    F = np.zeros((n + m, NF))
    for j in range(n):
        rows = np.nonzero(nzA[:, j])[0]
        a = A[rows, j]
        u = np.min(b[rows] / a) if len(rows) else 0.0
        fj = np.sum(a / b[rows])
        dj = c[j] + np.sum(a * y[rows])
        bb = abs(c[j]) / fj if fj > 0 else 0.0
        F[j, :12] = [0.0, c[j] / Cm, colsum[j], np.sqrt(np.sum(a * a)), len(rows), a.max() if len(rows) else 0.0,
                     abs(c[j]) / colsum[j] / Cm if colsum[j] > 0 else 0.0, u / Bm, abs(c[j]) * u / (Cm * Bm), fj,
                     bb / Cm, dj / Cm]
        F[j, 14] = t[rows].max() if len(rows) else 0.0
        F[j, 15] = y[rows].max() / Cm if len(rows) else 0.0
    for i in range(m):
        cols = np.nonzero(nzA[i])[0]
        a = A[i, cols]
        F[n + i, :10] = [1.0, 0.0, rowsum[i], np.sqrt(np.sum(a * a)), len(cols), a.max() if len(cols) else 0.0,
                         b[i] / Bm, t[i], y[i] / Cm, np.sum(a * np.abs(c[cols])) / b[i] / Cm]
    key1 = F[:n, 10] * Cm
    key2 = F[:n, 11] * Cm
    F[:n, 12] = rank(key1) / n
    F[:n, 13] = rank(key2) / n
    F[n:, 10] = rank(t) / m
    F[n:, 11] = rank(y) / m
    return F


def family_prior(labels, loo):
    """Column 16 of the learned warm start's features for a family of T LPs of the
    same shape: labels is T x N (1 = basic in that LP's training basis); the
    fraction of the training bases in which each variable is basic, over the
    other T - 1 LPs for a training row (loo=True), over all T for a prediction."""
    Y = np.asarray(labels, np.float64)
    T = Y.shape[0]
    cnt = Y.sum(0)
    return (cnt[None, :] - Y) / (T - 1) if loo else cnt / T


def rank(v):
    """Number of entries smaller (ties: of lower index) than each entry."""
    order = np.lexsort((np.arange(len(v)), v))
    r = np.empty(len(v), np.int64)
    r[order] = np.arange(len(v))
    return r.astype(np.float64)


def select_top_m(scores, m):
    """Indices of the m highest scores, ties to the lower index, in that order."""
    s = np.asarray(scores)
    order = np.lexsort((np.arange(len(s)), -s))
    return order[:m]


def gj_inverse(A, basis, n, repair_tol=0.0):
    """The batched kernel's factorization of a starting basis, step for step:
    in-place Gauss-Jordan with partial pivoting over [A | I] (A dense m x n),
    slack columns taken first as pure row exchanges, structural columns in
    position order; a structural column whose best pivot is <= repair_tol is
    replaced by the slack of the original row at the lowest free position.
    Returns (basis after repair, B^{-1} of it, number of substitutions), or
    raises ValueError for a singular basis without repair."""
    A = np.asarray(A, np.float64)
    m = A.shape[0]
    Aext = np.hstack([A, np.eye(m)])
    basis = [int(j) for j in basis]
    rpos = list(range(m))
    at = list(range(m))
    piv = [0] * m
    for s in range(m):
        j = basis[s]
        if j < n:
            continue
        r = j - n
        cur = rpos[r]
        piv[s] = cur
        if cur != s:
            r2 = at[s]
            at[s], at[cur] = r, r2
            rpos[r], rpos[r2] = s, cur
    used = [basis[s] >= n for s in range(m)]
    B = np.zeros((m, m))
    for i in range(m):
        col = Aext[:, basis[i]]
        for r in np.nonzero(col)[0]:
            B[rpos[r], i] = col[r]
    nrep = 0
    for s in range(m):
        if basis[s] >= n:
            continue
        cand = [(abs(B[r, s]), -r) for r in range(m) if not used[r]]
        bv, nbi = max(cand)
        bi = -nbi
        if repair_tol > 0.0 and not bv > repair_tol:
            assert used.index(False) == s  # every position below s is in use
            basis[s] = n + at[s]
            B[:, s] = 0.0
            B[s, s] = 1.0
            bi, bv = s, 1.0
            nrep += 1
        if not bv > 0.0:
            raise ValueError("singular basis")
        piv[s] = bi
        used[s] = True
        at[s], at[bi] = at[bi], at[s]
        pv = B[bi, s]
        rowS, rowB = B[s].copy(), B[bi].copy()
        prow = np.where(np.arange(m) == s, 1.0, rowB) / pv
        if bi != s:
            keep = B[bi, s]
            B[bi] = rowS
            B[bi, s] = keep
        w = np.array([0.0 if i == s else (rowS[s] if i == bi else B[i, s]) for i in range(m)])
        for i in range(m):
            if i == s:
                B[i] = prow
            else:
                a = B[i].copy()
                a[s] = 0.0
                B[i] = a - w[i] * prow
    cmap = list(range(m))
    for slack_pass in (False, True):
        for s in range(m - 1, -1, -1):
            if (basis[s] >= n) != slack_pass or piv[s] == s:
                continue
            cmap[piv[s]], cmap[s] = cmap[s], cmap[piv[s]]
    return basis, B[:, cmap], nrep
