"""GPU tests of the learned warm start: the BF16 MFMA network against the FP64 CPU
checker (forward, gradient, Adam), the feature / label / selection kernels, the
slack-substitution repair, and the gain of a trained network over the slack start
and over an untrained network on the same LPs."""
import os
import sys

import numpy as np
import pytest

from parallel_ds_amd import lp as lpmod
from parallel_ds_amd.batch import BatchSolver, default_batch_params
from parallel_ds_amd.warm import Mlp

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, os.path.join(ROOT, "oracle"))
import warm_ref as W  # noqa: E402

pytestmark = pytest.mark.gpu
CAPP = dict(max_iterations=100000)


def bf16(x):
    """Round to the nearest BF16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    r = ((a + 0x7FFF + ((a >> 16) & 1)) >> 16) << 16
    return (r & 0xFFFFFFFF).astype(np.uint32).view(np.float32)


def rounded_params(p):
    """The parameter vector with the three weight matrices rounded as the device uses them."""
    q = np.array(p, np.float64)
    lay, _ = W.layout()
    for k in ("W1", "W2", "W3"):
        o, s = lay[k]
        n = int(np.prod(s))
        q[o:o + n] = bf16(p[o:o + n])
    return q


def device_like_forward(p, F, mu, sd):
    """FP64 forward on the device's rounded inputs and weights (activations unrounded)."""
    X = bf16((np.asarray(F, np.float32) - np.asarray(mu, np.float32)) / np.asarray(sd, np.float32))
    return W.forward(rounded_params(p), X, np.zeros(F.shape[1]), np.ones(F.shape[1]))


def sample(rng, S):
    nf = W.NF
    F = rng.normal(0.0, 1.0, (S, nf)).astype(np.float32) * rng.uniform(0.5, 3.0, nf).astype(np.float32)
    mu = rng.normal(0.0, 0.3, nf).astype(np.float32)
    sd = rng.uniform(0.7, 2.5, nf).astype(np.float32)
    y = (rng.random(S) < 0.3).astype(np.float32)
    return F, mu, sd, y


def test_mlp_forward_matches_fp64(gpu):
    rng = np.random.default_rng(11)
    F, mu, sd, _ = sample(rng, 333)
    with Mlp(seed=5) as net:
        net.set_norm(mu, sd)
        p = net.get_params()
        z = net.infer(F)
    zr, _ = device_like_forward(p, F, mu, sd)
    z64, _ = W.forward(p, F, mu, sd)
    scale = np.max(np.abs(z64)) + 1.0
    # BF16 activations between layers: ~2^-8 relative per layer
    assert np.max(np.abs(z - zr)) <= 2e-2 * scale, np.max(np.abs(z - zr))
    assert np.max(np.abs(z - z64)) <= 5e-2 * scale, np.max(np.abs(z - z64))


@pytest.mark.parametrize("weighted", [False, True])
def test_mlp_gradient_matches_fp64(gpu, weighted):
    """Device gradient against (a) the checker at the device's storage precision
    (BF16 where the device stores activations and gradients): tight; (b) plain
    FP64: within what BF16 storage costs (a few per cent of the gradient norm)."""
    rng = np.random.default_rng(12 + weighted)
    F, mu, sd, y = sample(rng, 700)
    w = rng.uniform(0.5, 2.0, len(y)).astype(np.float32) if weighted else None
    with Mlp(seed=6) as net:
        net.set_norm(mu, sd)
        p = net.get_params()
        loss, g = net.grad(F, y, w)
    Ld, gd, _ = W.loss_grad_device(p, F, y, w, mu, sd)
    L64, g64 = W.loss_grad(p, F, y, w, mu, sd)
    assert abs(loss - Ld) <= 1e-4 * abs(Ld), (loss, Ld)
    assert abs(loss - L64) <= 1e-2 * abs(L64), (loss, L64)
    lay, _ = W.layout()
    report = {}
    for k, (o, s) in lay.items():
        n = int(np.prod(s))
        a = g[o:o + n].astype(np.float64)
        if k == "W1":  # inputs beyond the features are zero: no gradient there
            assert np.all(a.reshape(s)[:, W.NF:] == 0.0)
        for ref, name in ((gd[o:o + n], "dev"), (g64[o:o + n], "fp64")):
            nb = np.linalg.norm(ref)
            report[(k, name)] = (np.linalg.norm(a - ref) / nb, float(a @ ref) / (np.linalg.norm(a) * nb))
    print({k: ("%.2e" % v[0], "%.6f" % v[1]) for k, v in report.items()})
    for (k, name), (rel, cos) in report.items():
        if name == "dev":
            assert rel < 1e-2 and cos > 0.99995, (k, rel, cos)
        else:
            assert rel < 6e-2 and cos > 0.998, (k, rel, cos)


def test_adam_step_matches_fp64(gpu):
    rng = np.random.default_rng(13)
    F, mu, sd, y = sample(rng, 256)
    with Mlp(seed=7) as net:
        net.set_norm(mu, sd)
        p0 = net.get_params().astype(np.float64)
        _, g = net.grad(F, y)
        net.adam_step(1e-3)
        p1 = net.get_params().astype(np.float64)
    pr, _, _ = W.adam(p0, g.astype(np.float64), np.zeros_like(p0), np.zeros_like(p0), 1, 1e-3)
    np.testing.assert_allclose(p1, pr, rtol=2e-6, atol=2e-7)
    assert np.max(np.abs(p1 - p0)) > 5e-4  # the step moved the weights


def batch_family():
    lps = [lpmod.config3_instance(k) for k in range(5)]
    base = lpmod.gen_kcol(60, 150, 5, 4)
    lps += [lpmod.perturbed_instance(base, 0.1, s) for s in range(3)]
    return lps


def test_batch_features_match_oracle(gpu):
    lps = batch_family()
    with BatchSolver(lps) as B:
        F = B.features()
    for k, L in enumerate(lps):
        Fk = F[B.voff[k]:B.voff[k + 1]].astype(np.float64)
        R = W.features(L.m, L.n, L.col_ptr, L.row_idx, L.val, L.upper_range, L.costs)
        n, m = L.n, L.m
        dense = np.ones_like(R, bool)
        dense[:n, 12:14] = False
        dense[n:, 10:12] = False
        np.testing.assert_allclose(Fk[dense], R[dense], rtol=2e-5, atol=2e-5)
        # ranks from FP32 keys may swap near-equal neighbours
        assert np.max(np.abs(Fk[:n, 12:14] - R[:n, 12:14])) <= 2.0 / n + 1e-6
        assert np.max(np.abs(Fk[n:, 10:12] - R[n:, 10:12])) <= 2.0 / m + 1e-6


def test_batch_labels_and_top_m_selection(gpu):
    lps = batch_family()
    rng = np.random.default_rng(14)
    with BatchSolver(lps) as B:
        B.solve(default_batch_params(**CAPP))
        st, _, _, bases, _ = B.results()
        assert np.all(st == 0)
        y = B.labels()
        scores = rng.normal(0, 1, B.nvars).astype(np.float32)
        scores[B.voff[0]:B.voff[0] + 5] = 7.0  # ties resolved by index
        B.warm_from_scores(scores)
        warm = B.get_warm()
    for k, L in enumerate(lps):
        yk = y[B.voff[k]:B.voff[k + 1]]
        assert sorted(np.nonzero(yk)[0].tolist()) == sorted(bases[k].tolist())
        sk = scores[B.voff[k]:B.voff[k + 1]]
        assert warm[k].tolist() == W.select_top_m(sk, L.m).tolist()


def singular_start_lp():
    """A packing LP whose columns 0 and 1 are identical, and a start that holds both."""
    L = lpmod.config3_instance(2)
    cp, ri, v = L.col_ptr.copy(), L.row_idx.copy(), L.val.copy()
    c0 = slice(cp[0], cp[1])
    n0 = cp[1] - cp[0]
    rest_ri, rest_v = ri[cp[2]:], v[cp[2]:]
    ri2 = np.concatenate([ri[c0], ri[c0], rest_ri]).astype(np.int32)
    v2 = np.concatenate([v[c0], v[c0], rest_v])
    cp2 = np.concatenate([[0, n0], cp[2:] - cp[2] + 2 * n0]).astype(np.int32)
    L2 = lpmod.packing_lp(L.m, L.n, cp2, ri2, v2, L.upper_range, L.costs, "dup-col")
    rows = sorted(set(ri[c0].tolist()))[:2]
    start = [0, 1] + [L.n + i for i in range(L.m) if i not in rows]
    return L2, np.array(start, np.int32)


def test_repair_substitutes_slacks_for_a_singular_start(gpu):
    L, start = singular_start_lp()
    with BatchSolver([L, L]) as B:
        B.solve(default_batch_params(**CAPP))
        st0, obj0, it0, _, _ = B.results()
        B.set_warm([start, start])
        B.solve(default_batch_params(use_graph=0, **CAPP))
        st1, obj1, _, _, _ = B.results()
        rep = B.repairs()
        B.set_repair_tol(0.0)
        B.solve(default_batch_params(use_graph=0, **CAPP))
        st2, _, _, _, _ = B.results()
    assert np.all(st0 == 0) and np.all(st1 == 0), (st0, st1)
    np.testing.assert_allclose(obj1, obj0, rtol=1e-9, atol=0)
    assert np.all(rep >= 1), rep
    assert np.all(st2 == 5), st2  # without repair the start is rejected as singular


def test_trained_network_gains_over_slack_start(gpu):
    """Config-1 recipe at m=64, n=160, 5 nonzeros per column, b and c perturbed by
    lognormal(0, 0.1): train on 768 instances' GPU-solved bases, evaluate on 64 others."""
    base = lpmod.gen_kcol(64, 160, 5, 11)
    train = [lpmod.perturbed_instance(base, 0.1, s) for s in range(768)]
    held = [lpmod.perturbed_instance(base, 0.1, 5000 + s) for s in range(64)]
    prm = default_batch_params(use_graph=0, **CAPP)

    def say(msg):
        print(msg, flush=True)

    with BatchSolver(train) as T, Mlp(seed=21) as net, Mlp(seed=21) as raw:
        say("solve 768 training LPs from the slack basis")
        T.solve(prm)
        st, _, _, _, _ = T.results()
        assert np.all(st == 0)
        say("train")
        losses, _ = net.fit(T, epochs=6, minibatch=4096, lr=1e-3, seed=3)
        say("losses %s" % np.round(losses, 4))
        # the family prior (feature 16) makes the first epoch's loss already low
        assert losses[-1] < 0.9 * losses[0], losses
        # feature 16: fraction of the training bases in which each variable is basic
        labs = T.labels().reshape(len(train), base.m + base.n)
        np.testing.assert_allclose(net.prior(), W.family_prior(labs, loo=False), rtol=0, atol=1e-6)
        assert raw.prior().size == 0
        mu, sd = net.get_norm()
        raw.set_norm(mu, sd)
        with BatchSolver(held) as E:
            say("held-out: slack start")
            E.solve(prm)
            st_s, obj_s, it_s, _, _ = E.results()
            say("held-out: trained warm start")
            E.warm_start(net)
            E.solve(prm)
            st_w, obj_w, it_w, _, _ = E.results()
            rep_w = E.repairs()
            say("held-out: untrained warm start")
            E.warm_start(raw)
            E.solve(prm)
            st_u, obj_u, it_u, _, _ = E.results()
            say("repairs (trained) %s" % rep_w[:16])
    print("slack %.1f  untrained %.1f  trained %.1f mean iterations" % (it_s.mean(), it_u.mean(), it_w.mean()))
    assert np.array_equal(st_s, st_w) and np.all(st_s == 0)
    np.testing.assert_allclose(obj_w, obj_s, rtol=1e-9, atol=0)
    np.testing.assert_allclose(obj_u, obj_s, rtol=1e-9, atol=0)
    assert it_w.mean() < 0.5 * it_s.mean()
    assert it_w.mean() < it_u.mean()
