"""GPU tests of the batched solver (one persistent launch, one workgroup per LP)."""
import json
import os

import numpy as np
import pytest

import _oracle as ora
from parallel_ds_amd import lp as lpmod
from parallel_ds_amd.batch import BatchSolver, default_batch_params

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HI = json.load(open(os.path.join(G, "highs_objectives.json")))
CAPP = dict(max_iterations=50000)


@pytest.fixture(scope="module")
def family():
    return [lpmod.config3_instance(k) for k in range(48)]


def test_batch_matches_oracle_and_highs(gpu, family):
    with BatchSolver(family) as B:
        B.solve(default_batch_params(**CAPP))
        st, obj, it, bases, _ = B.results()
    for k, L in enumerate(family):
        r, x, b = ora.solve(L)
        assert st[k] == r.status == 0, (k, st[k], r.status)
        assert abs(obj[k] - r.Z) <= 1e-9 * abs(r.Z), (k, obj[k], r.Z)
        if k < len(HI["config3"]):
            g = HI["config3"][k]
            assert abs(obj[k] - g["objective"]) <= 1e-9 * abs(g["objective"])
            if g["nondegenerate"]:
                assert sorted(bases[k].tolist()) == g["basis"], k


@pytest.mark.parametrize("lds_m,block,occ", [(0, 256, 2), (96, 128, 4), (64, 512, 1)])
def test_batch_launch_shapes_agree(gpu, family, lds_m, block, occ):
    sub = family[:24]
    with BatchSolver(sub) as B:
        B.solve(default_batch_params(lds_inv_max_m=0, **CAPP))
        st0, obj0, _, _, _ = B.results()
        B.solve(default_batch_params(lds_inv_max_m=lds_m, block=block, wgs_per_cu=occ, **CAPP))
        st1, obj1, _, _, _ = B.results()
    assert np.array_equal(st0, st1)
    np.testing.assert_allclose(obj1, obj0, rtol=1e-9, atol=0)


def test_batch_graph_replays_match_eager(gpu, family):
    sub = family[:16]
    with BatchSolver(sub) as B:
        B.solve(default_batch_params(use_graph=0, **CAPP))
        ref = B.results(with_x=True)
        B.solve(default_batch_params(use_graph=1, **CAPP))
        for rep in range(3):
            # eager work and a synchronisation between replays
            B.solve(default_batch_params(use_graph=0, **CAPP))
            B.replay()
            st, obj, it, bases, xs = B.results(with_x=True)
            assert np.array_equal(st, ref[0]) and np.array_equal(obj, ref[1]) and np.array_equal(it, ref[2])
            for a, b in zip(xs, ref[4]):
                assert np.array_equal(a, b)


def test_batch_autotuner(gpu, family):
    with BatchSolver(family) as B:
        best, ms, tried = B.tune(sample=16, base=default_batch_params(**CAPP))
        assert tried >= 9 and ms > 0
        assert best.block in (128, 256, 512)
        B.solve(best)
        st, obj, _, _, _ = B.results()
    assert np.all(st == 0)


def test_batch_warm_start_from_optimal_basis(gpu, family):
    """Starting from the optimal basis, every LP finishes at once with the same optimum."""
    sub = family[:8]
    with BatchSolver(sub) as B:
        B.solve(default_batch_params(**CAPP))
        st, obj, it, bases, _ = B.results()
        B.set_warm(bases)
        B.solve(default_batch_params(**CAPP))
        st2, obj2, it2, _, _ = B.results()
    assert np.all(st2 == 0)
    np.testing.assert_allclose(obj2, obj, rtol=1e-9)
    assert np.all(it2 <= 2)


def test_batch_rejects_invalid_warm_basis(gpu, family):
    sub = family[:2]
    with BatchSolver(sub) as B:
        bad = [np.zeros(L.m, np.int32) for L in sub]  # repeated ids
        with pytest.raises(Exception):
            B.set_warm(bad)


def test_batch_cleanup_after_perturbation_matches_highs(gpu):
    """Learned-warm-start family instances against HiGHS.  Instance 20443 is perturbed on its
    slack-start path; with the default cleanup its objective still matches HiGHS to 1e-9."""
    gold = {g["seed"]: g["objective"] for g in HI["warm_family"]}
    base = lpmod.gen_kcol(256, 640, 5, 4)
    seeds = list(range(20000, 20016)) + [20443]
    lps = [lpmod.perturbed_instance(base, 0.1, s) for s in seeds]
    with BatchSolver(lps) as B:
        B.solve(default_batch_params(max_iterations=200000))
        st, obj, _, _, _ = B.results()
        B.solve(default_batch_params(max_iterations=200000, cleanup=0))
        st0, obj0, _, _, _ = B.results()
    for k, s in enumerate(seeds):
        assert st[k] == 0 and abs(obj[k] - gold[s]) <= 1e-9 * abs(gold[s]), (s, obj[k], gold[s])
        assert st0[k] == 0 and abs(obj0[k] - gold[s]) <= 1e-7 * abs(gold[s]), (s, obj0[k], gold[s])
