"""Warm starts through primal simplex iterations (params.warm_primal) on the GPU: the starting
basis is made primal feasible by moving violated bounds, primal iterations remove the dual
infeasibilities, the bounds are restored and the dual simplex finishes.  Every solve must end at
HiGHS' optimum (golden fixtures) with a feasible certificate, whatever the starting basis."""
import json
import os

import numpy as np
import pytest

import lpcheck
from parallel_ds_amd import default_params, solve_lp
from parallel_ds_amd import lp as lpmod
from parallel_ds_amd._native import DS_OPTIMAL

pytestmark = pytest.mark.gpu

G = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
HI = json.load(open(os.path.join(G, "highs_objectives.json")))


def check(L, sol, basis, r, ref):
    assert r.status == DS_OPTIMAL, r.status
    assert abs(sol.Z - ref) <= 1e-9 * abs(ref), (sol.Z, ref)
    pinf, dinf, _ = lpcheck.certificate(L, sol.x, basis)
    assert pinf <= 1e-7 and dinf <= 1e-7, (pinf, dinf)


def test_primal_phase_on_the_family_held_out_lps(gpu):
    """Family m=256, n=640 (5 nnz/col, b and c x lognormal(0, 0.1)): the optimal basis of one
    training instance starts held-out instances 20000..20007 (HiGHS objectives in the fixture)."""
    base = lpmod.gen_kcol(256, 640, 5, 4)
    _, start, r0 = solve_lp(lpmod.perturbed_instance(base, 0.1, 10000), default_params())
    assert r0.status == DS_OPTIMAL
    gold = {g["seed"]: g for g in HI["warm_family"]}
    prim = []
    for s in range(20000, 20008):
        L = lpmod.perturbed_instance(base, 0.1, s)
        sol, basis, r = solve_lp(L, default_params(warm_primal=1, max_iterations=50000), start=start)
        print("seed %d: %d iterations (%d primal, %d bounds moved), restarts %d"
              % (s, r.iterations, r.primal_iterations, r.shifted_bounds, r.restarts))
        check(L, sol, basis, r, gold[s]["objective"])
        prim.append(r.primal_iterations)
        assert r.phase1_iterations >= r.primal_iterations
    assert max(prim) > 0


def test_primal_phase_at_the_family_size(gpu):
    """The family's base LP at m=2000, n=5000 started from a perturbed instance's optimal basis."""
    base = lpmod.gen_kcol(2000, 5000, 5, 4)
    _, start, r0 = solve_lp(lpmod.perturbed_instance(base, 0.1, 10001), default_params(time_limit_ms=60000))
    assert r0.status == DS_OPTIMAL
    sol, basis, r = solve_lp(base, default_params(warm_primal=1, time_limit_ms=60000), start=start)
    print("%d iterations (%d primal, %d bounds moved)" % (r.iterations, r.primal_iterations, r.shifted_bounds))
    check(base, sol, basis, r, HI["kcol_m2000_n5000_k5_seed4"]["objective"])
    assert r.primal_iterations > 0 and r.shifted_bounds > 0


def test_primal_phase_from_the_slack_basis_is_the_primal_simplex(gpu):
    """From the slack basis of a packing LP (primal feasible, no bound to move) the primal
    iterations alone reach the optimum; the dual simplex then has nothing left to do."""
    L = lpmod.config0(0)
    start = np.arange(L.n, L.n + L.m, dtype=np.int32)
    sol, basis, r = solve_lp(L, default_params(warm_primal=1, max_iterations=20000), start=start)
    check(L, sol, basis, r, HI["config0"][0]["objective"])
    assert r.shifted_bounds == 0 and r.primal_iterations > 0
    assert r.iterations - r.primal_iterations <= 2


def test_models_with_boxed_variables_skip_the_primal_phase(gpu):
    L = lpmod.boxed_instance(120, 240, 0.1, 3)
    _, b0, r0 = solve_lp(L, default_params(max_iterations=20000))
    assert r0.status == DS_OPTIMAL
    sol, basis, r = solve_lp(L, default_params(warm_primal=1, max_iterations=20000), start=b0)
    assert r.status == DS_OPTIMAL and r.primal_iterations == 0
    assert abs(sol.Z - r0.Z) <= 1e-9 * max(1.0, abs(r0.Z))


def test_split_launches_give_the_same_solves(gpu):
    """Below a size limit, stages share launches: both ratio-test passes with BTRAN's set-up
    (m <= 8,192), and the entering column's FTRAN, its etas, tau's etas and the dual update
    (n <= 16,384).  params.split_launches runs them in the launches larger LPs take; the solves
    must be the same bit for bit: status, iterations, objective and final basis.  Warm starts with
    primal iterations on the m = 256 family, and slack starts with both ratio tests."""
    base = lpmod.gen_kcol(256, 640, 5, 4)
    _, start, r0 = solve_lp(lpmod.perturbed_instance(base, 0.1, 10000), default_params())
    assert r0.status == DS_OPTIMAL
    cases = [(lpmod.perturbed_instance(base, 0.1, s), start, dict(warm_primal=1)) for s in (20000, 20003)]
    cases += [(lpmod.config0(1), None, dict(ratio_test=rt)) for rt in (0, 1)]
    for L, st, kw in cases:
        outs = []
        for split in (0, 1):
            sol, basis, r = solve_lp(L, default_params(max_iterations=50000, split_launches=split, **kw), start=st)
            outs.append((r.status, r.iterations, r.primal_iterations, sol.Z, sorted(basis.tolist())))
        print(kw, "iterations", outs[0][1], outs[1][1])
        assert outs[0][0] == DS_OPTIMAL
        assert outs[0] == outs[1]
