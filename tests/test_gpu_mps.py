"""A model read from MPS by the library (ds_load_mps) solves on the GPU exactly as the same arrays
loaded with ds_load_problem, with the objective constant removed as getSolution does."""
import os

import numpy as np
import pytest

import _oracle as ora
import mixed_lps
from parallel_ds_amd import DualSimplex, default_params
from parallel_ds_amd import lp as lpmod
from parallel_ds_amd._native import DS_OPTIMAL

pytestmark = pytest.mark.gpu


def run(src, prm):
    with DualSimplex(src, prm) as s:
        s.initDualSimplex()
        if s.presolve("minInfeas") == 0:
            s.solve()
        return s.getSolution(), s.result()


@pytest.mark.parametrize("which", ["boxed", "config0", "transport"])
def test_mps_load_matches_array_load(gpu, tmp_path, which):
    L = {"boxed": lambda: lpmod.boxed_instance(200, 400, 0.05, 1), "config0": lambda: lpmod.config0(2),
         "transport": lambda: lpmod.gen_transport(30, 60, 5)}[which]()
    L.offset = 12.5
    p = tmp_path / "m.mps"
    lpmod.write_mps(L, p)
    prm = default_params(max_iterations=100000)
    s1, r1 = run(str(p), prm)
    s2, r2 = run(L, prm)
    assert r1.status == r2.status == DS_OPTIMAL
    assert r1.iterations == r2.iterations and abs(s1.Z - s2.Z) <= 1e-12 * abs(s2.Z)
    r_o, _, _ = ora.solve(lpmod.read_mps(p), tie_rule=1, cleanup=1)
    assert abs(s1.Z - (r_o.Z - L.offset)) <= 1e-9 * abs(s1.Z)
    hs, hz = mixed_lps.highs(L)
    assert hs == 0 and abs(s1.Z - (hz - L.offset)) <= 1e-9 * abs(hz)


def test_hand_written_mps_model(gpu):
    """The hand-written model with every section (ranges, negative upper bounds, free and fixed
    columns, maximisation, an objective constant) against the oracle and HiGHS on the parsed arrays."""
    p = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "mps", "features.mps")
    R = lpmod.read_mps(p)
    hs, hz = mixed_lps.highs(R)
    sol, r = run(p, default_params(max_iterations=10000))
    r_o, _, _ = ora.solve(R, max_iterations=10000, tie_rule=1)
    assert r.status == r_o.status
    if hs == 0 and r.status == DS_OPTIMAL:
        assert abs(sol.Z - (hz - R.offset)) <= 1e-9 * max(1.0, abs(hz))
