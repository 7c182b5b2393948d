"""The library's MPS reader (the model Problem::readMps takes from Clp, problem.tpp:331-415):
hand-derived known answers for every section and bound type, exact round trips of the seeded
workloads through the MPS writer, and a parsed model solved by the oracle against HiGHS."""
import os

import numpy as np
import pytest

import _oracle as ora
import mixed_lps
from parallel_ds_amd import lp as lpmod
from parallel_ds_amd._native import DS_BOXED, DS_FIXED, DS_FREE, DS_LOWER, DS_UPPER

HERE = os.path.dirname(os.path.abspath(__file__))
INF = np.inf


def test_known_answers_of_every_section():
    L = lpmod.read_mps(os.path.join(HERE, "golden", "mps", "features.mps"))
    assert (L.m, L.n) == (5, 7)  # the second N row (SPARE) is dropped
    # OBJSENSE MAX: costs negated; RHS on the objective row: constant = -(-7.5)
    np.testing.assert_array_equal(L.costs, [-1.0, -2.0, 1.0, 0.0, -0.5, 0.0, -3.0])
    assert L.offset == 7.5
    # rows: LIM1 L 4 range 2.5 -> [1.5, 4]; LIM2 G 1 range 1.5 -> [1, 2.5]; MYEQN E 7;
    # EQRNGP E 2 range +4 -> [2, 6]; EQRNGN E 3 range -4 -> [-1, 3]
    np.testing.assert_array_equal(L.lower_range, [1.5, 1.0, 7.0, 2.0, -1.0])
    np.testing.assert_array_equal(L.upper_range, [4.0, 2.5, 7.0, 6.0, 3.0])
    np.testing.assert_array_equal(L.range_type, [DS_BOXED, DS_BOXED, DS_FIXED, DS_BOXED, DS_BOXED])
    # bounds: UP 4; [-1, 1]; MI -> (-inf, +inf); UP -2 with lower 0 -> (-inf, -2]; FX 3.5; FR; BV [0, 1]
    np.testing.assert_array_equal(L.lower_bound, [0.0, -1.0, -INF, -INF, 3.5, -INF, 0.0])
    np.testing.assert_array_equal(L.upper_bound, [4.0, 1.0, INF, -2.0, 3.5, INF, 1.0])
    np.testing.assert_array_equal(L.bound_type, [DS_BOXED, DS_BOXED, DS_FREE, DS_UPPER, DS_FIXED, DS_FREE,
                                                 DS_BOXED])
    # matrix (rows LIM1=0, LIM2=1, MYEQN=2, EQRNGP=3, EQRNGN=4); X7 has only an objective entry
    A = np.zeros((5, 7))
    for j in range(7):
        for e in range(L.col_ptr[j], L.col_ptr[j + 1]):
            A[L.row_idx[e], j] = L.val[e]
    want = np.array([[1, 1, 0, 0, 0, .25, 0], [1, 0, 0, 1, 0, 0, 0], [0, -1, 1, 0, 0, 0, 0],
                     [0, 0, 1, 0, 2, 0, 0], [0, 0, 0, 1, 0, 0, 0]], float)
    np.testing.assert_array_equal(A, want)


def test_reader_rejects_malformed_files(tmp_path):
    from parallel_ds_amd._native import DsError
    bad = tmp_path / "bad.mps"
    bad.write_text("NAME X\nROWS\n N OBJ\n L R1\nCOLUMNS\n X1 R9 1.0\nENDATA\n")
    with pytest.raises(DsError) as e:
        lpmod.read_mps(bad)
    assert "unknown row" in str(e.value)
    bad.write_text("NAME X\nROWS\n N OBJ\n L R1\n L R2\nCOLUMNS\n X1 R1 1.0\nRHS\n RHS R1 1\nENDATA\n")
    with pytest.raises(DsError) as e:  # 2 rows, 1 column: checkConstraints' rejection
        lpmod.read_mps(bad)
    assert "constraint size" in str(e.value)
    bad.write_text("NAME X\nROWS\n N OBJ\n")
    with pytest.raises(DsError):
        lpmod.read_mps(bad)


def same(a, b):
    """Exact equality; a ranged row's second bound is rebuilt from RHS and RANGES (lo = up - r), so it
    may differ in the last bit."""
    for f in ("col_ptr", "row_idx", "val", "costs", "lower_bound", "upper_bound", "bound_type", "upper_range",
              "range_type"):
        x, y = np.asarray(getattr(a, f)), np.asarray(getattr(b, f))
        assert x.shape == y.shape and np.array_equal(x, y), f
    x, y = np.asarray(a.lower_range), np.asarray(b.lower_range)
    ranged = np.asarray(b.range_type) == DS_BOXED
    assert np.array_equal(x[~ranged], y[~ranged])
    np.testing.assert_allclose(x[ranged], y[ranged], rtol=0, atol=4e-16 * max(1.0, float(np.abs(b.upper_range).max())))
    assert a.offset == b.offset


@pytest.mark.parametrize("which", ["config0", "boxed", "transport", "mixed"])
def test_round_trip_of_seeded_workloads(tmp_path, which):
    if which == "config0":
        lps = [lpmod.config0(3)]
    elif which == "boxed":
        lps = [lpmod.boxed_instance(50, 100, 0.1, 1)]
    elif which == "transport":
        lps = [lpmod.gen_transport(10, 20, 2)]
    else:
        lps = [mixed_lps.mixed_lp(s, boxed=True) for s in range(40)]
        lps = [L for L in lps if L.m <= L.n]
    for k, L in enumerate(lps):
        L.offset = 0.25 * k
        p = tmp_path / ("%s%d.mps" % (which, k))
        lpmod.write_mps(L, p)
        same(lpmod.read_mps(p), L)


def test_parsed_model_solves_to_highs_optimum(tmp_path):
    """The oracle on a boxed instance read back from MPS: HiGHS' objective (plus the constant)."""
    L = lpmod.boxed_instance(50, 100, 0.1, 2)
    L.offset = -3.0
    p = tmp_path / "b.mps"
    lpmod.write_mps(L, p)
    R = lpmod.read_mps(p)
    res, x, basis = ora.solve(R, tie_rule=1)
    hs, hz = mixed_lps.highs(R)
    assert res.status == 0 and hs == 0
    assert abs((res.Z - R.offset) - (hz - L.offset)) <= 1e-9 * abs(hz)
