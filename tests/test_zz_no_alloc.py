"""The allocation-free CHUZR entry (ds_chuzr_noalloc) against ds_chuzr on the same device
arrays: identical position and delta on two calls in a row, the second on fewer variables
(the static scratch keeps the first call's partials and counter), and the size limit."""
import ctypes as C

import numpy as np
import pytest

import _oracle as ora
from test_gpu_chuzr import pricing_inputs, run_chuzr

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx(gpu):
    from parallel_ds_amd._native import lib
    c = C.c_void_p()
    assert lib().ds_create(0, C.byref(c)) == 0
    yield c
    lib().ds_destroy(c)


def test_noalloc_matches_allocating_entry(ctx):
    for m, seed in ((300000, 11), (777, 12)):
        bt, x, l, u, beta = pricing_inputs(m, seed)
        rc0, p0, d0 = run_chuzr("ds_chuzr", ctx, bt, x, l, u, beta)
        rc1, p1, d1 = run_chuzr("ds_chuzr_noalloc", ctx, bt, x, l, u, beta)
        assert rc0 == 0 and rc1 == 0
        assert (p1, d1) == (p0, d0) == ora.dse_pricing(bt, x, l, u, beta)


def test_noalloc_rejects_sizes_above_its_limit(ctx):
    from parallel_ds_amd._native import lib
    dummy = C.c_void_p(1)
    rc = lib().ds_chuzr_noalloc(ctx, (1 << 28) + 1, dummy, dummy, dummy, dummy, dummy, dummy, dummy)
    assert rc == -1
    assert b"exceeds" in lib().ds_last_error(ctx)
