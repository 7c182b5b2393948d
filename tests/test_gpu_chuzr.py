"""The standalone CHUZR entry (ds_chuzr) on device arrays against the oracle's DSEPricing:
identical position and delta (both sides compute delta^2/beta in the same FP64 operations)."""
import ctypes as C

import numpy as np
import pytest

import _oracle as ora

pytestmark = pytest.mark.gpu


def pricing_inputs(m, seed):
    """Basic variables of every bound type, about a third out of bounds, with tied weights."""
    rng = np.random.default_rng(seed)
    bt = rng.choice([0, 1, 2, 3, 4], m, p=[0.1, 0.05, 0.25, 0.2, 0.4]).astype(np.int8)
    l = np.where((bt == 4) | (bt == 2) | (bt == 0), rng.uniform(-1, 1, m), -np.inf)
    u = np.where(bt == 0, l, np.where((bt == 3) | (bt == 2), np.nan_to_num(l, neginf=0) + rng.uniform(0, 2, m),
                                      np.inf))
    x = rng.uniform(-3, 3, m)
    x = np.where(rng.random(m) < 0.6, np.clip(x, np.nan_to_num(l, neginf=-1e300), np.nan_to_num(u, posinf=1e300)), x)
    beta = rng.choice([0.5, 1.0, 2.0], m) * rng.uniform(0.5, 2.0, m)
    if m > 8:
        x[3], l[3], u[3], bt[3], beta[3] = -2.0, 0.0, np.inf, 4, 1.0
        x[m - 2], l[m - 2], u[m - 2], bt[m - 2], beta[m - 2] = -2.0, 0.0, np.inf, 4, 1.0  # same weight as 3
    return bt, x, l, u, beta


def run_chuzr(fn_name, ctx, bt, x, l, u, beta):
    import torch
    from parallel_ds_amd._native import lib
    dev = torch.device("cuda:0")
    t = [torch.from_numpy(np.ascontiguousarray(a)).to(dev) for a in (bt, x, l, u, beta)]
    out_i = torch.full((1,), -7, dtype=torch.int32, device=dev)
    out_d = torch.zeros(1, dtype=torch.float64, device=dev)
    torch.cuda.synchronize()
    rc = getattr(lib(), fn_name)(ctx, len(x), *[C.c_void_p(a.data_ptr()) for a in t], C.c_void_p(out_i.data_ptr()),
                                 C.c_void_p(out_d.data_ptr()))
    return rc, int(out_i.cpu()[0]), float(out_d.cpu()[0])


@pytest.fixture(scope="module")
def ctx(gpu):
    from parallel_ds_amd._native import lib
    c = C.c_void_p()
    assert lib().ds_create(0, C.byref(c)) == 0
    yield c
    lib().ds_destroy(c)


@pytest.mark.parametrize("m,seed", [(1, 0), (9, 1), (777, 2), (300000, 3)])
def test_chuzr_matches_oracle(ctx, m, seed):
    bt, x, l, u, beta = pricing_inputs(m, seed)
    want = ora.dse_pricing(bt, x, l, u, beta)
    rc, p, d = run_chuzr("ds_chuzr", ctx, bt, x, l, u, beta)
    assert rc == 0
    assert (p, d) == want


def test_chuzr_feasible_and_rejects_bad_sizes(ctx):
    from parallel_ds_amd._native import lib
    m = 1000
    bt = np.full(m, 4, np.int8)
    x = np.ones(m)
    rc, p, d = run_chuzr("ds_chuzr", ctx, bt, x, np.zeros(m), np.full(m, np.inf), np.ones(m))
    assert rc == 0 and (p, d) == (-1, 0.0) == ora.dse_pricing(bt, x, np.zeros(m), np.full(m, np.inf), np.ones(m))
    assert lib().ds_chuzr(ctx, 0, None, None, None, None, None, None, None) == -1
