"""Batched small-LP solves over the C-ABI (one persistent launch, one workgroup per LP)."""
import ctypes as C

import numpy as np

from ._native import BatchParams, DsError, lib, ptr


def default_batch_params(**kw):
    p = BatchParams()
    lib().ds_batch_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class BatchSolver:
    """Packing LPs (min c^T x, A x <= b, x >= 0) solved together on the GPU."""

    def __init__(self, lps, device=0):
        L = lib()
        self._ctx = C.c_void_p()
        rc = L.ds_create(device, C.byref(self._ctx))
        if rc != 0:
            raise DsError(rc, "ds_create", "(no usable GPU?)")
        self.lps = list(lps)
        self.count = len(self.lps)
        m = np.array([lp.m for lp in self.lps], np.int32)
        n = np.array([lp.n for lp in self.lps], np.int32)
        nnz = np.array([lp.nnz for lp in self.lps], np.int64)
        nnz_off = np.zeros(self.count + 1, np.int64)
        nnz_off[1:] = np.cumsum(nnz)
        col_ptr = np.concatenate([lp.col_ptr for lp in self.lps]).astype(np.int32)
        row_idx = np.concatenate([lp.row_idx for lp in self.lps]).astype(np.int32)
        val = np.concatenate([lp.val for lp in self.lps])
        b = np.concatenate([lp.upper_range for lp in self.lps])
        c = np.concatenate([lp.costs for lp in self.lps])
        self.m, self.n = m, n
        self.moff = np.concatenate([[0], np.cumsum(m)]).astype(np.int64)
        self.voff = np.concatenate([[0], np.cumsum(m + n)]).astype(np.int64)
        self._check(L.ds_batch_load(self._ctx, self.count, ptr(m), ptr(n), ptr(nnz_off), ptr(col_ptr),
                                    ptr(row_idx), ptr(val), ptr(b), ptr(c)), "ds_batch_load")

    def _check(self, rc, where):
        if rc != 0:
            detail = lib().ds_last_error(self._ctx)
            raise DsError(rc, where, detail.decode() if detail else "")

    def close(self):
        if self._ctx:
            lib().ds_destroy(self._ctx)
            self._ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_warm(self, bases):
        """bases: list of m_k-length variable-id arrays, or None for the slack basis."""
        if bases is None:
            self._check(lib().ds_batch_set_warm(self._ctx, None), "ds_batch_set_warm")
            return
        flat = np.concatenate([np.asarray(b, np.int32) for b in bases]).astype(np.int32)
        self._check(lib().ds_batch_set_warm(self._ctx, ptr(flat)), "ds_batch_set_warm")

    def set_bases(self, bases):
        """Bases found by another solver (list of m_k-length variable-id arrays): the labels of
        the learned warm start, in place of a batched solve."""
        flat = np.concatenate([np.asarray(b, np.int32) for b in bases]).astype(np.int32)
        self._check(lib().ds_batch_set_bases(self._ctx, ptr(flat)), "ds_batch_set_bases")

    def solve(self, params=None):
        p = params if params is not None else default_batch_params()
        ms = C.c_double()
        self._check(lib().ds_batch_solve(self._ctx, C.byref(p), C.byref(ms)), "ds_batch_solve")
        return ms.value

    def replay(self):
        ms = C.c_double()
        self._check(lib().ds_batch_replay(self._ctx, C.byref(ms)), "ds_batch_replay")
        return ms.value

    def results(self, with_x=False):
        st = np.zeros(self.count, np.int32)
        obj = np.zeros(self.count)
        it = np.zeros(self.count, np.int32)
        basis = np.zeros(int(self.moff[-1]), np.int32)
        x = np.zeros(int(self.voff[-1])) if with_x else None
        self._check(lib().ds_batch_results(self._ctx, ptr(st), ptr(obj), ptr(it), ptr(basis), ptr(x)),
                    "ds_batch_results")
        bases = [basis[self.moff[k]:self.moff[k + 1]] for k in range(self.count)]
        xs = [x[self.voff[k]:self.voff[k + 1]] for k in range(self.count)] if with_x else None
        return st, obj, it, bases, xs

    def tune(self, sample=256, base=None):
        best = base if base is not None else default_batch_params()
        ms = C.c_double()
        tried = C.c_int()
        self._check(lib().ds_batch_tune(self._ctx, sample, C.byref(best), C.byref(ms), C.byref(tried)),
                    "ds_batch_tune")
        return best, ms.value, tried.value

    # ---- learned warm start -------------------------------------------------
    @property
    def nvars(self):
        return int(self.voff[-1])

    def features(self):
        """Per-variable features, (sum(n_k+m_k), 17) FP32, computed on the device (column 16, the
        family prior, is filled only when a trained network predicts)."""
        F = np.zeros((self.nvars, 17), np.float32)
        self._check(lib().ds_batch_features(self._ctx, ptr(F)), "ds_batch_features")
        return F

    def labels(self):
        """1.0 for the variables of each LP's last solved basis, else 0.0."""
        y = np.zeros(self.nvars, np.float32)
        self._check(lib().ds_batch_labels(self._ctx, ptr(y)), "ds_batch_labels")
        return y

    def warm_from_scores(self, scores):
        s = np.ascontiguousarray(scores, np.float32)
        if s.shape != (self.nvars,):
            raise ValueError("one score per variable expected")
        self._check(lib().ds_batch_warm_from_scores(self._ctx, ptr(s)), "ds_batch_warm_from_scores")

    def warm_start(self, net):
        """Starting bases from the network (features, inference, top-m on the device).
        Returns (wall ms of the three steps, inference kernel ms)."""
        ms, ims = C.c_double(), C.c_double()
        self._check(lib().ds_batch_warm_start(self._ctx, net._h, C.byref(ms), C.byref(ims)), "ds_batch_warm_start")
        return ms.value, ims.value

    def get_warm(self):
        b = np.zeros(int(self.moff[-1]), np.int32)
        self._check(lib().ds_batch_get_warm(self._ctx, ptr(b)), "ds_batch_get_warm")
        return [b[self.moff[k]:self.moff[k + 1]] for k in range(self.count)]

    def set_repair_tol(self, tol):
        self._check(lib().ds_batch_set_repair_tol(self._ctx, float(tol)), "ds_batch_set_repair_tol")

    def repairs(self):
        r = np.zeros(self.count, np.int32)
        self._check(lib().ds_batch_repairs(self._ctx, ptr(r)), "ds_batch_repairs")
        return r

