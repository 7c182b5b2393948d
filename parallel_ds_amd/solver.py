"""Host-side mirror of the reference's solver API over the C-ABI.

    solver = DualSimplex(lp)           # Problem(...) + transformToComputeForm()
    solver.initDualSimplex()           # SequentialDualSimplex::initDualSimplex
    solver.presolve("minInfeas")       # phase 1 (baseDualSimplex.tpp:5-30)
    solver.solve()                     # phase 2 + postProcess (baseDualSimplex.tpp:123-180)
    sol = solver.getSolution()         # Problem::getSolution -> LPsolution

The arithmetic runs in libpds.so on the GPU; there is no CPU path here.
"""
import ctypes as C
from dataclasses import dataclass

import numpy as np

from ._native import DsError, Params, Result, lib, ptr


@dataclass
class LPsolution:
    """The reference's LPsolution record (src/common/LPsolution.hpp:22-44)."""
    Z: float
    solved: bool
    x: np.ndarray
    message: str
    iterations: int
    status: int


def default_params(**kw):
    p = Params()
    lib().ds_default_params(C.byref(p))
    for k, v in kw.items():
        setattr(p, k, v)
    return p


class _Size:
    def __init__(self, m, n):
        self.m, self.n = m, n


class DualSimplex:
    """`lp` is an LP of parallel_ds_amd.lp (the 8 arrays of Problem's constructor) or the path of
    an MPS file (Problem::readMps, read by the library's parser)."""

    def __init__(self, lp, params=None, device=0):
        L = lib()
        self._ctx = C.c_void_p()
        rc = L.ds_create(device, C.byref(self._ctx))
        if rc != 0:
            raise DsError(rc, "ds_create", "(no usable GPU?)")
        if isinstance(lp, (str, bytes)) or hasattr(lp, "__fspath__"):
            m, n = C.c_int(), C.c_int()
            self._check(L.ds_load_mps(self._ctx, str(lp).encode(), C.byref(m), C.byref(n)), "ds_load_mps")
            self.lp = _Size(m.value, n.value)
        else:
            self.lp = lp
            self._check(L.ds_load_problem(self._ctx, lp.m, lp.n, *[ptr(a) for a in lp.arrays()]), "ds_load_problem")
            if getattr(lp, "offset", 0.0):
                self._check(L.ds_set_offset(self._ctx, float(lp.offset)), "ds_set_offset")
        self.params = params if params is not None else default_params()
        self._check(L.ds_set_params(self._ctx, C.byref(self.params)), "ds_set_params")

    def _check(self, rc, where):
        if rc != 0:
            detail = lib().ds_last_error(self._ctx) if self._ctx else b""
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

    # reference call sequence (src/main.cpp:57-68)
    def initDualSimplex(self, start=None):
        """From the slack basis, or from `start` (m variable ids), whose dependent columns the
        first factorization replaces by slacks (Result.repairs)."""
        if start is None:
            self._check(lib().ds_init(self._ctx), "ds_init")
        else:
            b = np.ascontiguousarray(start, np.int32)
            if b.shape != (self.lp.m,):
                raise ValueError("a starting basis lists m variables")
            self._check(lib().ds_init_basis(self._ctx, ptr(b)), "ds_init_basis")

    def presolve(self, method="minInfeas"):
        st = C.c_int()
        self._check(lib().ds_presolve(self._ctx, method.encode(), C.byref(st)), "ds_presolve")
        return st.value

    def solve(self):
        self._check(lib().ds_solve(self._ctx), "ds_solve")

    def result(self):
        r = Result()
        self._check(lib().ds_get_result(self._ctx, C.byref(r)), "ds_get_result")
        return r

    def getSolution(self):
        Z = C.c_double()
        solved = C.c_int()
        msg = C.create_string_buffer(64)
        x = np.zeros(self.lp.n + self.lp.m)
        self._check(lib().ds_get_solution(self._ctx, C.byref(Z), C.byref(solved), msg, 64, ptr(x)),
                    "ds_get_solution")
        r = self.result()
        return LPsolution(Z.value, bool(solved.value), x, msg.value.decode(), r.iterations, r.status)

    def basis(self):
        b = np.zeros(self.lp.m, np.int32)
        self._check(lib().ds_get_basis(self._ctx, ptr(b)), "ds_get_basis")
        return b

    def nonbasis(self):
        b = np.zeros(self.lp.n, np.int32)
        self._check(lib().ds_get_nonbasis(self._ctx, ptr(b)), "ds_get_nonbasis")
        return b

    # hook-level entries
    def set_basis(self, basis):
        b = np.ascontiguousarray(basis, np.int32)
        self._check(lib().ds_set_basis(self._ctx, ptr(b)), "ds_set_basis")

    def refactor(self):
        self._check(lib().ds_refactor(self._ctx), "ds_refactor")

    def btran(self, p_idx):
        rho = np.zeros(self.lp.m)
        self._check(lib().ds_btran(self._ctx, p_idx, ptr(rho)), "ds_btran")
        return rho

    def ftran(self, b):
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(self.lp.m)
        self._check(lib().ds_ftran(self._ctx, ptr(b), ptr(x)), "ds_ftran")
        return x

    def ftran_column(self, j):
        x = np.zeros(self.lp.m)
        self._check(lib().ds_ftran_column(self._ctx, j, ptr(x)), "ds_ftran_column")
        return x

    def pivot_row(self, rho):
        rho = np.ascontiguousarray(rho, np.float64)
        a = np.zeros(self.lp.n)
        self._check(lib().ds_pivot_row(self._ctx, ptr(rho), ptr(a)), "ds_pivot_row")
        return a

    def time_pivot_row(self, rho, reps=100):
        rho = np.ascontiguousarray(rho, np.float64)
        ms = C.c_double()
        by = C.c_double()
        self._check(lib().ds_time_pivot_row(self._ctx, ptr(rho), reps, C.byref(ms), C.byref(by)),
                    "ds_time_pivot_row")
        return ms.value, by.value

    def time_refactor(self, reps=3):
        """Factorizations of the current basis: mean ms, trailing-update (FP64 MFMA) ms and flops, k."""
        r, g, f, k = C.c_double(), C.c_double(), C.c_double(), C.c_int()
        self._check(lib().ds_time_refactor(self._ctx, reps, C.byref(r), C.byref(g), C.byref(f), C.byref(k)),
                    "ds_time_refactor")
        return {"refactor_ms": r.value, "gemm_ms": g.value, "gemm_flops": f.value, "k": k.value}

    def time_pivot_row_ext(self, rho, reps=100, flush_bytes=640 << 20):
        """(warm ms, cold ms, algorithmic bytes) per SpMV launch; cold flushes caches before each launch."""
        rho = np.ascontiguousarray(rho, np.float64)
        w, c, by = C.c_double(), C.c_double(), C.c_double()
        self._check(lib().ds_time_pivot_row_ext(self._ctx, ptr(rho), reps, flush_bytes, C.byref(w), C.byref(c),
                                                C.byref(by)), "ds_time_pivot_row_ext")
        return w.value, c.value, by.value


def solve_lp(lp, params=None, device=0, start=None):
    """load + initDualSimplex + presolve("minInfeas") + solve; returns (LPsolution, basis, Result)."""
    with DualSimplex(lp, params, device) as s:
        s.initDualSimplex(start)
        st = s.presolve("minInfeas")
        if st == 0:
            s.solve()
        return s.getSolution(), s.basis(), s.result()
