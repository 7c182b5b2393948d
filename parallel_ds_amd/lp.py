"""LP instances in the reference's model form and the seeded synthetic workloads.

An LP is ``min c^T x`` over structurals x with column bounds and row ranges, exactly the
arrays of the reference's 8-array Problem constructor (src/common/problem.tpp:143-173);
A is held in CSC with int32 indices and FP64 values.
"""
from dataclasses import dataclass

import numpy as np

from ._native import DS_BOXED, DS_FIXED, DS_LOWER, DS_UPPER, lib, ptr


@dataclass
class LP:
    m: int
    n: int
    col_ptr: np.ndarray
    row_idx: np.ndarray
    val: np.ndarray
    costs: np.ndarray
    lower_bound: np.ndarray
    upper_bound: np.ndarray
    bound_type: np.ndarray
    lower_range: np.ndarray
    upper_range: np.ndarray
    range_type: np.ndarray
    name: str = ""
    offset: float = 0.0   # objective constant (Problem::_offset), removed from Z

    @property
    def nnz(self):
        return int(self.col_ptr[-1])

    def arrays(self):
        """Argument tuple for ds_load_problem / ora_solve (after m, n)."""
        return (self.col_ptr, self.row_idx, self.val, self.costs, self.lower_bound, self.upper_bound,
                self.bound_type, self.lower_range, self.upper_range, self.range_type)

    def dense(self):
        A = np.zeros((self.m, self.n))
        for j in range(self.n):
            for e in range(self.col_ptr[j], self.col_ptr[j + 1]):
                A[self.row_idx[e], j] += self.val[e]
        return A


def packing_lp(m, n, col_ptr, row_idx, val, b, c, name=""):
    """min c^T x, A x <= b, x >= 0."""
    return LP(m, n, np.ascontiguousarray(col_ptr, np.int32), np.ascontiguousarray(row_idx, np.int32),
              np.ascontiguousarray(val, np.float64), np.ascontiguousarray(c, np.float64),
              np.zeros(n), np.full(n, np.inf), np.full(n, DS_LOWER, np.int32),
              np.full(m, -np.inf), np.ascontiguousarray(b, np.float64), np.full(m, DS_UPPER, np.int32), name)


def equality_lp(m, n, col_ptr, row_idx, val, rhs, c, name=""):
    """min c^T x, A x = rhs, x >= 0."""
    rhs = np.ascontiguousarray(rhs, np.float64)
    return LP(m, n, np.ascontiguousarray(col_ptr, np.int32), np.ascontiguousarray(row_idx, np.int32),
              np.ascontiguousarray(val, np.float64), np.ascontiguousarray(c, np.float64),
              np.zeros(n), np.full(n, np.inf), np.full(n, DS_LOWER, np.int32),
              rhs.copy(), rhs.copy(), np.full(m, DS_FIXED, np.int32), name)


def gen_bernoulli(m, n, density, seed):
    """Random packing LP: Bernoulli(density) positions, values U(0.1,1), b = A x0 + U(0,1), c = -U(0,1)."""
    cap = max(16, int(m * n * density * 1.5) + 64)
    while True:
        cp = np.zeros(n + 1, np.int32)
        ri = np.zeros(cap, np.int32)
        v = np.zeros(cap)
        b = np.zeros(m)
        c = np.zeros(n)
        nnz = lib().ds_gen_bernoulli_lp(m, n, density, seed, ptr(cp), ptr(ri), ptr(v), cap, ptr(b), ptr(c))
        if nnz == -2:
            cap *= 2
            continue
        if nnz < 0:
            raise ValueError("ds_gen_bernoulli_lp failed")
        return packing_lp(m, n, cp, ri[:nnz].copy(), v[:nnz].copy(), b, c, "bernoulli-m%d-n%d-s%d" % (m, n, seed))


def gen_kcol(m, n, k, seed):
    """Random packing LP with exactly k uniformly random rows per column."""
    cp = np.zeros(n + 1, np.int32)
    ri = np.zeros(n * k, np.int32)
    v = np.zeros(n * k)
    b = np.zeros(m)
    c = np.zeros(n)
    rc = lib().ds_gen_kcol_lp(m, n, k, seed, ptr(cp), ptr(ri), ptr(v), ptr(b), ptr(c))
    if rc < 0:
        raise ValueError("ds_gen_kcol_lp failed")
    return packing_lp(m, n, cp, ri, v, b, c, "kcol-m%d-n%d-k%d-s%d" % (m, n, k, seed))


def gen_transport(S, D, seed):
    """Balanced transportation LP, S sources x D sinks (equality rows)."""
    n = S * D
    m = S + D
    cp = np.zeros(n + 1, np.int32)
    ri = np.zeros(2 * n, np.int32)
    v = np.zeros(2 * n)
    rhs = np.zeros(m)
    c = np.zeros(n)
    rc = lib().ds_gen_transport(S, D, seed, ptr(cp), ptr(ri), ptr(v), ptr(rhs), ptr(c))
    if rc < 0:
        raise ValueError("ds_gen_transport failed")
    return equality_lp(m, n, cp, ri, v, rhs, c, "transport-%dx%d-s%d" % (S, D, seed))


def lognormal(base, sigma, seed):
    base = np.ascontiguousarray(base, np.float64)
    out = np.zeros_like(base)
    if lib().ds_gen_lognormal(len(base), sigma, seed, ptr(base), ptr(out)) != 0:
        raise ValueError("ds_gen_lognormal failed")
    return out


def perturbed_instance(base, sigma, seed):
    """Family instance: b and c of `base` times exp(sigma * N(0,1)), elementwise."""
    b = lognormal(base.upper_range, sigma, seed * 2 + 1)
    c = lognormal(base.costs, sigma, seed * 2 + 2)
    return packing_lp(base.m, base.n, base.col_ptr, base.row_idx, base.val, b, c,
                      "%s-pert%d" % (base.name, seed))


# Workloads of the benchmark, named by what they are.
def config0(seed=0):
    return gen_bernoulli(200, 400, 0.05, seed)


def config1(seed=1, m=20000, n=40000, k=10):
    return gen_kcol(m, n, k, seed)


def config2(seed=2, S=500, D=1000):
    return gen_transport(S, D, seed)


def config3_instance(k):
    """Instance k of the batched family: m ~ U{32..256}, n = 2m, density 10%, seed 3000 + k."""
    seed = 3000 + k
    m = 32 + _splitmix_first_int(seed, 225)
    return gen_bernoulli(m, 2 * m, 0.10, seed)


def _splitmix_first_int(seed, span):
    """floor(U * span) of the first draw of a SplitMix64 stream seeded with seed ^ 0x9e37."""
    s = (seed ^ 0x9E37) & 0xFFFFFFFFFFFFFFFF
    s = (s + 0x9E3779B97F4A7C15) & 0xFFFFFFFFFFFFFFFF
    z = s
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & 0xFFFFFFFFFFFFFFFF
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & 0xFFFFFFFFFFFFFFFF
    z = z ^ (z >> 31)
    u = (z >> 11) * (1.0 / 9007199254740992.0)
    return min(span - 1, int(u * span))


def boxed_instance(m, n, density, seed):
    """The small-LP recipe with boxed structurals and ranged rows: 0 <= x_j <= u_j,
    u_j ~ U(0.3, 1), and 0.2 b_i <= (A x)_i <= b_i (a Boxed slack per row).  With c < 0 most
    structurals end at their upper bound, so the dual simplex flips bounds throughout."""
    base = gen_bernoulli(m, n, density, seed)
    rng = np.random.default_rng(seed + 7919)
    ub = rng.uniform(0.3, 1.0, n)
    b = base.upper_range
    return LP(m, n, base.col_ptr, base.row_idx, base.val, base.costs, np.zeros(n), ub,
              np.full(n, DS_BOXED, np.int32), 0.2 * b, b.copy(), np.full(m, DS_BOXED, np.int32),
              "boxed-m%d-n%d-s%d" % (m, n, seed))


def read_mps(path):
    """Problem::readMps (src/common/problem.tpp:331-415) with this library's MPS parser: the
    8 arrays of the model before transformToComputeForm, and the objective constant."""
    import ctypes as C
    from ._native import DsError
    L = lib()
    h = C.c_void_p()
    err = C.create_string_buffer(512)
    rc = L.ds_mps_read(str(path).encode(), C.byref(h), err, 512)
    if rc != 0:
        raise DsError(rc, "ds_mps_read", err.value.decode())
    try:
        m, n, nnz = C.c_int(), C.c_int(), C.c_int()
        L.ds_mps_dims(h, C.byref(m), C.byref(n), C.byref(nnz))
        m, n, nnz = m.value, n.value, nnz.value
        cp = np.zeros(n + 1, np.int32)
        ri = np.zeros(max(nnz, 1), np.int32)
        v = np.zeros(max(nnz, 1))
        c, lb, ub = np.zeros(n), np.zeros(n), np.zeros(n)
        bt = np.zeros(n, np.int32)
        lr, ur = np.zeros(m), np.zeros(m)
        rt = np.zeros(m, np.int32)
        off = C.c_double()
        L.ds_mps_get(h, ptr(cp), ptr(ri), ptr(v), ptr(c), ptr(lb), ptr(ub), ptr(bt), ptr(lr), ptr(ur), ptr(rt),
                     C.byref(off))
    finally:
        L.ds_mps_free(h)
    import os
    return LP(m, n, cp, ri[:nnz].copy(), v[:nnz].copy(), c, lb, ub, bt, lr, ur, rt,
              os.path.basename(str(path)), off.value)


def write_mps(lp, path, name="LP"):
    """Free-format MPS of an LP (test and interchange helper): rows E/L/G with RANGES for ranged
    rows, BOUNDS for every column whose bounds differ from [0, +inf)."""
    inf = 1e30
    with open(path, "w") as f:
        f.write("NAME %s\nROWS\n N COST\n" % name)
        kind = []
        for i in range(lp.m):
            lo, up = lp.lower_range[i], lp.upper_range[i]
            if np.isfinite(lo) and np.isfinite(up) and lo == up:
                kind.append("E")
            elif np.isfinite(up):
                kind.append("L")
            elif np.isfinite(lo):
                kind.append("G")
            else:
                raise ValueError("free rows are not written")
            f.write(" %s R%d\n" % (kind[-1], i))
        f.write("COLUMNS\n")
        for j in range(lp.n):
            if lp.costs[j] != 0.0:
                f.write(" X%d COST %.17g\n" % (j, lp.costs[j]))
            for e in range(lp.col_ptr[j], lp.col_ptr[j + 1]):
                f.write(" X%d R%d %.17g\n" % (j, lp.row_idx[e], lp.val[e]))
            if lp.costs[j] == 0.0 and lp.col_ptr[j] == lp.col_ptr[j + 1]:
                f.write(" X%d COST 0\n" % j)
        f.write("RHS\n")
        for i in range(lp.m):
            rhs = lp.lower_range[i] if kind[i] == "G" else lp.upper_range[i]
            if rhs != 0.0:
                f.write(" RHS R%d %.17g\n" % (i, rhs))
        if lp.offset != 0.0:
            f.write(" RHS COST %.17g\n" % (-lp.offset))
        rng = [(i, lp.upper_range[i] - lp.lower_range[i]) for i in range(lp.m)
               if kind[i] != "E" and np.isfinite(lp.lower_range[i]) and np.isfinite(lp.upper_range[i])]
        if rng:
            f.write("RANGES\n")
            for i, r in rng:
                f.write(" RNG R%d %.17g\n" % (i, r))
        f.write("BOUNDS\n")
        for j in range(lp.n):
            lo, up = lp.lower_bound[j], lp.upper_bound[j]
            if not np.isfinite(lo) and not np.isfinite(up):
                f.write(" FR BND X%d\n" % j)
                continue
            if np.isfinite(lo) and np.isfinite(up) and lo == up:
                f.write(" FX BND X%d %.17g\n" % (j, lo))
                continue
            if not np.isfinite(lo):
                f.write(" MI BND X%d\n" % j)
            elif lo != 0.0:
                f.write(" LO BND X%d %.17g\n" % (j, lo))
            if np.isfinite(up):
                f.write(" UP BND X%d %.17g\n" % (j, up))
        f.write("ENDATA\n")
    _ = inf
