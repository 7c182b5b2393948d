"""ctypes binding of libpds.so (the C-ABI declared in include/ds_abi.h and include/ds_gen.h).

The library is built in-tree by ``parallel_ds_amd.build``; importing this module never
falls back to anything else: a missing library raises immediately.
"""
import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libpds.so")

DS_FIXED, DS_FREE, DS_BOXED, DS_UPPER, DS_LOWER = 0, 1, 2, 3, 4
DS_OPTIMAL, DS_PRIMAL_INFEASIBLE, DS_DUAL_INFEASIBLE, DS_ITERATION_LIMIT, DS_NOT_SOLVED = 0, 1, 2, 3, 4
ERRORS = {0: "DS_OK", -1: "DS_EINVAL", -2: "DS_EHIP", -3: "DS_ESINGULAR", -4: "DS_EUNSUPPORTED",
          -5: "DS_ESTATE", -6: "DS_EINTERNAL"}


class Params(C.Structure):
    _fields_ = [("refactor_freq", C.c_int), ("max_iterations", C.c_int), ("ratio_test", C.c_int),
                ("perturb", C.c_int), ("perturb_seed", C.c_uint64), ("chunk", C.c_int), ("verbose", C.c_int),
                ("mark_iteration", C.c_int), ("time_kernels", C.c_int), ("cleanup", C.c_int),
                ("profile", C.c_int), ("force_event", C.c_int), ("force_at", C.c_int),
                ("time_limit_ms", C.c_double), ("candidates", C.c_int), ("candidate_psi", C.c_double),
                ("warm_primal", C.c_int), ("warm_shift", C.c_double), ("warm_primal_cap", C.c_int),
                ("split_launches", C.c_int)]


DS_FORCE_NONE, DS_FORCE_STABILITY, DS_FORCE_DUAL_RESTART, DS_FORCE_SINGULAR = 0, 1, 2, 3


# Steps of Result.step_ms (the reference's TimeProfiler parts, common/types.hpp:12-27).
PARTS = ("Init", "Factor", "RestoreProc", "Btran", "PivotRow", "RatioTestPart1", "RatioTestPart2", "Ftran",
         "FtranBfrt", "FtranBeta", "Pricing", "UpdateRedCosts", "BasisUpdate")


class Result(C.Structure):
    _fields_ = [("status", C.c_int), ("solved", C.c_int), ("iterations", C.c_int),
                ("phase1_iterations", C.c_int), ("perturbed", C.c_int), ("refactors", C.c_int),
                ("restarts", C.c_int), ("Z", C.c_double), ("solve_ms", C.c_double), ("refactor_ms", C.c_double),
                ("mark_ms", C.c_double), ("spmv_ms", C.c_double), ("spmv_launches", C.c_int), ("kernel_k", C.c_int),
                ("nnz_nonbasic_mark", C.c_double), ("nnz_nonbasic_end", C.c_double), ("cleanups", C.c_int),
                ("repairs", C.c_int), ("step_ms", C.c_double * 13), ("primal_iterations", C.c_int),
                ("shifted_bounds", C.c_int)]


class BatchParams(C.Structure):
    _fields_ = [("refactor_freq", C.c_int), ("max_iterations", C.c_int), ("perturb", C.c_int),
                ("perturb_seed", C.c_uint64), ("block", C.c_int), ("lds_inv_max_m", C.c_int),
                ("wgs_per_cu", C.c_int), ("use_graph", C.c_int), ("cleanup", C.c_int)]


class DsError(RuntimeError):
    def __init__(self, code, where, detail=""):
        super().__init__("%s failed: %s %s" % (where, ERRORS.get(code, code), detail))
        self.code = code


_lib = None

VP = C.c_void_p
I32 = C.c_int
F64 = C.c_double
U64 = C.c_uint64

_SIGS = {
    "ds_default_params": (None, [C.POINTER(Params)]),
    "ds_create": (I32, [I32, C.POINTER(VP)]),
    "ds_destroy": (I32, [VP]),
    "ds_last_error": (C.c_char_p, [VP]),
    "ds_load_problem": (I32, [VP, I32, I32, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    "ds_set_params": (I32, [VP, C.POINTER(Params)]),
    "ds_init": (I32, [VP]),
    "ds_init_basis": (I32, [VP, VP]),
    "ds_load_mps": (I32, [VP, C.c_char_p, C.POINTER(I32), C.POINTER(I32)]),
    "ds_set_offset": (I32, [VP, F64]),
    "ds_mps_read": (I32, [C.c_char_p, C.POINTER(VP), C.c_char_p, I32]),
    "ds_mps_dims": (I32, [VP, C.POINTER(I32), C.POINTER(I32), C.POINTER(I32)]),
    "ds_mps_get": (I32, [VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, C.POINTER(F64)]),
    "ds_mps_free": (I32, [VP]),
    "ds_presolve": (I32, [VP, C.c_char_p, C.POINTER(I32)]),
    "ds_solve": (I32, [VP]),
    "ds_get_result": (I32, [VP, C.POINTER(Result)]),
    "ds_get_solution": (I32, [VP, C.POINTER(F64), C.POINTER(I32), C.c_char_p, I32, VP]),
    "ds_get_basis": (I32, [VP, VP]),
    "ds_get_nonbasis": (I32, [VP, VP]),
    "ds_chuzr": (I32, [VP, I32, VP, VP, VP, VP, VP, VP, VP]),
    "ds_chuzr_noalloc": (I32, [VP, I32, VP, VP, VP, VP, VP, VP, VP]),
    "ds_solve_lp": (I32, [VP, I32, I32, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, C.POINTER(Params),
                          C.POINTER(Result), VP, VP]),
    "ds_set_basis": (I32, [VP, VP]),
    "ds_refactor": (I32, [VP]),
    "ds_btran": (I32, [VP, I32, VP]),
    "ds_ftran": (I32, [VP, VP, VP]),
    "ds_ftran_column": (I32, [VP, I32, VP]),
    "ds_pivot_row": (I32, [VP, VP, VP]),
    "ds_time_pivot_row": (I32, [VP, VP, I32, C.POINTER(F64), C.POINTER(F64)]),
    "ds_time_pivot_row_ext": (I32, [VP, VP, I32, C.c_longlong, C.POINTER(F64), C.POINTER(F64), C.POINTER(F64)]),
    "ds_debug_gemm_tile": (I32, [VP, VP, VP, I32, I32]),
    "ds_time_refactor": (I32, [VP, I32, C.POINTER(F64), C.POINTER(F64), C.POINTER(F64), C.POINTER(I32)]),
    "ds_debug_copy": (I32, [C.c_longlong, C.POINTER(F64)]),
    "ds_batch_default_params": (None, [C.POINTER(BatchParams)]),
    "ds_batch_load": (I32, [VP, I32, VP, VP, VP, VP, VP, VP, VP, VP]),
    "ds_batch_set_warm": (I32, [VP, VP]),
    "ds_batch_set_bases": (I32, [VP, VP]),
    "ds_batch_solve": (I32, [VP, C.POINTER(BatchParams), C.POINTER(F64)]),
    "ds_batch_replay": (I32, [VP, C.POINTER(F64)]),
    "ds_batch_results": (I32, [VP, VP, VP, VP, VP, VP]),
    "ds_batch_tune": (I32, [VP, I32, C.POINTER(BatchParams), C.POINTER(F64), C.POINTER(I32)]),
    "ds_mlp_create": (I32, [I32, I32, U64, C.POINTER(VP)]),
    "ds_mlp_destroy": (I32, [VP]),
    "ds_mlp_last_error": (C.c_char_p, [VP]),
    "ds_mlp_param_count": (C.c_longlong, []),
    "ds_mlp_get_params": (I32, [VP, VP]),
    "ds_mlp_set_params": (I32, [VP, VP]),
    "ds_mlp_set_norm": (I32, [VP, VP, VP]),
    "ds_mlp_get_norm": (I32, [VP, VP, VP]),
    "ds_mlp_get_prior": (I32, [VP, VP, I32, C.POINTER(I32)]),
    "ds_mlp_infer": (I32, [VP, I32, VP, VP]),
    "ds_mlp_grad": (I32, [VP, I32, VP, VP, VP, C.POINTER(F64), VP]),
    "ds_mlp_adam_step": (I32, [VP, C.c_float]),
    "ds_mlp_fit": (I32, [VP, VP, I32, I32, C.c_float, U64, VP, C.POINTER(F64)]),
    "ds_mlp_train_stats": (I32, [VP, C.POINTER(F64), C.POINTER(C.c_longlong)]),
    "ds_batch_features": (I32, [VP, VP]),
    "ds_batch_labels": (I32, [VP, VP]),
    "ds_batch_warm_from_scores": (I32, [VP, VP]),
    "ds_batch_warm_start": (I32, [VP, VP, C.POINTER(F64), C.POINTER(F64)]),
    "ds_batch_get_warm": (I32, [VP, VP]),
    "ds_batch_set_repair_tol": (I32, [VP, F64]),
    "ds_batch_repairs": (I32, [VP, VP]),
    "ds_gen_bernoulli_lp": (I32, [I32, I32, F64, U64, VP, VP, VP, I32, VP, VP]),
    "ds_gen_kcol_lp": (I32, [I32, I32, I32, U64, VP, VP, VP, VP, VP]),
    "ds_gen_transport": (I32, [I32, I32, U64, VP, VP, VP, VP, VP]),
    "ds_gen_lognormal": (I32, [I32, F64, U64, VP, VP]),
}


def lib():
    """Load libpds.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libpds.so not built: run `python -m parallel_ds_amd.build` (%s)" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def ptr(a):
    """Data pointer of a contiguous numpy array (or None)."""
    return None if a is None else a.ctypes.data_as(VP)
