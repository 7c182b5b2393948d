"""CPU checks of the C-ABI library: it loads, exports every declared entry, validates its
arguments, and its generators reproduce the pinned workloads.  No GPU work is launched."""
import ctypes as C
import hashlib
import json
import os
import re

import numpy as np

from parallel_ds_amd import lp as lpmod
from parallel_ds_amd._native import LIB_PATH, lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in os.listdir(os.path.join(ROOT, "include")):
        if not h.endswith(".h"):
            continue
        txt = open(os.path.join(ROOT, "include", h)).read()
        txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
        for m in re.finditer(r"^\s*(?:const\s+)?[A-Za-z_][\w\s\*]*?\b(ds_\w+)\s*\(", txt, flags=re.M):
            names.add(m.group(1))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = C.CDLL(LIB_PATH)
    names = declared_functions()
    assert len(names) >= 25
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_no_gpu_means_loud_failure():
    """Without a usable device ds_create fails with an error code; nothing falls back."""
    ctx = C.c_void_p()
    rc = lib().ds_create(0, C.byref(ctx))
    try:
        import torch  # noqa: F401
        has_gpu = torch.cuda.is_available()
    except Exception:
        has_gpu = False
    if not has_gpu:
        assert rc != 0 and not ctx.value
    else:
        lib().ds_destroy(ctx)


def test_null_context_is_rejected():
    assert lib().ds_init(None) == -1
    assert lib().ds_solve(None) == -1


def digest(L):
    h = hashlib.sha256()
    for a in (L.col_ptr, L.row_idx, L.val, L.costs, L.upper_range):
        h.update(np.ascontiguousarray(a).tobytes())
    return h.hexdigest()


def test_generators_reproduce_pinned_digests():
    d = json.load(open(os.path.join(ROOT, "tests", "golden", "generator_digests.json")))
    assert digest(lpmod.config0(0)) == d["config0_seed0"]
    assert digest(lpmod.gen_kcol(50, 100, 5, 9)) == d["kcol_m50_n100_k5_seed9"]
    assert digest(lpmod.gen_transport(10, 20, 2)) == d["transport_10x20_seed2"]
    assert digest(lpmod.config3_instance(0)) == d["config3_instance0"]
    assert digest(lpmod.config1()) == d["config1_seed1"]


def test_generator_shapes():
    L = lpmod.gen_kcol(100, 300, 7, 3)
    assert L.nnz == 2100
    for j in range(L.n):
        rows = L.row_idx[L.col_ptr[j]:L.col_ptr[j + 1]]
        assert len(set(rows.tolist())) == 7 and np.all(np.diff(rows) > 0)
    assert np.all((L.val >= 0.1) & (L.val < 1.0)) and np.all(L.costs < 0)
    T = lpmod.gen_transport(20, 30, 4)
    assert T.upper_range[:20].sum() == T.upper_range[20:].sum()
    assert np.all(T.costs >= 1) and np.all(T.costs <= 1000)
    B = lpmod.config3_instance(5)
    assert 32 <= B.m <= 256 and B.n == 2 * B.m and np.all(np.diff(B.col_ptr) >= 1)


def test_python_struct_mirrors_match_the_c_layout():
    """The ctypes mirrors of ds_params, ds_result and ds_batch_params have the C compiler's size
    and the offset of every field (compiled with gcc against include/ds_abi.h)."""
    import shutil
    import subprocess
    import tempfile

    import pytest

    from parallel_ds_amd._native import BatchParams, Params, Result

    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("gcc not found")
    mirrors = [("ds_params", Params), ("ds_result", Result), ("ds_batch_params", BatchParams)]
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "ds_abi.h"', "int main(void) {"]
    for cname, py in mirrors:
        lines.append('  printf("%%s %%zu\\n", "%s", sizeof(%s));' % (cname, cname))
        for f in py._fields_:
            lines.append('  printf("%%s.%%s %%zu\\n", "%s", "%s", offsetof(%s, %s));' % (cname, f[0], cname, f[0]))
    lines.append("  return 0;\n}\n")
    with tempfile.TemporaryDirectory() as d:
        src, exe = os.path.join(d, "layout.c"), os.path.join(d, "layout")
        open(src, "w").write("\n".join(lines))
        subprocess.run([gcc, "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True).stdout.split("\n")
    got = dict(l.split() for l in out if l)
    for cname, py in mirrors:
        assert int(got[cname]) == C.sizeof(py), cname
        for f in py._fields_:
            assert int(got["%s.%s" % (cname, f[0])]) == getattr(py, f[0]).offset, (cname, f[0])
