"""The C-ABI library loads and exports every entry point that include/flof.h
declares; the ctypes mirrors match the C struct layouts (no GPU needed)."""
import ctypes as C
import os
import re

import pytest

from ofblend_amd import _lib
from ofblend_amd.abi import OfParams, SolveInfo, default_params

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared():
    src = open(os.path.join(ROOT, "include", "flof.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(flof_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.lib()
    names = declared()
    assert len(names) >= 50
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # and the Python bindings cover the whole header
    assert set(names) == set(_lib.SIGNATURES), set(names) ^ set(_lib.SIGNATURES)


def test_header_is_plain_c_and_links(tmp_path):
    """include/flof.h compiles as strict C99 (a C, Go/cgo or FFI caller
    includes it as is) and a C program that takes the address of every
    declared entry point links against libflof_hip.so."""
    import shutil
    import subprocess
    gcc = shutil.which("gcc")
    if gcc is None:
        pytest.skip("no gcc")
    names = declared()
    src = tmp_path / "uses_all.c"
    src.write_text('#include "flof.h"\n#include <stdio.h>\n'
                   "typedef void (*any_fn)(void);\n"
                   "static const any_fn entry_points[] = {\n"
                   + "".join("    (any_fn)&%s,\n" % n for n in names)
                   + "};\nint main(void) {\n"
                   '    printf("%d\\n", (int)(sizeof entry_points / sizeof entry_points[0]));\n'
                   "    return 0;\n}\n")
    inc = os.path.join(ROOT, "include")
    r = subprocess.run([gcc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror",
                        "-I", inc, "-c", str(src), "-o", str(tmp_path / "u.o")],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    libdir = os.path.dirname(_lib.LIB_PATH)
    r = subprocess.run([gcc, str(tmp_path / "u.o"), "-o", str(tmp_path / "u"), "-L", libdir,
                        "-lflof_hip", "-Wl,--unresolved-symbols=ignore-in-shared-libs"],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_struct_layouts_match():
    L = _lib.lib()
    assert L.flof_abi_struct_size(0) == C.sizeof(OfParams)
    assert L.flof_abi_struct_size(1) == C.sizeof(SolveInfo)


def test_default_params_match_reference_defaults():
    p = OfParams()
    _lib.lib().flof_of_params_default(C.byref(p))
    q = default_params()
    for name, _ in OfParams._fields_:
        assert getattr(p, name) == getattr(q, name), name
    # optflow4d.cpp:2182-2189 defaults
    assert (q.wSmooth, q.wEnergy, q.postVelBlur, q.cgAccuracy, q.cfl,
            q.resetBndWidth, q.multiStep, q.projSizeThresh, q.minGridSize,
            q.doFinalProject) == (0.0, 0.0, 0.0, 1e-4, 999.0, -1.0, 1, 9999, 10, 0)


def test_errors_are_reported_not_thrown():
    L = _lib.lib()
    # null arguments return an error code instead of crashing
    assert L.flof_grid4d_create(None, 4, 4, 4, 4, 1, 0, None) != 0
    assert L.flof_sync(None) != 0
    assert L.flof_last_error(None) == b"null context"
    # the allocation-free entry point checks its arguments before any HIP call
    rc = L.flof_calc_ls_diff4d_noalloc(None, None, None, 4, 4, 4, 4, 1.0, 0, None, None)
    assert rc == -1 and b"null argument" in L.flof_last_error(None)


def test_bench_refuses_more_than_one_gpu():
    """bench.py measures one GPU: --gpus N > 1 exits with a message before
    any GPU work (runs on the CPU)."""
    import subprocess
    import sys
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    p = subprocess.run([sys.executable, os.path.join(root, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""
