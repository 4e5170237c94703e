"""The allocation-free calcLsDiff entry point (flof_calc_ls_diff4d_noalloc,
no counterpart in the reference) against the context entry point
flof_calc_ls_diff4d through the C-ABI: bit-identical error value and
per-cell output, on two calls in a row, the second on a smaller grid (the
shared per-device scratch keeps the first call's partials)."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from ofblend_amd import _lib
from ofblend_amd import manta as M

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def run_pair(shape, seeds, corr, bnd):
    T, Z, Y, X = shape
    s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
    L = _lib.lib()
    a = O.drop_sdf(X, Y, Z, T, seed=seeds[0], nspheres=2)
    b = O.drop_sdf(X, Y, Z, T, seed=seeds[1], nspheres=2)
    ga, gb = s.create(M.Grid4Real), s.create(M.Grid4Real)
    ga.from_numpy(a)
    gb.from_numpy(b)
    out_ref, out_na = s.create(M.Grid4Real), s.create(M.Grid4Real)
    out_na.setConst(7.0)
    out_ref.setConst(7.0)
    ref = M.calcLsDiff4d(ga, gb, out_ref, correction=corr, bnd=bnd)
    res = M.DeviceArray((1,))  # the caller's output slot
    ctx = M.context()
    stream = L.flof_ctx_stream(ctx)
    pa = L.flof_grid4d_device_ptr(ga.h)
    pb = L.flof_grid4d_device_ptr(gb.h)
    po = L.flof_grid4d_device_ptr(out_na.h)
    say("noalloc calcLsDiff", shape, "bnd", bnd)
    rc = L.flof_calc_ls_diff4d_noalloc(pa, pb, po, X, Y, Z, T, corr, bnd, stream, res.ptr)
    assert rc == 0, L.flof_last_error(None)
    M.sync()
    got = float(res.to_numpy()[0])
    assert np.float32(got).tobytes() == np.float32(ref).tobytes(), (got, ref)
    assert np.array_equal(out_na.to_numpy(), out_ref.to_numpy())
    # and the oracle's value (the reference's arithmetic)
    with O.sum_mode(1):
        assert got == float(np.float32(O.ls_diff(a, b, correction=corr, bnd=bnd)))
    return got


def test_noalloc_ls_diff_matches_context_entry_twice():
    big = run_pair((24, 30, 28, 26), (1, 2), 0.005, 1)
    small = run_pair((6, 9, 8, 7), (3, 4), 0.3, 0)
    assert big > 0 and small > 0


def test_noalloc_ls_diff_single_time_slice():
    """A grid of one time slice: the 3D scaling of calcLsDiff4d (x1000 once).
    (A border band applies to t as well, so with one slice only bnd 0
    leaves cells to count.)"""
    assert run_pair((1, 12, 13, 14), (5, 6), 0.02, 0) > 0


def test_noalloc_ls_diff_rejects_bad_arguments():
    L = _lib.lib()
    res = M.DeviceArray((1,))
    # 2^31 cells: rejected before anything is launched
    rc = L.flof_calc_ls_diff4d_noalloc(res.ptr, res.ptr, None, 1 << 11, 1 << 10, 1 << 10,
                                       1, 1.0, 0, None, res.ptr)
    assert rc == -7 and b"2^31" in L.flof_last_error(None)
    rc = L.flof_calc_ls_diff4d_noalloc(None, res.ptr, None, 4, 4, 4, 4, 1.0, 0, None, res.ptr)
    assert rc == -1
