"""The measurement entry points through the C-ABI: per-kernel timing of the
PCG launch kinds (tagged by level) and the streaming-bandwidth probe."""
import ctypes as C

import numpy as np
import pytest

from ofblend_amd import _lib
from ofblend_amd import manta as M
from ofblend_amd.abi import (FLOF_KT_CG_APPLY, FLOF_KT_CG_DIR, FLOF_KT_CG_UPDATE, FLOF_MIX_COPY,
                             FLOF_MIX_DIR, FLOF_MIX_DIR_X, FLOF_MIX_UPDATE)
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def test_stream_bandwidth_reports_and_rejects():
    L = _lib.lib()
    ctx = M.context()
    for mix in (FLOF_MIX_UPDATE, FLOF_MIX_DIR, FLOF_MIX_COPY, FLOF_MIX_DIR_X):
        g = C.c_double()
        assert L.flof_stream_bandwidth(ctx, mix, 1 << 22, 3, C.byref(g)) == 0
        print("mix", mix, "GB/s", g.value, flush=True)
        assert 100.0 < g.value < 8000.0
    g = C.c_double()
    assert L.flof_stream_bandwidth(ctx, 4, 1 << 20, 3, C.byref(g)) != 0
    assert L.flof_stream_bandwidth(ctx, FLOF_MIX_COPY, 0, 3, C.byref(g)) != 0


def test_preconditioner_reciprocal_is_the_ieee_division():
    """The fast reciprocal of the Jacobi preconditioner (hardware reciprocal
    plus one Newton step) equals 1.f / d bit for bit for every float in
    [2^-126, 2^126), the range the set-up requires of every diagonal entry
    before the kernels use it, and differs outside it (denormals, d >= 2^126),
    where the kernels keep the division."""
    L = _lib.lib()
    ctx = M.context()
    bad = C.c_uint64(99)
    assert L.flof_check_reciprocal(ctx, 0x00800000, 0x7e7fffff, C.byref(bad)) == 0
    print("normal range: %d mismatches" % bad.value, flush=True)
    assert bad.value == 0
    for lo, hi in ((0x00000001, 0x007fffff), (0x7e800000, 0x7f7fffff)):
        assert L.flof_check_reciprocal(ctx, lo, hi, C.byref(bad)) == 0
        print("bits [%08x, %08x]: %d mismatches" % (lo, hi, bad.value), flush=True)
        assert bad.value > 0
    assert L.flof_check_reciprocal(ctx, 5, 4, C.byref(bad)) != 0


def test_timing_tags_count_the_level_launches():
    """Every PCG iteration of a single-level solve is one launch of each of
    the three kinds, tagged with the level's cell count."""
    L = _lib.lib()
    ctx = M.context()
    X, Y, Z, T = 20, 16, 12, 10
    s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
    a, b = s.create(M.Grid4Real), s.create(M.Grid4Real)
    a.from_numpy(O.drop_sdf(X, Y, Z, T, seed=61).astype(np.float32))
    b.from_numpy(O.drop_sdf(X, Y, Z, T, seed=62).astype(np.float32))
    v = s.create(M.Grid4Vec4)
    kinds = (FLOF_KT_CG_APPLY, FLOF_KT_CG_UPDATE, FLOF_KT_CG_DIR)
    for k in kinds:
        L.flof_ctx_timing(ctx, k, None, None)
        L.flof_ctx_enable_timing(ctx, k, 1)
    try:
        M.opticalFlow4d(v, a, b, wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5)
        it = M.getLastSolveInfo().summary()["levels"][0]["iters"][0]
        for k in kinds:
            n, ms = C.c_int64(), C.c_double()
            assert L.flof_ctx_timing_tag(ctx, k, X * Y * Z * T, C.byref(n), C.byref(ms)) == 0
            # chunks of 16 iterations, the host checks the state one chunk
            # behind: launches after convergence (return at once) are counted
            assert it <= n.value <= it + 32, (k, it, n.value)
            assert ms.value > 0.0
            n2 = C.c_int64()
            assert L.flof_ctx_timing_tag(ctx, k, 7, C.byref(n2), None) == 0
            assert n2.value == 0
    finally:
        for k in kinds:
            L.flof_ctx_enable_timing(ctx, k, 0)
            L.flof_ctx_timing(ctx, k, None, None)
