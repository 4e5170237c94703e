"""Batched multiscale solves (worker streams) give exactly the results of
one-at-a-time solves, for inputs of different sizes."""
import numpy as np
import pytest

from oracle import oracle as O
from ofblend_amd import manta as M

pytestmark = pytest.mark.gpu

P = dict(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, minGridSize=8, multiStep=1,
         postVelBlur=0.0, doFinalProject=False, resetBndWidth=-1.0, cfl=999.0)


def test_batch_equals_single_solves():
    shapes = [(16, 16, 16, 16), (20, 18, 16, 12), (12, 12, 12, 20), (24, 16, 16, 16),
              (16, 16, 16, 16)]
    jobs = []
    for n, (X, Y, Z, T) in enumerate(shapes):
        print("batch case", n, (X, Y, Z, T), flush=True)
        s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
        a = O.drop_sdf(X, Y, Z, T, seed=10 + 2 * n).astype(np.float32)
        b = O.drop_sdf(X, Y, Z, T, seed=11 + 2 * n).astype(np.float32)
        g0, g1, v = s.create(M.Grid4Real), s.create(M.Grid4Real), s.create(M.Grid4Vec4)
        g0.from_numpy(a)
        g1.from_numpy(b)
        jobs.append((s, g0, g1, v))
    ref = []
    for s, g0, g1, v in jobs:
        v.clear()
        e = M.opticalFlowMultiscale4d(vel=v, i0=g0, i1=g1, **P)
        ref.append((e, v.to_numpy(), M.getLastSolveInfo().summary()))
    for s, g0, g1, v in jobs:
        v.clear()
    errs, infos = M.opticalFlowMultiscale4dBatch([j[3] for j in jobs], [j[1] for j in jobs],
                                                 [j[2] for j in jobs], concurrency=3, **P)
    for n, (s, g0, g1, v) in enumerate(jobs):
        e, vr, inf = ref[n]
        assert errs[n] == e
        assert np.array_equal(v.to_numpy(), vr), n
        assert infos[n].summary()["levels"] == inf["levels"]


def test_autotune_keeps_results():
    """The launch-shape autotuner changes timing only: a solve with it on is
    identical to one with it off, and it records a candidate per shape."""
    from ofblend_amd import _lib
    X, Y, Z, T = 24, 24, 24, 20
    print("autotune", (X, Y, Z, T), flush=True)
    s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
    g0, g1, v = s.create(M.Grid4Real), s.create(M.Grid4Real), s.create(M.Grid4Vec4)
    g0.from_numpy(O.drop_sdf(X, Y, Z, T, seed=31).astype(np.float32))
    g1.from_numpy(O.drop_sdf(X, Y, Z, T, seed=32).astype(np.float32))
    M.opticalFlowMultiscale4d(vel=v, i0=g0, i1=g1, **P)
    ref = v.to_numpy()
    L = _lib.lib()
    ctx = M.context()
    L.flof_ctx_set_autotune(ctx, 1)
    try:
        v.clear()
        M.opticalFlowMultiscale4d(vel=v, i0=g0, i1=g1, **P)
        assert np.array_equal(v.to_numpy(), ref)
        assert L.flof_ctx_tuned_blocks(ctx, X, Y, Z, T) in (2, 4, 8, 16)
    finally:
        L.flof_ctx_set_autotune(ctx, 0)
