"""Batched multiscale solves (worker streams) give exactly the results of
one-at-a-time solves, for inputs of different sizes."""
import numpy as np
import pytest

from oracle import oracle as O
from ofblend_amd import manta as M
from ofblend_amd.abi import default_params

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


@pytest.mark.parametrize("shape", [(24, 24, 24, 20), (70, 16, 16, 24)])
def test_autotune_keeps_results(shape):
    """The launch-shape autotuner changes timing only: a solve with it on is
    identical to one with it off, and it records a valid choice for the
    finest level's shape (t-marching workgroups per CU; LDS-tiled or
    t-marching stencil kernels and the tiles' t-segment length).  The
    second shape has a ragged second 64-wide x tile."""
    import ctypes as C
    from ofblend_amd import _lib
    X, Y, Z, T = shape
    print("autotune", shape, flush=True)
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
        tb, lb, sl = C.c_int(), C.c_int(), C.c_int()
        assert L.flof_ctx_tuned_shape(ctx, X, Y, Z, T, C.byref(tb), C.byref(lb),
                                      C.byref(sl)) == 0
        print("tuned: t-march", tb.value, "LDS work items per CU", lb.value,
              "t segment", sl.value, flush=True)
        assert tb.value in (2, 4, 8, 16) and lb.value in (0, 1, 2, 4, 8)
        assert (sl.value == 0) == (lb.value == 0) and 0 <= sl.value <= T
        # a second solve reuses the choice and gives the same result again
        v.clear()
        M.opticalFlowMultiscale4d(vel=v, i0=g0, i1=g1, **P)
        assert np.array_equal(v.to_numpy(), ref)
    finally:
        L.flof_ctx_set_autotune(ctx, 0)


def _jobs(shapes, seed0, nspheres=1):
    jobs = []
    for n, (X, Y, Z, T) in enumerate(shapes):
        s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
        a = O.drop_sdf(X, Y, Z, T, seed=seed0 + 2 * n, nspheres=nspheres, rmin=0.08,
                       rmax=0.18).astype(np.float32)
        b = O.drop_sdf(X, Y, Z, T, seed=seed0 + 1 + 2 * n, nspheres=nspheres, rmin=0.08,
                       rmax=0.18).astype(np.float32)
        g0, g1, v = s.create(M.Grid4Real), s.create(M.Grid4Real), s.create(M.Grid4Vec4)
        g0.from_numpy(a)
        g1.from_numpy(b)
        jobs.append((s, g0, g1, v, a, b))
    return jobs


@pytest.mark.parametrize("precision", [0, 2])
def test_fused_batch_matches_oracle(precision):
    """The fused batch (shared PCG launches per level) on uneven small
    problems with different pyramid depths (3, 2 and 1 levels): each flow,
    iteration record and error is the oracle's (correctly rounded sums), bit
    for bit; also in the mixed-precision form."""
    shapes = [(20, 18, 16, 14), (12, 12, 12, 10), (8, 9, 10, 11), (17, 16, 16, 16)]
    jobs = _jobs(shapes, 40, nspheres=2)
    prm = dict(P, minGridSize=6)
    print("fused batch vs oracle, precision", precision, flush=True)
    errs, infos = M.opticalFlowMultiscale4dFused([j[3] for j in jobs], [j[1] for j in jobs],
                                                 [j[2] for j in jobs], precision=precision,
                                                 **prm)
    for n, (s, g0, g1, v, a, b) in enumerate(jobs):
        with O.sum_mode(1):
            ref, rerr, rinfo = O.multiscale4d(np.zeros(a.shape + (4,), np.float32), a, b,
                                              default_params(precision=precision, **prm))
        got = infos[n].summary()["levels"]
        want = rinfo.summary()["levels"]
        print(n, shapes[n], "iters", [l["iters"] for l in got], flush=True)
        assert [l["iters"] for l in got] == [l["iters"] for l in want]
        assert [l["dims"] for l in got] == [l["dims"] for l in want]
        assert np.array_equal(v.to_numpy(), ref), n
        assert errs[n] == np.float32(rerr)


def test_fused_batch_equals_single_solves():
    """Larger uneven problems (forward and backward pairs): the fused batch
    gives exactly the flows, records and errors of single solves."""
    shapes = [(48, 48, 48, 32), (40, 40, 40, 24), (32, 32, 32, 48)]
    jobs = _jobs(shapes, 60)
    jobs += [(s, g1, g0, s.create(M.Grid4Vec4), b, a) for (s, g0, g1, v, a, b) in jobs]
    prm = dict(P, minGridSize=16)
    ref = []
    for s, g0, g1, v, a, b in jobs:
        v.clear()
        e = M.opticalFlowMultiscale4d(vel=v, i0=g0, i1=g1, **prm)
        ref.append((e, v.to_numpy(), M.getLastSolveInfo().summary()))
        v.clear()
    print("fused batch vs single solves", len(jobs), flush=True)
    errs, infos = M.opticalFlowMultiscale4dFused([j[3] for j in jobs], [j[1] for j in jobs],
                                                 [j[2] for j in jobs], **prm)
    for n, j in enumerate(jobs):
        e, vr, inf = ref[n]
        assert errs[n] == e
        assert np.array_equal(j[3].to_numpy(), vr), n
        assert infos[n].summary()["levels"] == inf["levels"]


def test_fused_batch_with_a_level_solved_on_its_own():
    """A finest level of 64^3 x 32 (8.4M cells, full 64-wide tiles, 16-step
    t segments) runs the LDS-tiled stencil kernels on its own while the
    other problems' levels share the fused launches: every result is still
    the single solve's, bit for bit."""
    shapes = [(64, 64, 64, 32), (24, 20, 18, 16), (40, 40, 40, 24)]
    jobs = _jobs(shapes, 80)
    prm = dict(P, minGridSize=16)
    ref = []
    for s, g0, g1, v, a, b in jobs:
        v.clear()
        e = M.opticalFlowMultiscale4d(vel=v, i0=g0, i1=g1, **prm)
        ref.append((e, v.to_numpy(), M.getLastSolveInfo().summary()))
        v.clear()
    print("fused batch with a solo level", flush=True)
    errs, infos = M.opticalFlowMultiscale4dFused([j[3] for j in jobs], [j[1] for j in jobs],
                                                 [j[2] for j in jobs], **prm)
    for n, j in enumerate(jobs):
        e, vr, inf = ref[n]
        assert errs[n] == e
        assert np.array_equal(j[3].to_numpy(), vr), n
        assert infos[n].summary()["levels"] == inf["levels"]


def test_fused_batch_at_batch_workload_sizes():
    """Two of the 16-solve batch's problems at full size (80^3 x 64 and
    48^3 x 32, forward and backward, 2 spheres): the fused batch equals the
    single solves bit for bit and every level reaches the residual target.
    (The alignment error need not fall: with two spheres far apart the
    linearised flow can raise calcLsDiff above the unwarped pair's.)"""
    shapes = [(80, 80, 80, 64), (48, 48, 48, 32)]
    jobs = _jobs(shapes, 100, nspheres=2)
    jobs += [(s, g1, g0, s.create(M.Grid4Vec4), b, a) for (s, g0, g1, v, a, b) in jobs]
    prm = dict(P, minGridSize=16, cgMaxIter=20000)
    ref = []
    for s, g0, g1, v, a, b in jobs:
        v.clear()
        e = M.opticalFlowMultiscale4d(vel=v, i0=g0, i1=g1, **prm)
        ref.append((e, v.to_numpy(), M.getLastSolveInfo().summary()))
        v.clear()
    print("fused batch at batch-workload sizes", flush=True)
    errs, infos = M.opticalFlowMultiscale4dFused([j[3] for j in jobs], [j[1] for j in jobs],
                                                 [j[2] for j in jobs], **prm)
    for n, j in enumerate(jobs):
        e, vr, inf = ref[n]
        got = infos[n].summary()
        assert errs[n] == e
        assert np.array_equal(j[3].to_numpy(), vr), n
        assert got["levels"] == inf["levels"]
        assert all(r <= 1e-5 for l in got["levels"] for r in l["residual"])


def test_fused_batch_of_one_problem():
    """A batch of one is the single solve."""
    jobs = _jobs([(20, 18, 16, 12)], 120, nspheres=2)
    s, g0, g1, v, a, b = jobs[0]
    prm = dict(P, minGridSize=8)
    e = M.opticalFlowMultiscale4d(vel=v, i0=g0, i1=g1, **prm)
    ref, inf = v.to_numpy(), M.getLastSolveInfo().summary()
    v.clear()
    errs, infos = M.opticalFlowMultiscale4dFused([v], [g0], [g1], **prm)
    assert errs[0] == e and np.array_equal(v.to_numpy(), ref)
    assert infos[0].summary()["levels"] == inf["levels"]


def test_fused_batch_rejects_other_parameters():
    s = M.Solver(gridSize=M.vec3(8, 8, 8), fourthDim=8)
    g0, g1, v = s.create(M.Grid4Real), s.create(M.Grid4Real), s.create(M.Grid4Vec4)
    for bad in (dict(multiStep=3), dict(postVelBlur=2.0), dict(resetBndWidth=0.1),
                dict(doFinalProject=True)):
        prm = dict(P)
        prm.update(bad)
        with pytest.raises(M.FlofError):
            M.opticalFlowMultiscale4dFused([v], [g0], [g1], **prm)
