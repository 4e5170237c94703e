"""Reduced-precision storage of the CG search direction (bfloat16, an
extension): the GPU result is bit-identical to the oracle run with the same
storage (correctly rounded sums), and its distance to the FP64 oracle is
bounded and reported."""
import numpy as np
import pytest

from oracle import oracle as O
from ofblend_amd import manta as M
from ofblend_amd.abi import FLOF_PSTORE_BF16, default_params

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _grids(s, a, b):
    g0, g1 = s.create(M.Grid4Real), s.create(M.Grid4Real)
    g0.from_numpy(a)
    g1.from_numpy(b)
    return g0, g1


def test_bf16_direction_level_solve_bit_exact():
    N, T = 14, 12
    i0 = O.drop_sdf(N, N, N, T, seed=1)
    i1 = O.drop_sdf(N, N, N, T, seed=2)
    say("bf16 level solve", (N, N, N, T))
    s = M.Solver(gridSize=M.vec3(N, N, N), fourthDim=T)
    g0, g1 = _grids(s, i0, i1)
    vel = s.create(M.Grid4Vec4)
    M.opticalFlow4d(vel, g0, g1, wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5,
                    precision=FLOF_PSTORE_BF16)
    it_gpu = M.getLastSolveInfo().summary()["levels"][0]["iters"]
    p = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5,
                       precision=FLOF_PSTORE_BF16)
    with O.sum_mode(1):
        ref, it, res, st = O.optical_flow4d(np.zeros(i0.shape + (4,), np.float32), i0, i1, p)
    say("iterations gpu", it_gpu, "oracle", it)
    assert it_gpu == [it]
    assert np.array_equal(vel.to_numpy(), ref)


def test_bf16_direction_multiscale_bit_exact_and_error():
    N, T = 24, 20
    i0 = O.drop_sdf(N, N, N, T, seed=5, nspheres=2, rmin=0.08, rmax=0.18)
    i1 = O.drop_sdf(N, N, N, T, seed=6, nspheres=2, rmin=0.08, rmax=0.18)
    say("bf16 multiscale", (N, N, N, T))
    s = M.Solver(gridSize=M.vec3(N, N, N), fourthDim=T)
    g0, g1 = _grids(s, i0, i1)
    vel = s.create(M.Grid4Vec4)
    prm = dict(O.P_PLAIN, minGridSize=8)
    M.opticalFlowMultiscale4d(vel=vel, i0=g0, i1=g1, precision=FLOF_PSTORE_BF16, **prm)
    v = vel.to_numpy()
    with O.sum_mode(1):
        ref, _, _ = O.multiscale4d(np.zeros(v.shape, np.float32), i0, i1,
                                   default_params(precision=FLOF_PSTORE_BF16, **prm))
    assert np.array_equal(v, ref)
    f64, _, _ = O.multiscale4d(np.zeros(v.shape), i0.astype(np.float64),
                               i1.astype(np.float64), default_params(**prm))
    d = v - f64
    say("bf16 vs FP64 oracle: max %.3e rms %.3e (|v| max %.2f)" % (
        np.abs(d).max(), np.sqrt((d ** 2).mean()), np.abs(f64).max()))
    assert np.abs(d).max() < 5e-2


# ---------------------------------------------------------------------------
# the mixed-precision solve: bfloat16 images, binary16 gradients of i1, FP32
# stencil arithmetic, FP64 CG scalars (precision FLOF_PREC_IMG_BF16)

def test_mixed_images_level_solve_bit_exact():
    from ofblend_amd.abi import FLOF_PREC_IMG_BF16
    N, T = 14, 12
    i0 = O.drop_sdf(N, N, N, T, seed=7, nspheres=2, rmin=0.08, rmax=0.18)
    i1 = O.drop_sdf(N, N, N, T, seed=8, nspheres=2, rmin=0.08, rmax=0.18)
    say("mixed level solve", (N, N, N, T))
    s = M.Solver(gridSize=M.vec3(N, N, N), fourthDim=T)
    g0, g1 = _grids(s, i0, i1)
    for prec in (FLOF_PREC_IMG_BF16, FLOF_PREC_IMG_BF16 | FLOF_PSTORE_BF16):
        vel = s.create(M.Grid4Vec4)
        M.opticalFlow4d(vel, g0, g1, wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5,
                        precision=prec)
        it_gpu = M.getLastSolveInfo().summary()["levels"][0]["iters"]
        p = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, precision=prec)
        with O.sum_mode(1):
            ref, it, res, st = O.optical_flow4d(np.zeros(i0.shape + (4,), np.float32),
                                                i0, i1, p)
        say("precision", prec, "iterations gpu", it_gpu, "oracle", it)
        assert it_gpu == [it]
        assert np.array_equal(vel.to_numpy(), ref)


@pytest.mark.parametrize("multiStep,blur", [(1, 0.0), (3, 2.0)])
def test_mixed_images_multiscale_bit_exact_and_error(multiStep, blur):
    """24^3 x 20, the mixed-precision generator (2 spheres, radii 0.08N-0.18N): the
    mixed-precision solve equals the oracle that rounds at the same points
    (images stored as bfloat16 at the start, at each pyramid level and after
    every warp; gradients rounded to binary16); its distance to the FP64
    oracle is reported next to the FP32 solve's."""
    from ofblend_amd.abi import FLOF_PREC_IMG_BF16
    N, T = 24, 20
    i0 = O.drop_sdf(N, N, N, T, seed=9, nspheres=2, rmin=0.08, rmax=0.18)
    i1 = O.drop_sdf(N, N, N, T, seed=10, nspheres=2, rmin=0.08, rmax=0.18)
    say("mixed multiscale", (N, N, N, T), "multiStep", multiStep)
    s = M.Solver(gridSize=M.vec3(N, N, N), fourthDim=T)
    g0, g1 = _grids(s, i0, i1)
    vel = s.create(M.Grid4Vec4)
    prm = dict(O.P_PLAIN, minGridSize=8)
    prm.update(multiStep=multiStep, postVelBlur=blur)
    err = M.opticalFlowMultiscale4d(vel=vel, i0=g0, i1=g1, precision=FLOF_PREC_IMG_BF16,
                                    **prm)
    v = vel.to_numpy()
    with O.sum_mode(1):
        ref, rerr, _ = O.multiscale4d(np.zeros(v.shape, np.float32), i0, i1,
                                      default_params(precision=FLOF_PREC_IMG_BF16, **prm))
        f32, _, _ = O.multiscale4d(np.zeros(v.shape, np.float32), i0, i1,
                                   default_params(**prm))
    assert np.array_equal(v, ref)
    assert err == np.float32(rerr)
    f64, _, _ = O.multiscale4d(np.zeros(v.shape), i0.astype(np.float64),
                               i1.astype(np.float64), default_params(**prm))
    dm, d32 = v - f64, f32 - f64
    say("vs FP64 oracle: mixed max %.3e rms %.3e, FP32 max %.3e rms %.3e (|v| max %.2f)" % (
        np.abs(dm).max(), np.sqrt((dm ** 2).mean()), np.abs(d32).max(),
        np.sqrt((d32 ** 2).mean()), np.abs(f64).max()))


def test_mixed_images_golden_digest():
    """The mixed-precision solve of the 32^3 x 32 two-sphere pair (seeds 7,
    8) is bit-identical to the oracle's (digest recorded by
    tests/golden/make_digests.py, with both solves' distances to FP64)."""
    import hashlib
    import json
    import os
    from ofblend_amd.abi import FLOF_PREC_IMG_BF16
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "digests.json")
    d = json.load(open(path)).get("mixed32_l3")
    if d is None:
        pytest.skip("no golden digest for mixed32_l3")
    c = d["config"]
    N, T = c["N"], c["T"]
    gen = dict(nspheres=2, rmin=0.08, rmax=0.18)
    i0 = O.drop_sdf(N, N, N, T, seed=d["seeds"][0], **gen)
    i1 = O.drop_sdf(N, N, N, T, seed=d["seeds"][1], **gen)
    s = M.Solver(gridSize=M.vec3(N, N, N), fourthDim=T)
    g0, g1 = _grids(s, i0, i1)
    vel = s.create(M.Grid4Vec4)
    say("mixed golden", (N, N, N, T))
    err = M.opticalFlowMultiscale4d(vel=vel, i0=g0, i1=g1, precision=FLOF_PREC_IMG_BF16,
                                    minGridSize=c["minGridSize"], cgMaxIter=c["cgMaxIter"],
                                    **O.P_PLAIN)
    iters = [l["iters"][0] for l in M.getLastSolveInfo().summary()["levels"]]
    say("iters gpu", iters, "oracle", d["iters"], "distances", d["distances"])
    assert iters == d["iters"]
    assert hashlib.sha256(vel.to_numpy().tobytes()).hexdigest() == d["sha256"]
    assert err == np.float32(d["err"])
