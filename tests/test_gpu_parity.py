"""GPU parity of the HIP path (through the C-ABI) against the CPU oracle.

The GPU library reproduces the reference's FP32 arithmetic operation by
operation; the only deliberate difference is that FP64 sums (CG dot
products, calcLsDiff) are correctly rounded instead of summed serially.  With
the oracle switched to the same summation (sum_mode(1)) results must be
bit-identical; against the serial order of the reference (sum_mode(0)) the
FP32 CG result differs by the reference's own sensitivity to summation order,
bounded below.
"""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from ofblend_amd import _lib
from ofblend_amd import manta as M
from ofblend_amd.abi import default_params

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(1234)


def solver(X, Y, Z, T):
    return M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)


def real(s, a):
    g = s.create(M.Grid4Real)
    g.from_numpy(a)
    return g


def vec4(s, a):
    g = s.create(M.Grid4Vec4)
    g.from_numpy(a)
    return g


def say(*a):
    print(*a, flush=True)


def test_grid_ops_known_answers():
    """tools/tests/test_0032_grid4dop.py: closed-form constants."""
    s = solver(10, 20, 30, 12)
    r1, r2, r3 = (s.create(M.Grid4Real) for _ in range(3))
    r1.setConst(1.0); r2.setConst(2.4); r3.setConst(9.6)
    r1.addConst(0.1); r2.multConst(0.5); r3.copyFrom(r1); r3.add(r2)
    r3.addScaled(r2, 0.5)
    assert np.abs(r1.to_numpy() - 1.1).max() <= 1e-7
    assert np.abs(r2.to_numpy() - 1.2).max() <= 1e-7
    assert np.abs(r3.to_numpy() - 2.9).max() <= 1e-7
    i1, i2, i3 = (s.create(M.Grid4Int) for _ in range(3))
    i1.setConst(123); i2.setConst(2); i3.setConst(9)
    i1.addConst(2); i2.multConst(3); i3.copyFrom(i1); i3.add(i2)
    i3.addScaled(i2, 2)
    assert (i1.to_numpy() == 125).all() and (i2.to_numpy() == 6).all()
    assert (i3.to_numpy() == 143).all()
    f1, f2, f3 = (s.create(M.Grid4Vec4) for _ in range(3))
    f1.setConst(M.Vec4(1.0)); f2.setConst(M.Vec4(1.0)); f3.setConst(M.Vec4(9.0))
    f1.addConst(M.Vec4(0.2)); f2.multConst(M.Vec4(0.5)); f3.copyFrom(f1)
    f3.add(f2); f3.addScaled(f2, M.Vec4(0.5))
    assert np.abs(f1.to_numpy() - 1.2).max() <= 5e-7
    assert np.abs(f2.to_numpy() - 0.5).max() <= 5e-7
    assert np.abs(f3.to_numpy() - 1.95).max() <= 5e-7
    # getMaxAbsValue of the Vec4 grids (comments of the reference test)
    assert abs(f1.getMaxAbsValue() - 2.40000009537) < 1e-6
    assert abs(f3.getMaxAbsValue() - 3.90000009537) < 1e-6


def test_advect_bit_exact():
    X, Y, Z, T = 13, 11, 9, 8
    s = solver(X, Y, Z, T)
    a = RNG.standard_normal((T, Z, Y, X)).astype(np.float32)
    av = RNG.standard_normal((T, Z, Y, X, 4)).astype(np.float32)
    v = (RNG.standard_normal((T, Z, Y, X, 4)) * 2.5).astype(np.float32)
    vel = vec4(s, v)
    g = real(s, a)
    M.advect4d(vel, g, dtFac=0.7)
    ref = O.advect4d(v, a, dt=1.0, dtFac=0.7)
    assert np.array_equal(g.to_numpy(), ref)
    gv = vec4(s, av)
    M.advect4d(vel, gv)
    assert np.array_equal(gv.to_numpy(), O.advect4d(v, av))


def test_resample_and_pyramid_bit_exact():
    s = solver(12, 10, 8, 6)
    src = RNG.standard_normal((6, 8, 10, 12)).astype(np.float32)
    sv = RNG.standard_normal((6, 8, 10, 12, 4)).astype(np.float32)
    g = real(s, src)
    gv = vec4(s, sv)
    for shape in [(3, 4, 5, 6), (12, 16, 20, 24), (5, 7, 9, 11)]:
        T, Z, Y, X = shape
        t = solver(X, Y, Z, T)
        tg = t.create(M.Grid4Real)
        M.interpolateGrid4d(tg, g, offset=M.Vec4(0.3, -0.2, 0.1, 0.0),
                            scale=M.Vec4(1.1, 0.9, 1.0, 1.2))
        ref = O.interpolate_grid4d(shape, src, offset=(0.3, -0.2, 0.1, 0.0),
                                   scale=(1.1, 0.9, 1.0, 1.2))
        assert np.array_equal(tg.to_numpy(), ref)
        tv = t.create(M.Grid4Vec4)
        M.interpolateGrid4dVec(tv, gv)
        assert np.array_equal(tv.to_numpy(), O.interpolate_grid4d(shape, sv))


def test_interpol4d_chain():
    """tools/tests/test_0042_interpol4d.py: 20 -> 40 -> 80 -> 40 -> 20."""
    sizes = [20, 40, 80, 40, 20]
    sv = [solver(n, n, n, n) for n in sizes]
    rs, re = 20 * 0.3, 20 * 0.7
    d = sv[0].create(M.Grid4Real)
    M.setRegion4d(d, M.Vec4(rs), M.Vec4(re), 1.0)
    a = d.to_numpy()
    cur = d
    for s in sv[1:]:
        nxt = s.create(M.Grid4Real)
        M.interpolateGrid4d(nxt, cur)
        a = O.interpolate_grid4d(nxt.np_shape(), a)
        assert np.array_equal(nxt.to_numpy(), a)
        cur = nxt


def test_bounds_and_lsdiff_bit_exact():
    X, Y, Z, T = 10, 9, 8, 7
    s = solver(X, Y, Z, T)
    a = RNG.standard_normal((T, Z, Y, X)).astype(np.float32)
    b = RNG.standard_normal((T, Z, Y, X)).astype(np.float32) * 3
    g = real(s, a)
    g.setBoundNeumann(0)
    assert np.array_equal(g.to_numpy(), O.set_bound_neumann(a, 0))
    g = real(s, a)
    g.setBound(2.5, 1)
    assert np.array_equal(g.to_numpy(), O.set_bound(a, 2.5, 1))
    ga, gb = real(s, a), real(s, b)
    out = s.create(M.Grid4Real)
    for bnd in (0, 1, 2):
        e = M.calcLsDiff4d(ga, gb, out=out, correction=0.3, bnd=bnd)
        with O.sum_mode(1):
            ref, ro = O.ls_diff(a, b, correction=0.3, bnd=bnd, want_out=True)
        assert e == np.float32(ref)
        if bnd == 0:
            assert np.array_equal(out.to_numpy(), ro)


def test_drop_generator_matches_oracle():
    s = solver(24, 24, 24, 16)
    g = s.create(M.Grid4Real)
    M.dropSdf4d(g, seed=7, nspheres=2, rmin=0.08, rmax=0.18)
    ref = O.drop_sdf(24, 24, 24, 16, seed=7, nspheres=2, rmin=0.08, rmax=0.18)
    assert np.abs(g.to_numpy() - ref).max() <= 1e-6
    fr = M.DeviceArray((4, 32, 32, 32))
    M.dropFrames3d(fr, 16, 16, 16, 4, 2, seed=3)
    ref = O.drop_sdf(16, 16, 16, 4, seed=3, scale=2)
    assert np.abs(fr.to_numpy() - ref).max() <= 1e-6


def _drops(N, T, seeds=(1, 2)):
    return [O.drop_sdf(N, N, N, T, seed=sd) for sd in seeds]


def test_optical_flow4d_level_bit_exact():
    N, T = 14, 12
    i0, i1 = _drops(N, T)
    s = solver(N, N, N, T)
    vin = (RNG.standard_normal((T, N, N, N, 4)) * 0.5).astype(np.float32)
    vel = vec4(s, vin)
    rhsT = s.create(M.Grid4Real)
    M.opticalFlow4d(vel, real(s, i0), real(s, i1), rhsT=rhsT, wSmooth=0.1,
                    wEnergy=1e-4, cgAccuracy=1e-5)
    info = M.getLastSolveInfo().summary()
    p = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5)
    with O.sum_mode(1):
        ref, it, res, st, rhs = O.optical_flow4d(vin, i0, i1, p, want_rhs=True)
    say("level solve iterations gpu", info["levels"][0]["iters"], "oracle", it)
    assert info["levels"][0]["iters"] == [it]
    # ||r||_2 / ||b||_2 of the final residual: converged, so well below 1
    l2 = info["levels"][0]["residual_l2"][0]
    say("residual: signed-max ratio", info["levels"][0]["residual"][0], "l2 ratio", l2)
    assert 0.0 < l2 < 1e-2
    assert np.array_equal(rhsT.to_numpy(), rhs)
    assert np.array_equal(vel.to_numpy(), ref)


def _multiscale_case(N, T, mg, **kw):
    i0, i1 = _drops(N, T)
    s = solver(N, N, N, T)
    vel = s.create(M.Grid4Vec4)
    prm = dict(O.P_PLAIN, minGridSize=mg)
    prm.update(kw)
    err = M.opticalFlowMultiscale4d(vel=vel, i0=real(s, i0), i1=real(s, i1),
                                    **{k: v for k, v in prm.items()})
    info = M.getLastSolveInfo().summary()
    p = default_params(**prm)
    return vel.to_numpy(), err, info, (i0, i1, p)


def test_multiscale_small_bit_exact():
    v, err, info, (i0, i1, p) = _multiscale_case(16, 16, 4)
    with O.sum_mode(1):
        ref, rerr, rinfo = O.multiscale4d(np.zeros(v.shape, np.float32), i0, i1, p)
    r = rinfo.summary()
    say("iters gpu", [l["iters"] for l in info["levels"]],
        "oracle", [l["iters"] for l in r["levels"]])
    assert [l["iters"] for l in info["levels"]] == [l["iters"] for l in r["levels"]]
    assert np.array_equal(v, ref)
    assert err == np.float32(rerr)


def _digests():
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden",
                        "digests.json")
    return json.load(open(path)) if os.path.exists(path) else {}


@pytest.mark.parametrize("case", ["drop32_l3", "drop64_l3_cap1000"])
def test_multiscale_golden_digest(case):
    """The 32^4 and 64^4 drop-pair solves (3 levels), P_plain: the GPU
    flow field is bit-identical to the oracle's (correctly rounded sums),
    whose SHA-256 tests/golden/make_digests.py recorded; the digest file also
    records the serial-sum oracle's distance to it."""
    import hashlib
    d = _digests().get(case)
    if d is None:
        pytest.skip("no golden digest for %s" % case)
    c = d["config"]
    N = c["N"]
    say("case", case, "N", N)
    i0, i1 = _drops(N, N)
    s = solver(N, N, N, N)
    vel = s.create(M.Grid4Vec4)
    err = M.opticalFlowMultiscale4d(vel=vel, i0=real(s, i0), i1=real(s, i1),
                                    minGridSize=c["minGridSize"],
                                    cgMaxIter=c["cgMaxIter"], **O.P_PLAIN)
    info = M.getLastSolveInfo().summary()
    v = vel.to_numpy()
    iters = [l["iters"][0] for l in info["levels"]]
    say(case, "iters gpu", iters, "oracle", d["iters"], "err", err, d["err"],
        "serial-sum oracle distance", d["serial_sum_oracle"])
    assert iters == d["iters"]
    assert hashlib.sha256(v.tobytes()).hexdigest() == d["sha256"]
    assert err == np.float32(d["err"])
    if "distances" in d:
        # the flow is the sum_mode-1 oracle's bit for bit, so the recorded
        # distances are the GPU's: to the reference's serial order and, for
        # both, to the FP64 oracle
        dd = d["distances"]
        say(case, "max|gpu - serial order|", dd["gpu_vs_serial"]["max_abs_diff"],
            "max|serial - fp64|", dd["serial_vs_fp64"]["max_abs_diff"],
            "max|gpu - fp64|", dd["gpu_vs_fp64"]["max_abs_diff"])
        if d["fp64_oracle"]["iters"] == d["iters"]:  # every level converged alike
            assert dd["gpu_vs_serial"]["max_abs_diff"] <= dd["serial_vs_fp64"]["max_abs_diff"]
            assert dd["gpu_vs_fp64"]["rms_diff"] <= 1.1 * dd["serial_vs_fp64"]["rms_diff"]
    if "warp" in d:
        _golden_warp(d, vel, N)


def _golden_warp(d, vel, N):
    """The frame warp of the golden case on the GPU: the drops as N frames at
    twice the resolution, placed 1:1 onto t, all N frames warped by the
    solved flow in one launch (alpha 1, defoFactor (2, 2, 2, 1), zeroed
    frames); the first, middle and last frame match the oracle's digest, and
    the middle one the oracle's warp of the GPU flow directly."""
    import hashlib
    w = d["warp"]
    F = w["scale"]
    frames = O.drop_sdf(N, N, N, N, seed=1, scale=F)
    sp = solver(N * F, N * F, N * F, N)
    phi = sp.create(M.Grid4Real)
    M.loadPlaceGrid4dFrames(M.DeviceArray.from_numpy(frames), phi, M.Vec4(0.), M.Vec4(1.))
    del frames
    dst = M.DeviceArray.from_numpy(np.zeros((N, N * F, N * F, N * F), np.float32))
    f4 = lambda *c: (C.c_float * 4)(*c)
    L = _lib.lib()
    say("warp", N, "frames of", N * F, "^3")
    _lib.check(L.flof_advect_time_slices_opt(
        vel.h, dst.ptr, N * F, N * F, N * F, N, 0.0, 1.0, phi.h, 1.0, 1.0,
        f4(0, 0, 0, 0), f4(1, 1, 1, 1), f4(2, 2, 2, 1), 1, 1.0), vel.ctx)
    out = dst.to_numpy()
    sel = np.stack([out[t] for t in w["frames"]])
    assert hashlib.sha256(sel.tobytes()).hexdigest() == w["sha256"]
    mid = w["frames"][1]
    ref = O.warp_slice(vel.to_numpy(), phi.to_numpy(), np.zeros(out.shape[1:], np.float32),
                       float(mid), blendAlpha=1.0, defoFactor=(2, 2, 2, 1))
    assert np.array_equal(out[mid], ref)
    say("warp max|gpu - fp64|", w["gpu_vs_fp64"]["max_abs_diff"],
        "max|serial - fp64|", w["serial_vs_fp64"]["max_abs_diff"])


@pytest.mark.parametrize("N", [16, 24])
def test_parity_against_reference_order(N):
    """The reference sums its CG dot products serially in FP64
    (optflow4d.cpp:234-241); the GPU sums them correctly rounded.  Against
    the oracle in the reference's order, the GPU flow is closer than that
    order's own FP32 flow is to the FP64 oracle, and the GPU's distance to
    the FP64 oracle matches the reference order's (rms within 10%)."""
    v, err, info, (i0, i1, p) = _multiscale_case(N, N, N // 4)
    with O.sum_mode(0):
        ser, _, sinfo = O.multiscale4d(np.zeros(v.shape, np.float32), i0, i1, p)
        d0 = O.drop_sdf(N, N, N, N, seed=1, dtype=np.float64)
        d1 = O.drop_sdf(N, N, N, N, seed=2, dtype=np.float64)
        v64, _, _ = O.multiscale4d(np.zeros(v.shape, np.float64), d0, d1, p)
    gs = np.abs(v.astype(np.float64) - ser).max()
    s64 = np.abs(ser - v64).max()
    g64 = np.abs(v - v64)
    say(N, "max|gpu - serial|", gs, "max|serial - fp64|", s64, "max|gpu - fp64|", g64.max())
    assert gs <= s64
    rms = lambda a: float(np.sqrt((a ** 2).mean()))
    assert rms(g64) <= 1.1 * rms(ser - v64)


def test_frames_place_and_warp_bit_exact():
    N, T, F = 12, 10, 2
    frames = O.drop_sdf(N, N, N, T, seed=1, scale=F)
    s = solver(N * F, N * F, N * F, T)
    phi = s.create(M.Grid4Real)
    fd = M.DeviceArray.from_numpy(frames)
    M.loadPlaceGrid4dFrames(fd, phi, M.Vec4(0.), M.Vec4(1.))
    ref_phi = O.place_frames(phi.np_shape(), frames)
    assert np.array_equal(phi.to_numpy(), ref_phi)
    sd = solver(N, N, N, T)
    dv = (RNG.standard_normal((T, N, N, N, 4)) * 1.5).astype(np.float32)
    defo = vec4(sd, dv)
    dst0 = RNG.standard_normal((N * F, N * F, N * F)).astype(np.float32)
    for time in (0.0, 3.0, 4.6, T - 1.0):
        dst = M.DeviceArray.from_numpy(dst0)
        M.loadAdvectTimeSliceFromDefo(defo, dst, phi, time, 0.8,
                                      defoFactor=M.Vec4(2, 2, 2, 1))
        ref = O.warp_slice(dv, ref_phi, dst0, time, blendAlpha=0.8,
                           defoFactor=(2, 2, 2, 1))
        assert np.array_equal(dst.to_numpy(), ref)


@pytest.mark.parametrize("sigma,it,vec", [(1.0, 1, True), (2.0, 1, True), (0.3, 2, False),
                                          (1.6, 1, False)])
def test_gaussian_blur_bit_exact(sigma, it, vec):
    """The (2s+1)^4 window summed in the reference's order, weights from the
    host's exp: bit-identical to the oracle (Vec4 and Real grids, 1 and 2
    iterations, windows clipped at the grid's faces)."""
    X, Y, Z, T = 11, 10, 9, 8
    s = solver(X, Y, Z, T)
    shp = (T, Z, Y, X, 4) if vec else (T, Z, Y, X)
    a = RNG.standard_normal(shp).astype(np.float32)
    g = vec4(s, a) if vec else real(s, a)
    say("gaussian blur", sigma, it, vec)
    M.gaussianBlur4d(g, sigma, it)
    ref = O.gaussian_blur(a, sigma, it)
    assert np.array_equal(g.to_numpy(), ref)


def test_multiscale_flof_params_bit_exact():
    """flof.py's parameters (multiStep 3, blur 4, final projection, border
    0.1) on a small grid: bit-identical to the oracle (the blur sums the 4D
    window in the reference's order)."""
    N, T = 24, 24
    i0, i1 = _drops(N, T)
    s = solver(N, N, N, T)
    vel = s.create(M.Grid4Vec4)
    prm = dict(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-3, postVelBlur=4,
               cfl=999, multiStep=3, minGridSize=10, doFinalProject=True,
               resetBndWidth=0.1)
    err = M.opticalFlowMultiscale4d(vel=vel, i0=real(s, i0), i1=real(s, i1),
                                    **prm)
    with O.sum_mode(1):
        ref, rerr, rinfo = O.multiscale4d(
            np.zeros((T, N, N, N, 4), np.float32), i0, i1, default_params(**prm))
    d = np.abs(vel.to_numpy() - ref)
    say("P_flof 24^4: max diff", d.max(), "rms", np.sqrt((d ** 2).mean()),
        "err gpu", err, "oracle", rerr)
    assert np.array_equal(vel.to_numpy(), ref)
    assert err == np.float32(rerr)


def test_warm_start_bit_exact_and_fallback():
    """Warm start (extension): the guess w = (warm - incoming flow) of every
    level is scaled by gamma = b.w / w.A w and kept only when
    ||b - gamma A w|| <= ||b||.  Bit-exact against the oracle with the same
    guess; a guess that fails the test (NaN) falls back to the cold start
    (identical to the plain solve)."""
    N, T, mg = 16, 16, 4
    i0, i1 = _drops(N, T)
    s = solver(N, N, N, T)
    g0, g1 = real(s, i0), real(s, i1)
    prm = dict(O.P_PLAIN, minGridSize=mg)
    cold = s.create(M.Grid4Vec4)
    M.opticalFlowMultiscale4d(vel=cold, i0=g0, i1=g1, **prm)
    cold_info = M.getLastSolveInfo().summary()
    vcold = cold.to_numpy()
    # a perturbed converged flow as the guess
    guess = (vcold + RNG.standard_normal(vcold.shape).astype(np.float32) * 0.05)
    w = vec4(s, guess)
    vel = s.create(M.Grid4Vec4)
    M.opticalFlowMultiscale4d(vel=vel, i0=g0, i1=g1, warm=w, **prm)
    info = M.getLastSolveInfo().summary()
    with O.sum_mode(1):
        ref, _, rinfo = O.multiscale4d(np.zeros_like(vcold), i0, i1,
                                       default_params(**prm), warm=guess)
    r = rinfo.summary()
    say("warm iters", [l["iters"] for l in info["levels"]], "oracle",
        [l["iters"] for l in r["levels"]], "cold", [l["iters"] for l in cold_info["levels"]],
        "warm used", [l["warm_used"] for l in info["levels"]])
    assert [l["iters"] for l in info["levels"]] == [l["iters"] for l in r["levels"]]
    assert [l["warm_used"] for l in info["levels"]] == [l["warm_used"] for l in r["levels"]]
    assert np.array_equal(vel.to_numpy(), ref)
    assert info["levels"][0]["iters"][0] < cold_info["levels"][0]["iters"][0]
    # a wild guess is scaled down to almost nothing: still bit-exact
    wild = (RNG.standard_normal(vcold.shape) * 1e3).astype(np.float32)
    vb = s.create(M.Grid4Vec4)
    M.opticalFlowMultiscale4d(vel=vb, i0=g0, i1=g1, warm=vec4(s, wild), **prm)
    with O.sum_mode(1):
        refw, _, _ = O.multiscale4d(np.zeros_like(vcold), i0, i1, default_params(**prm),
                                    warm=wild)
    assert np.array_equal(vb.to_numpy(), refw)
    # a NaN guess fails the residual test on every level: the cold solve
    nan = np.full(vcold.shape, np.nan, np.float32)
    vn = s.create(M.Grid4Vec4)
    M.opticalFlowMultiscale4d(vel=vn, i0=g0, i1=g1, warm=vec4(s, nan), **prm)
    ninfo = M.getLastSolveInfo().summary()
    assert [l["warm_used"] for l in ninfo["levels"]] == [0] * len(ninfo["levels"])
    assert np.array_equal(vn.to_numpy(), vcold)


def test_multiscale_ragged_sizes_bit_exact():
    """Uneven X, Y, Z, T (odd halvings in the pyramid) and the early exits:
    identical inputs give a zero right-hand side (status 3, no iterations)."""
    X, Y, Z, T = 21, 18, 16, 13
    a = O.drop_sdf(X, Y, Z, T, seed=41)
    b = O.drop_sdf(X, Y, Z, T, seed=42)
    s = solver(X, Y, Z, T)
    prm = dict(O.P_PLAIN, minGridSize=6)
    vel = s.create(M.Grid4Vec4)
    err = M.opticalFlowMultiscale4d(vel=vel, i0=real(s, a), i1=real(s, b), **prm)
    info = M.getLastSolveInfo().summary()
    with O.sum_mode(1):
        ref, rerr, rinfo = O.multiscale4d(np.zeros((T, Z, Y, X, 4), np.float32), a, b,
                                          default_params(**prm))
    say("ragged dims", [l["dims"] for l in info["levels"]], "iters",
        [l["iters"] for l in info["levels"]])
    assert [l["dims"] for l in info["levels"]] == [l["dims"] for l in rinfo.summary()["levels"]]
    assert [l["iters"] for l in info["levels"]] == [l["iters"] for l in rinfo.summary()["levels"]]
    assert np.array_equal(vel.to_numpy(), ref)
    assert err == np.float32(rerr)
    v2 = s.create(M.Grid4Vec4)
    M.opticalFlow4d(v2, real(s, a), real(s, a), wSmooth=0.1, wEnergy=1e-4)
    lv = M.getLastSolveInfo().summary()["levels"][0]
    assert lv["status"] == [3] and lv["iters"] == [0]
    assert not v2.to_numpy().any()


@pytest.mark.parametrize("kind", ["lds", "tmarch", "qarray"])
def test_stencil_kernels_bit_exact(kind):
    """Every form of the PCG stencil passes (LDS tiles of 64x4x4 cells, the
    t-marching kernels reading neighbours through the caches, the streaming
    kernels with a stored q) gives the oracle's result, on a level whose
    sizes leave every tile ragged (X > 64, Y, Z not multiples of 4) and with
    an incoming flow; and with the bfloat16 direction storage the two fused
    forms agree bit for bit."""
    from ofblend_amd import _lib
    from ofblend_amd.abi import (FLOF_PSTORE_BF16, FLOF_STENCIL_AUTO, FLOF_STENCIL_LDS,
                                 FLOF_STENCIL_QARRAY, FLOF_STENCIL_TMARCH)
    X, Y, Z, T = 70, 9, 10, 11
    a = O.drop_sdf(X, Y, Z, T, seed=51)
    b = O.drop_sdf(X, Y, Z, T, seed=52)
    vin = (np.random.default_rng(53).standard_normal((T, Z, Y, X, 4)) * 0.5).astype(np.float32)
    L = _lib.lib()
    ctx = M.context()
    s = solver(X, Y, Z, T)
    ga, gb = real(s, a), real(s, b)
    say("stencil kernels", kind, (X, Y, Z, T))
    out = {}
    try:
        for k, code in (("lds", FLOF_STENCIL_LDS), ("tmarch", FLOF_STENCIL_TMARCH),
                        ("qarray", FLOF_STENCIL_QARRAY)):
            assert L.flof_ctx_set_stencil(ctx, code) == 0
            for prec in ((0,) if k == "qarray" else (0, FLOF_PSTORE_BF16)):
                vel = vec4(s, vin)
                M.opticalFlow4d(vel, ga, gb, wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5,
                                precision=prec)
                out[k, prec] = (vel.to_numpy(), M.getLastSolveInfo().summary())
    finally:
        L.flof_ctx_set_stencil(ctx, FLOF_STENCIL_AUTO)
    assert L.flof_ctx_set_stencil(ctx, 7) != 0
    p = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5)
    with O.sum_mode(1):
        ref, it, res, st, rhs = O.optical_flow4d(vin, a, b, p, want_rhs=True)
    v, info = out[kind, 0]
    say("iterations", info["levels"][0]["iters"], "oracle", it)
    assert info["levels"][0]["iters"] == [it]
    assert np.array_equal(v, ref)
    assert np.array_equal(out["lds", FLOF_PSTORE_BF16][0], out["tmarch", FLOF_PSTORE_BF16][0])


@pytest.mark.parametrize("cap", [1, 2, 5, 6, 16, 17])
def test_cg_cap_at_both_iteration_parities(cap):
    """The x update of an even PCG iteration is deferred to the next one: a
    solve that stops at the iteration cap after an even or an odd iteration
    (caps 1, 5, 17 / 2, 6, 16; 16 and 17 end at and just after the host
    loop's 16-iteration chunk) must still hold every x update, in the
    LDS-tiled and the t-marching form, bit for bit as the oracle with the
    same cap."""
    from ofblend_amd import _lib
    from ofblend_amd.abi import FLOF_STENCIL_AUTO, FLOF_STENCIL_LDS, FLOF_STENCIL_TMARCH
    X, Y, Z, T = 70, 9, 10, 11
    a = O.drop_sdf(X, Y, Z, T, seed=61)
    b = O.drop_sdf(X, Y, Z, T, seed=62)
    L = _lib.lib()
    ctx = M.context()
    s = solver(X, Y, Z, T)
    ga, gb = real(s, a), real(s, b)
    p = default_params(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, cgMaxIter=cap)
    zero = np.zeros((T, Z, Y, X, 4), np.float32)
    with O.sum_mode(1):
        ref, it, res, st = O.optical_flow4d(zero, a, b, p)
    assert it == cap and st == 1
    try:
        for code in (FLOF_STENCIL_LDS, FLOF_STENCIL_TMARCH):
            assert L.flof_ctx_set_stencil(ctx, code) == 0
            vel = s.create(M.Grid4Vec4)
            say("cap", cap, "stencil", code, (X, Y, Z, T))
            M.opticalFlow4d(vel, ga, gb, wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5,
                            cgMaxIter=cap)
            info = M.getLastSolveInfo().summary()
            assert info["levels"][0]["iters"] == [cap]
            assert np.array_equal(vel.to_numpy(), ref)
    finally:
        L.flof_ctx_set_stencil(ctx, FLOF_STENCIL_AUTO)


def test_stencil_kernels_degenerate_shapes():
    """The LDS-tiled kernels (forced) and the streaming q-array kernels on
    shapes with at most a few cells in some dimension, tiles sticking out on
    every side and one-step t segments: the same flow and iterations as the
    t-marching kernels."""
    from ofblend_amd import _lib
    from ofblend_amd.abi import (FLOF_STENCIL_AUTO, FLOF_STENCIL_LDS, FLOF_STENCIL_QARRAY,
                                 FLOF_STENCIL_TMARCH)
    L = _lib.lib()
    ctx = M.context()
    try:
        for n, (X, Y, Z, T) in enumerate([(3, 2, 1, 5), (65, 5, 5, 2), (8, 8, 8, 1),
                                          (130, 3, 6, 4), (5, 9, 3, 17)]):
            say("degenerate shape", (X, Y, Z, T))
            a = O.drop_sdf(X, Y, Z, T, seed=70 + 2 * n)
            b = O.drop_sdf(X, Y, Z, T, seed=71 + 2 * n)
            s = solver(X, Y, Z, T)
            ga, gb = real(s, a), real(s, b)
            res = []
            for code in (FLOF_STENCIL_TMARCH, FLOF_STENCIL_LDS, FLOF_STENCIL_QARRAY):
                assert L.flof_ctx_set_stencil(ctx, code) == 0
                vel = s.create(M.Grid4Vec4)
                M.opticalFlow4d(vel, ga, gb, wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5)
                res.append((vel.to_numpy(), M.getLastSolveInfo().summary()["levels"][0]["iters"]))
            for r in res[1:]:
                assert r[1] == res[0][1]
                assert np.array_equal(r[0], res[0][0])
    finally:
        L.flof_ctx_set_stencil(ctx, FLOF_STENCIL_AUTO)


@pytest.mark.parametrize("where,bad", [("i1", np.nan), ("i0", np.nan), ("i1", np.inf),
                                       ("i0", np.inf)])
def test_non_finite_image_gives_the_references_zero_flow(where, bad):
    """A NaN or Inf in an input image makes the CG's first sigma NaN: the
    reference stops before the first iteration and keeps a zero increment
    (optflow4d.cpp:297-301, 508-514).  The GPU level solve returns the
    oracle's flow bit for bit, status 2 and 0 iterations.  (A NaN that
    appeared only later would end the GPU's CG at once with the zero
    increment, where the reference iterates to its cap first and then
    resets; the flow is the same.)"""
    N, T = 12, 10
    i0, i1 = _drops(N, T)
    imgs = {"i0": i0.copy(), "i1": i1.copy()}
    imgs[where][T // 2, N // 2, N // 2, N // 2 + 1] = bad
    say("non-finite", where, bad)
    s = solver(N, N, N, T)
    vin = (RNG.standard_normal((T, N, N, N, 4)) * 0.5).astype(np.float32)
    vel = vec4(s, vin)
    prm = dict(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, cgMaxIter=60)
    M.opticalFlow4d(vel, real(s, imgs["i0"]), real(s, imgs["i1"]), **prm)
    lv = M.getLastSolveInfo().summary()["levels"][0]
    p = default_params(**prm)
    with O.sum_mode(1):
        ref, it, res, st = O.optical_flow4d(vin, imgs["i0"], imgs["i1"], p)
    say("gpu iters", lv["iters"], "status", lv["status"], "oracle", it, st)
    got = vel.to_numpy()
    assert np.array_equal(got, ref, equal_nan=True)
    assert np.isfinite(got).all()
    assert lv["status"] == [st] and st == 2
    assert lv["iters"] == [it] == [0]
