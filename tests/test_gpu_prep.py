"""GPU parity of the data-preparation and file functions around the solver
(repeatFrame4d, extrap4dLsSimple, shiftForwGrid4d, Grid4d save/load,
loadPlaceGrid4d from .uni files, loadAdvectTimeSlice_OptInit/OptRun/Finish)
against the CPU oracle, through the C-ABI.  All results are bit-exact."""
import gzip
import struct

import numpy as np
import pytest

from oracle import oracle as O
from ofblend_amd import manta as M

pytestmark = pytest.mark.gpu

RNG = np.random.default_rng(77)
HDR = struct.Struct("<6i256sQ")


def say(*a):
    print(*a, flush=True)


def solver(X, Y, Z, T):
    return M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)


def real(s, a):
    g = s.create(M.Grid4Real)
    g.from_numpy(a)
    return g


def write_mnt2(path, a):
    """A 3D REAL grid (Z, Y, X) in the reference's MNT2 layout."""
    dz, dy, dx = a.shape
    with gzip.open(path, "wb", compresslevel=1) as f:
        f.write(b"MNT2")
        f.write(HDR.pack(dx, dy, dz, 1, 1, 4, b"frame", 0))
        f.write(np.ascontiguousarray(a, dtype=np.float32).tobytes())


@pytest.mark.parametrize("srct,rng,bnd", [(3.0, 2.0, 0), (2.25, 1.0, 1), (5.6, 3.0, 2)])
def test_repeat_frame_bit_exact(srct, rng, bnd):
    say("repeatFrame4d", srct, rng, bnd)
    T, Z, Y, X = 9, 7, 8, 10
    a = RNG.standard_normal((T, Z, Y, X)).astype(np.float32)
    g = real(solver(X, Y, Z, T), a)
    M.repeatFrame4d(g, srct, rng, bnd)
    assert np.array_equal(g.to_numpy(), O.repeat_frame4d(a, srct, rng, bnd))


@pytest.mark.parametrize("inside", [False, True])
def test_extrap_ls_simple_bit_exact(inside):
    say("extrap4dLsSimple inside", inside)
    N, T = 14, 6
    a = O.drop_sdf(N, N, N, T, seed=3, nspheres=2).astype(np.float32)
    g = real(solver(N, N, N, T), a)
    M.extrap4dLsSimple(g, distance=4, inside=inside)
    assert np.array_equal(g.to_numpy(), O.extrap4d_ls_simple(a, 4, inside))


def test_shift_forw_grid():
    say("shiftForwGrid4d")
    T, N = 10, 6
    a = RNG.standard_normal((T, N, N, N)).astype(np.float32)
    for good in (0, 3, 6, 9):
        g = real(solver(N, N, N, T), a)
        M.shiftForwGrid4d(g, good)
        ref = a.copy()
        shift = T - good
        s = 0
        while s < good:  # optflow4d.cpp:1607-1615
            ln = good - s if s + shift >= good else shift
            ref[s:s + ln] = ref[s + shift:s + shift + ln]
            s += shift
        assert np.array_equal(g.to_numpy(), ref)


def test_grid_save_load_round_trip(tmp_path):
    say("Grid4d save/load")
    s = solver(7, 6, 5, 4)
    a = RNG.standard_normal((4, 5, 6, 7)).astype(np.float32)
    v = RNG.standard_normal((4, 5, 6, 7, 4)).astype(np.float32)
    r = real(s, a)
    g4 = s.create(M.Grid4Vec4)
    g4.from_numpy(v)
    r.save(tmp_path / "r.uni")
    g4.save(tmp_path / "v.uni")
    r2 = s.create(M.Grid4Real)
    v2 = s.create(M.Grid4Vec4)
    r2.load(tmp_path / "r.uni")
    v2.load(tmp_path / "v.uni")
    assert np.array_equal(r2.to_numpy(), a) and np.array_equal(v2.to_numpy(), v)
    bad = solver(7, 6, 5, 3).create(M.Grid4Real)
    with pytest.raises(M.FlofError):
        bad.load(tmp_path / "r.uni")


def test_load_place_grid4d_from_files(tmp_path):
    """Frames written as reference MNT2 files are placed exactly like the
    in-memory oracle placement; only frames with weight are read."""
    say("loadPlaceGrid4d from files")
    N, T, F = 10, 8, 4
    frames = RNG.standard_normal((F, N, N, N)).astype(np.float32)
    for f in range(F):
        write_mnt2(tmp_path / ("fr_%04d.uni" % f), frames[f])
    pat = str(tmp_path / "fr_%04d.uni")
    s = solver(N, N, N, T)
    phi = s.create(M.Grid4Real)
    M.loadPlaceGrid4d(pat, phi, M.Vec4(0.), M.Vec4(1.), fileIdxStart=0, fileIdxEnd=F)
    ref = O.place_frames(phi.np_shape(), frames, fileIdxStart=0, fileIdxEnd=F)
    assert np.array_equal(phi.to_numpy(), ref)
    # rescaled level-set values with an iso offset, repeated first frame
    kw = dict(fileIdxStart=0, fileIdxEnd=F, rescaleSdfValues=True, sdfIsoOff=0.5,
              repeatStartFrame=0.25)
    M.loadPlaceGrid4d(pat, phi, M.Vec4(0.), M.Vec4(1.), **kw)
    ref = O.place_frames(phi.np_shape(), frames, **kw)
    assert np.array_equal(phi.to_numpy(), ref)
    # debugSkipLoad stops before the missing files are needed
    (tmp_path / "fr_0003.uni").unlink()
    M.loadPlaceGrid4d(pat, phi, M.Vec4(0.), M.Vec4(1.), fileIdxStart=0, fileIdxEnd=F,
                      debugSkipLoad=3)
    ref = O.place_frames(phi.np_shape(), frames[:3], fileIdxStart=0, fileIdxEnd=F,
                         debugSkipLoad=3)
    assert np.array_equal(phi.to_numpy(), ref)
    with pytest.raises(M.FlofError):
        M.loadPlaceGrid4d(pat, phi, M.Vec4(0.), M.Vec4(1.), fileIdxStart=0, fileIdxEnd=F)


def test_advect_time_slice_from_file(tmp_path):
    """OptInit reads the M4T2 deformation; OptRun equals the oracle warp."""
    say("loadAdvectTimeSlice_OptInit/OptRun/Finish")
    N, T, F = 10, 8, 2
    frames = O.drop_sdf(N, N, N, T, seed=2, scale=F).astype(np.float32)
    s = solver(N * F, N * F, N * F, T)
    phi = s.create(M.Grid4Real)
    fd = M.DeviceArray.from_numpy(frames)
    M.loadPlaceGrid4dFrames(fd, phi, M.Vec4(0.), M.Vec4(1.))
    ref_phi = O.place_frames(phi.np_shape(), frames)
    dv = (RNG.standard_normal((T, N, N, N, 4)) * 1.5).astype(np.float32)
    M.writeUni4d(tmp_path / "defo.uni", dv)
    M.loadAdvectTimeSlice_OptInit(7, str(tmp_path / "defo.uni"))
    dst0 = RNG.standard_normal((N * F,) * 3).astype(np.float32)
    try:
        for time in (0.0, 2.5, T - 1.0):
            dst = M.DeviceArray.from_numpy(dst0)
            M.loadAdvectTimeSlice_OptRun(7, "", dst, phi, time, 0.9, 1.0, M.Vec4(0.),
                                         M.Vec4(1.), M.Vec4(2, 2, 2, 1))
            ref = O.warp_slice(dv, ref_phi, dst0, time, blendAlpha=0.9,
                               defoFactor=(2, 2, 2, 1))
            assert np.array_equal(dst.to_numpy(), ref)
    finally:
        M.loadAdvectTimeSlice_Finish(7)
    with pytest.raises(M.FlofError):
        M.loadAdvectTimeSlice_OptRun(7, "", M.DeviceArray.from_numpy(dst0), phi, 0., 1.,
                                     1., M.Vec4(0.), M.Vec4(1.), M.Vec4(1.))


def test_advect_time_slice_non_cached(tmp_path):
    """loadAdvectTimeSlice (file read per call, bnd=1 interior, no border
    skip) equals the oracle's non-cached warp."""
    say("loadAdvectTimeSlice (non-cached)")
    N, T, F = 8, 6, 2
    frames = O.drop_sdf(N, N, N, T, seed=4, scale=F).astype(np.float32)
    s = solver(N * F, N * F, N * F, T)
    phi = real(s, O.place_frames((T, N * F, N * F, N * F), frames))
    ref_phi = phi.to_numpy()
    dv = (RNG.standard_normal((T, N, N, N, 4)) * 1.2).astype(np.float32)
    M.writeUni4d(tmp_path / "defo.uni", dv)
    dst0 = RNG.standard_normal((N * F,) * 3).astype(np.float32)
    for time in (0.5, 3.0):
        dst = M.DeviceArray.from_numpy(dst0)
        M.loadAdvectTimeSlice(0, str(tmp_path / "defo.uni"), dst, phi, time, 0.8, 1.0,
                              M.Vec4(0.), M.Vec4(1.), M.Vec4(2, 2, 2, 1))
        ref = O.warp_slice(dv, ref_phi, dst0, time, blendAlpha=0.8,
                           defoFactor=(2, 2, 2, 1), optimized=False)
        assert np.array_equal(dst.to_numpy(), ref)


@pytest.mark.parametrize("nv,aligned", [(2, False), (3, False), (2, True)])
def test_advect_time_slice_defo_volumes(tmp_path, nv, aligned):
    """loadAdvectTimeSlice_OptInit(useDefoVols) / _OptAdd / _OptRun: the
    windowed blend of 2 or 3 deformation volumes equals the oracle bit for bit
    (grid entry point and file-based handle)."""
    say("defo volumes", nv, "aligned", aligned)
    N, T = 24, 12
    phi0 = RNG.standard_normal((T, N, N, N)).astype(np.float32)
    vols = [(RNG.standard_normal((T, N, N, N, 4)) * 0.6).astype(np.float32)
            for _ in range(nv)]
    s = solver(N, N, N, T)
    phi = real(s, phi0)
    gv = []
    for v in vols:
        g = s.create(M.Grid4Vec4)
        g.from_numpy(v)
        gv.append(g)
    dst0 = RNG.standard_normal((N, N, N)).astype(np.float32)
    kw = dict(blendAlpha=0.7, thirdAlpha=0.4, fourthAlpha=0.2 if nv == 3 else 0.0)
    for time in (0.0, 5.3, T - 1.0):
        dst = M.DeviceArray.from_numpy(dst0)
        M.loadAdvectTimeSliceFromDefoVols(gv, dst, phi, time, kw["blendAlpha"],
                                          thirdAlpha=kw["thirdAlpha"],
                                          fourthAlpha=kw["fourthAlpha"], bordSkip=0,
                                          doAligned=aligned, partialLoadFac=0.3)
        ref = O.warp_slice_vols(vols, phi0, dst0, time, bordSkip=0, doAligned=aligned,
                                partialLoadFac=0.3, **kw)
        assert np.array_equal(dst.to_numpy(), ref), time
    # the same through the file-based handle
    for q, v in enumerate(vols):
        M.writeUni4d(tmp_path / ("d%d.uni" % q), v)
    M.loadAdvectTimeSlice_OptInit(3, str(tmp_path / "d0.uni"), useDefoVols=True,
                                  doAligned=aligned, partialLoadFac=0.3)
    try:
        for q in range(1, nv):
            M.loadAdvectTimeSlice_OptAdd(3, str(tmp_path / ("d%d.uni" % q)))
        dst = M.DeviceArray.from_numpy(dst0)
        M.loadAdvectTimeSlice_OptRun(3, "", dst, phi, 5.3, kw["blendAlpha"], 1.0, M.Vec4(0.),
                                     M.Vec4(1.), M.Vec4(1.), thirdAlpha=kw["thirdAlpha"],
                                     bordSkip=0, fourthAlpha=kw["fourthAlpha"])
        ref = O.warp_slice_vols(vols, phi0, dst0, 5.3, bordSkip=0, doAligned=aligned,
                                partialLoadFac=0.3, **kw)
        assert np.array_equal(dst.to_numpy(), ref)
    finally:
        M.loadAdvectTimeSlice_Finish(3)


def test_load_vel_rescaled(tmp_path):
    """loadVelRescaled (ofHelpers.py:91-100) as flof.py calls it: a
    deformation solved at a coarser resolution is resampled to the phi grid
    (interpolateGrid4dVec with the initOffsets defoScale) and scaled by
    defoFactor -- against the oracle's resample and an FP32 product."""
    say("loadVelRescaled")
    gsd = M.initRes(8, 1, 1.5, 1.)[1]
    gs = M.initRes(12, 1, 1.5, 1.)[1]
    (_, _, dOff, dScale, dFac) = M.initOffsets(gs, gsd, 0.1, 1.2, 1.)
    Xd, Yd, Zd, Td = [int(v) for v in gsd]
    v = (RNG.standard_normal((Td, Zd, Yd, Xd, 4)) * 2).astype(np.float32)
    M.writeUni4d(str(tmp_path / "pair_vel.uni"), v)
    sd = solver(Xd, Yd, Zd, Td)
    s = solver(*[int(c) for c in gs])
    vel = s.create(M.Grid4Vec4)
    M.loadVelRescaled(vel, str(tmp_path / "pair"), sd, dOff, dScale, dFac)
    ref = O.interpolate_grid4d(vel.np_shape()[:4], v, offset=dOff, scale=dScale)
    ref = (ref * np.asarray(dFac, dtype=np.float32)).astype(np.float32)
    assert np.array_equal(vel.to_numpy(), ref)
    assert np.abs(ref).max() > 1.0
    # a deformation solver of another size than the file is an error
    with pytest.raises(M.FlofError):
        M.loadVelRescaled(vel, str(tmp_path / "pair"), solver(Xd + 1, Yd, Zd, Td),
                          dOff, dScale, dFac)
