"""GPU checks of the remaining entry points (through the C-ABI): slices,
the max-difference helpers, the CFL-substepped advection, the level-set
projection on its own, the all-frames warp, and the context utilities."""
import ctypes as C

import numpy as np
import pytest

from oracle import oracle as O
from ofblend_amd import _lib
from ofblend_amd import manta as M

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_slice_and_place_overlap_and_out_of_range():
    """getSliceFrom4d / placeGrid3d (grid4d.cpp:486-524): the overlap of the
    two grids is copied, a time outside the 4D grid changes nothing."""
    rng = np.random.default_rng(7)
    src = rng.standard_normal((5, 6, 7, 8)).astype(np.float32)
    s = M.Solver(gridSize=M.vec3(8, 7, 6), fourthDim=5)
    g = s.create(M.Grid4Real)
    g.from_numpy(src)
    d0 = rng.standard_normal((5, 9, 6)).astype(np.float32)
    d = M.DeviceArray.from_numpy(d0)
    M.getSliceFrom4d(g, 3, d)
    ref = d0.copy()
    ref[:5, :7, :6] = src[3, :5, :7, :6]
    assert np.array_equal(d.to_numpy(), ref)
    M.getSliceFrom4d(g, 7, d)
    assert np.array_equal(d.to_numpy(), ref)
    p0 = rng.standard_normal((7, 5, 9)).astype(np.float32)
    M.placeGrid3d(M.DeviceArray.from_numpy(p0), g, 2)
    r4 = src.copy()
    r4[2, :6, :5, :8] = p0[:6, :5, :8]
    assert np.array_equal(g.to_numpy(), r4)
    M.placeGrid3d(M.DeviceArray.from_numpy(p0), g, 9)
    assert np.array_equal(g.to_numpy(), r4)


def test_max_diff_real_and_vec4():
    """grid4dMaxDiff: max |a - b| of the FP32 differences; the Vec4 form sums
    the component differences in double (grid4d.cpp:419-464); both return
    the reference's Real."""
    rng = np.random.default_rng(8)
    s = M.Solver(gridSize=M.vec3(9, 8, 7), fourthDim=6)
    a, b = (rng.standard_normal((6, 7, 8, 9)).astype(np.float32) for _ in range(2))
    ga, gb = s.create(M.Grid4Real), s.create(M.Grid4Real)
    ga.from_numpy(a)
    gb.from_numpy(b)
    assert M.grid4dMaxDiff(ga, gb) == float(np.abs(a - b).max())
    va, vb = (rng.standard_normal((6, 7, 8, 9, 4)).astype(np.float32) for _ in range(2))
    gva, gvb = s.create(M.Grid4Vec4), s.create(M.Grid4Vec4)
    gva.from_numpy(va)
    gvb.from_numpy(vb)
    d = np.abs(va.astype(np.float64) - vb.astype(np.float64))
    ref = (((d[..., 0] + d[..., 1]) + d[..., 2]) + d[..., 3]).max()
    assert M.grid4dMaxDiffVec4(gva, gvb) == float(np.float32(ref))


@pytest.mark.parametrize("cfl", [999.0, 0.7])
def test_advect_cfl_bit_exact(cfl):
    """advectCflHelper4d (optflow4d.cpp:2170-2180): with cfl 0.7 the flow is
    split into substeps; the GPU equals the oracle bit for bit."""
    X, Y, Z, T = 11, 10, 9, 8
    rng = np.random.default_rng(9)
    vel = (rng.standard_normal((T, Z, Y, X, 4)) * 1.3).astype(np.float32)
    grid = rng.standard_normal((T, Z, Y, X)).astype(np.float32)
    s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
    gv, gg = s.create(M.Grid4Vec4), s.create(M.Grid4Real)
    gv.from_numpy(vel)
    gg.from_numpy(grid)
    L = _lib.lib()
    say("advect cfl", cfl)
    _lib.check(L.flof_advect_cfl4d(gv.h, gg.h, cfl, 1.0, 1.0), gv.ctx)
    ref = O.advect_cfl4d(vel, grid, cfl=cfl, velFactor=1.0, dt=1.0)
    assert np.array_equal(gg.to_numpy(), ref)


@pytest.mark.parametrize("thresh,maxIter", [(4.0, 40), (1.5, 60), (4.0, 0), (1.0, 7),
                                            (0.3, 45)])
def test_corr_vels_of4d_far_cells_bit_exact(thresh, maxIter):
    """The projection's line search on level sets whose offsets span 0 to
    about 20 cells: cells that start out of reach of threshPhi, cells whose
    search runs away, cells that end just inside or outside threshPhi, both
    initial steps (maxIter below and above 50) and no search at all
    (maxIter 0): the field and vel equal the oracle's bit for bit."""
    X, Y, Z, T = 24, 11, 10, 9
    rng = np.random.default_rng(21)
    t, z, y, x = np.meshgrid(*[np.arange(n) + 0.5 for n in (T, Z, Y, X)], indexing="ij")
    base = np.sqrt((x - 9.0) ** 2 + (y - 5.0) ** 2 + (z - 4.5) ** 2) - 3.0 + 0.2 * t
    phiOrg = (base + 0.05 * rng.standard_normal(base.shape)).astype(np.float32)
    phiTgt = (2.5 * base - 2.3 + 0.4 * np.sin(0.9 * x) + 0.05 *
              rng.standard_normal(base.shape)).astype(np.float32)
    vel0 = (rng.standard_normal((T, Z, Y, X, 4)) * 0.7).astype(np.float32)
    s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
    dst, vel = s.create(M.Grid4Vec4), s.create(M.Grid4Vec4)
    vel.from_numpy(vel0)
    po, pt = s.create(M.Grid4Real), s.create(M.Grid4Real)
    po.from_numpy(phiOrg)
    pt.from_numpy(phiTgt)
    say("corrVelsOf4d far cells", (X, Y, Z, T), "thresh", thresh, "maxIter", maxIter)
    M.corrVelsOf4d(dst, vel, po, po, pt, threshPhi=thresh, postVelBlur=0.0, maxIter=maxIter)
    rd, rv = O.corr_vels_of4d(vel0, phiOrg, phiTgt, threshPhi=thresh, postVelBlur=0.0,
                              maxIter=maxIter)
    assert np.abs(rd).max() > 0.0
    assert np.array_equal(dst.to_numpy(), rd)
    assert np.array_equal(vel.to_numpy(), rv)


@pytest.mark.parametrize("vx", [0.0, 1.0])
def test_corr_vels_of4d_planar_and_oracle(vx):
    """corrVelsOf4d on its own: planar level sets offset by a known distance
    give the offset along the normal, and the field equals the oracle's.
    The extrapolation runs max|vel| + 4 box passes: 4 for vx = 0 and 5 for
    vx = 1, so both parities of the passes' buffer alternation are covered."""
    X, Y, Z, T = 16, 10, 10, 10
    a, b = 7.3, 5.1
    x = np.arange(X, dtype=np.float64) + 0.5
    phiOrg = np.broadcast_to(x - a, (T, Z, Y, X)).astype(np.float32).copy()
    phiTgt = np.broadcast_to(x - b, (T, Z, Y, X)).astype(np.float32).copy()
    vel0 = np.zeros((T, Z, Y, X, 4), np.float32)
    vel0[..., 0] = vx
    s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
    dst, vel = s.create(M.Grid4Vec4), s.create(M.Grid4Vec4)
    vel.from_numpy(vel0)
    po, pc, pt = s.create(M.Grid4Real), s.create(M.Grid4Real), s.create(M.Grid4Real)
    po.from_numpy(phiOrg)
    pc.from_numpy(phiOrg)
    pt.from_numpy(phiTgt)
    M.corrVelsOf4d(dst, vel, po, pc, pt, threshPhi=4.0, postVelBlur=0.0, maxIter=40)
    rd, rv = O.corr_vels_of4d(vel0, phiOrg, phiTgt, threshPhi=4.0, postVelBlur=0.0,
                              maxIter=40)
    gd = dst.to_numpy()
    if vx == 0.0:
        core = gd[2:-2, 2:-2, 2:-2, 4:-4]
        assert np.abs(core[..., 1:]).max() == 0.0
        assert np.abs(core[..., 0] - (a - b)).max() <= 0.125
    d = np.abs(gd - rd).max()
    say("corrVelsOf4d gpu vs oracle max diff", d)
    assert d == 0.0
    assert np.array_equal(vel.to_numpy(), rv)


def test_advect_time_slices_opt_all_frames_bit_exact():
    """flof_advect_time_slices_opt: nframes output frames in one launch, each
    equal to the oracle's loadAdvectTimeSlice_OptRun at its time."""
    N, T, F, nf = 10, 8, 2, 5
    rng = np.random.default_rng(10)
    frames = O.drop_sdf(N, N, N, T, seed=3, scale=F)
    phi_np = O.place_frames((T, N * F, N * F, N * F), frames)
    s = M.Solver(gridSize=M.vec3(N * F, N * F, N * F), fourthDim=T)
    phi = s.create(M.Grid4Real)
    phi.from_numpy(phi_np)
    sd = M.Solver(gridSize=M.vec3(N, N, N), fourthDim=T)
    dv = (rng.standard_normal((T, N, N, N, 4)) * 1.2).astype(np.float32)
    defo = sd.create(M.Grid4Vec4)
    defo.from_numpy(dv)
    dst0 = rng.standard_normal((nf, N * F, N * F, N * F)).astype(np.float32)
    dst = M.DeviceArray.from_numpy(dst0)
    f4 = lambda *v: (C.c_float * 4)(*v)
    L = _lib.lib()
    t0, dt = 0.5, 1.25
    _lib.check(L.flof_advect_time_slices_opt(
        defo.h, dst.ptr, N * F, N * F, N * F, nf, t0, dt, phi.h, 0.9, 1.0,
        f4(0, 0, 0, 0), f4(1, 1, 1, 1), f4(2, 2, 2, 1), 1, 1.0), defo.ctx)
    out = dst.to_numpy()
    for f in range(nf):
        ref = O.warp_slice(dv, phi_np, dst0[f], t0 + f * dt, blendAlpha=0.9,
                           defoFactor=(2, 2, 2, 1))
        assert np.array_equal(out[f], ref), f


def test_context_utilities():
    """A second context on the same device: its stream, grids with their
    dims and device address, trimming the pool, destruction."""
    L = _lib.lib()
    ctx = C.c_void_p()
    assert L.flof_ctx_create(C.byref(ctx), 0) == 0
    try:
        assert L.flof_ctx_stream(ctx)
        g = C.c_void_p()
        assert L.flof_grid4d_create(ctx, 5, 4, 3, 2, 1, 0, C.byref(g)) == 0
        dims = (C.c_int * 4)()
        elem, dtype = C.c_int(), C.c_int()
        assert L.flof_grid4d_dims(g, dims, C.byref(elem), C.byref(dtype)) == 0
        assert list(dims) == [5, 4, 3, 2] and elem.value == 1 and dtype.value == 0
        assert L.flof_grid4d_device_ptr(g)
        assert L.flof_grid4d_destroy(g) == 0
        assert L.flof_ctx_trim(ctx) == 0
    finally:
        assert L.flof_ctx_destroy(ctx) == 0


def test_reductions_mult_and_vec4_region():
    """getMax / getMin / getMaxAbs (grid4d.cpp:266-290: Real values; Vec4 the
    square root of the max / min squared norm summed x, y, z, t in FP32),
    mult (grid4d.cpp:256-259) and setRegion4dVec4 (grid4d.cpp:466-483)."""
    rng = np.random.default_rng(12)
    shp = (5, 6, 7, 8)
    s = M.Solver(gridSize=M.vec3(8, 7, 6), fourthDim=5)
    a = rng.standard_normal(shp).astype(np.float32)
    g = s.create(M.Grid4Real)
    g.from_numpy(a)
    assert g.getMax() == float(a.max()) and g.getMin() == float(a.min())
    assert g.getMaxAbs() == float(max(abs(a.min()), abs(a.max())))
    v = rng.standard_normal(shp + (4,)).astype(np.float32)
    gv = s.create(M.Grid4Vec4)
    gv.from_numpy(v)
    f = np.float32
    ns = ((v[..., 0] * v[..., 0] + v[..., 1] * v[..., 1]) + v[..., 2] * v[..., 2]) \
        + v[..., 3] * v[..., 3]
    assert ns.dtype == np.float32
    assert gv.getMax() == float(np.sqrt(f(ns.max())))
    assert gv.getMin() == float(np.sqrt(f(ns.min())))
    assert gv.getMaxAbs() == gv.getMax()
    b = rng.standard_normal(shp).astype(np.float32)
    gb = s.create(M.Grid4Real)
    gb.from_numpy(b)
    g.mult(gb)
    assert np.array_equal(g.to_numpy(), a * b)
    M.setRegion4dVec4(gv, M.Vec4(1, 2, 1, 0), M.Vec4(5, 4, 3, 2), M.Vec4(1, 2, 3, 4))
    ref = v.copy()
    ref[0:3, 1:4, 2:5, 1:6] = np.float32([1, 2, 3, 4])
    assert np.array_equal(gv.to_numpy(), ref)


@pytest.mark.parametrize("vec", [True, False])
@pytest.mark.parametrize("X", [40, 70, 100])
@pytest.mark.parametrize("sigma", [1.0, 2.0])
def test_gaussian_blur_two_x_tiles_bit_exact(vec, X, sigma):
    """gaussianBlur4d at sigma 1 and 2 (the LDS column marches over x tiles
    with 3^4 and 5^4 windows) on grids 40, 70 and 100 wide (one ragged
    64-wide tile, three 32-wide tiles, two 64-wide tiles): the seams between
    the x tiles and the ragged last tile equal the oracle bit for bit, Vec4
    and Real."""
    Y, Z, T = 9, 10, 11
    rng = np.random.default_rng(12)
    shp = (T, Z, Y, X, 4) if vec else (T, Z, Y, X)
    a = rng.standard_normal(shp).astype(np.float32)
    s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
    g = s.create(M.Grid4Vec4 if vec else M.Grid4Real)
    g.from_numpy(a)
    say("gaussian blur", (X, Y, Z, T), "sigma", sigma, "vec", vec)
    M.gaussianBlur4d(g, sigma, 1)
    assert np.array_equal(g.to_numpy(), O.gaussian_blur(a, sigma, 1))


@pytest.mark.parametrize("shape", [(3, 6, 4, 5), (7, 3, 5, 3), (5, 5, 5, 6)])
def test_gaussian_blur_5x5_window_small_dims_bit_exact(shape):
    """The 5^4 window (sigma 2) on grids with 3 to 7 cells per dimension: the
    windows clipped on one side, on both sides ([-1, 1] where a dimension has
    3 cells) and the weight sums of every such clipping pattern equal the
    oracle bit for bit, Vec4 (two iterations) and Real."""
    X, Y, Z, T = shape
    rng = np.random.default_rng(14)
    s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
    say("gaussian blur 5^4 window", shape)
    a4 = rng.standard_normal((T, Z, Y, X, 4)).astype(np.float32)
    g4 = s.create(M.Grid4Vec4)
    g4.from_numpy(a4)
    M.gaussianBlur4d(g4, 2.0, 2)
    assert np.array_equal(g4.to_numpy(), O.gaussian_blur(a4, 2.0, 2))
    a1 = rng.standard_normal((T, Z, Y, X)).astype(np.float32)
    g1 = s.create(M.Grid4Real)
    g1.from_numpy(a1)
    M.gaussianBlur4d(g1, 1.7, 1)
    assert np.array_equal(g1.to_numpy(), O.gaussian_blur(a1, 1.7, 1))


@pytest.mark.parametrize("vx", [0.0, 1.0])
@pytest.mark.parametrize("X", [40, 70, 100])
def test_corr_vels_of4d_two_x_tiles_bit_exact(vx, X):
    """corrVelsOf4d on grids 40, 70 and 100 wide (the box passes on one
    ragged 64-wide x tile, three 32-wide tiles, two 64-wide tiles): 4 and 5
    extrapolation passes, random level sets, bit-exact to the oracle."""
    Y, Z, T = 10, 9, 8
    rng = np.random.default_rng(13)
    t, z, y, x = np.meshgrid(*[np.arange(s) + 0.5 for s in (T, Z, Y, X)], indexing="ij")
    base = x - 0.45 * X + 0.3 * np.sin(0.7 * y) + 0.2 * np.cos(0.5 * z + 0.3 * t)
    phiOrg = (base + 0.05 * rng.standard_normal(base.shape)).astype(np.float32)
    phiTgt = (base - 1.7 + 0.05 * rng.standard_normal(base.shape)).astype(np.float32)
    vel0 = np.zeros((T, Z, Y, X, 4), np.float32)
    vel0[..., 0] = vx
    s = M.Solver(gridSize=M.vec3(X, Y, Z), fourthDim=T)
    dst, vel = s.create(M.Grid4Vec4), s.create(M.Grid4Vec4)
    vel.from_numpy(vel0)
    po, pt = s.create(M.Grid4Real), s.create(M.Grid4Real)
    po.from_numpy(phiOrg)
    pt.from_numpy(phiTgt)
    say("corrVelsOf4d", (X, Y, Z, T), "vx", vx)
    M.corrVelsOf4d(dst, vel, po, po, pt, threshPhi=4.0, postVelBlur=0.0, maxIter=40)
    rd, rv = O.corr_vels_of4d(vel0, phiOrg, phiTgt, threshPhi=4.0, postVelBlur=0.0,
                              maxIter=40)
    assert np.abs(rd).max() > 0.5
    assert np.array_equal(dst.to_numpy(), rd)
    assert np.array_equal(vel.to_numpy(), rv)
