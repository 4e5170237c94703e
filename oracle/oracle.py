"""ctypes front-end of the CPU oracle (oracle/liborc_flof.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by __graft_entry__.smoke() and by
the CPU-baseline leg of bench.py, always as the checker -- never as the thing
measured or shipped.  The product path (ofblend_amd) never imports it.

Arrays use numpy layout: a Real grid is (T, Z, Y, X), a Vec4 grid is
(T, Z, Y, X, 4) -- the reference's memory order (source/grid4d.h:92-104,
x fastest, Vec4 as AoS).  float32 arrays select the FP32-faithful variant,
float64 arrays the FP64 (DOUBLEPRECISION) variant.

What pins each oracle function (tests named in brackets):
  sample4 / sample3 (interp4, interp3, util/vector4d.h:422-516,
      util/interpol.h:57-116) -- scipy.ndimage.map_coordinates(order=1,
      mode='nearest') on random fields and random positions incl. outside the
      grid; exactness on multilinear fields [test_oracle_interp.py]
  advect4d / advect_cfl4d (optflow4d.cpp:1275-1321,2170-2180) --
      map_coordinates backtrace; zero flow = identity on the interior with a
      zero outer layer; integer shifts [test_oracle_interp.py]
  interpolate_grid4d / pyramid_resample (grid4d.cpp:531-569) -- the
      reference's own resample chain test (tools/tests/test_0042_interpol4d.py)
      with brute-force separable expectations; restriction = 16-cell mean
      [test_oracle_interp.py]
  set_bound / set_bound_neumann (grid4d.cpp:355-412) -- brute force
      [test_oracle_interp.py]
  ls_diff (optflow4d.cpp:895-927) -- numpy brute force [test_oracle_solve.py]
  gaussian_blur (optflow4d.cpp:127-174) -- separable normalised Gaussian
      (textbook) with the bnd=1 / swap boundary rule [test_oracle_solve.py]
  optical_flow4d (optflow4d.cpp:361-553) -- scipy.sparse assembly of the
      4D system solved by spsolve in FP64, and an independent numpy PCG with
      the signed-max rule (same iteration count) [test_oracle_solve.py]
  multiscale4d (optflow4d.cpp:936-1173) -- an independent numpy/scipy
      composition of restrict / spsolve / prolong / map_coordinates warp
      [test_oracle_solve.py]
  corr_vels_of4d (optflow4d.cpp:651-802) -- planar level sets with a known
      offset (the line search must land within its 0.1 step) and marker /
      fill invariants [test_oracle_solve.py]
  place_frames / warp_slice (optflow4d.cpp:1445-1595,1951-2105) -- 1:1
      placement equals the frames' interior; zero defo = identity lookup;
      constant defo = exact shift; map_coordinates [test_oracle_frames.py]
  drop_sdf -- closed form numpy restatement of the stated generator
      [test_oracle_frames.py]
"""
import ctypes as C
import os
import sys

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(_HERE))
from ofblend_amd.abi import OfParams, SolveInfo, default_params  # noqa: E402

_LIB = None


def build():
    """Compile liborc_flof.so with the committed Makefile."""
    import subprocess
    subprocess.check_call(["make", "-s", "-C", _HERE])


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liborc_flof.so")
        if not os.path.exists(path):
            build()
        _LIB = C.CDLL(path)
        _declare(_LIB)
    return _LIB


_P = C.c_void_p
_I = C.c_int
_D = C.c_double
_I64 = C.c_int64


def _declare(L):
    L.orc_set_sum_mode.argtypes = [C.c_int]
    L.orc_set_sum_mode.restype = None
    L.orc_get_sum_mode.argtypes = []
    L.orc_get_sum_mode.restype = C.c_int
    L.orc_create_mesh.argtypes = [_P, _I, _I, _I, _P, _P, _P, _I64, _I64,
                                  C.POINTER(_I64), C.POINTER(_I64)]
    L.orc_create_mesh.restype = _I
    L.orc_vertex_normals.argtypes = [_P, _I64, _P, _I64, _P]
    L.orc_vertex_normals.restype = None
    for name in ("orc_round_bf16", "orc_round_f16"):
        getattr(L, name).argtypes = [C.c_float]
        getattr(L, name).restype = C.c_float
    for suf in ("_f32", "_f64"):
        sig = {
            "orc_sample4_real": ([_P, _I, _I, _I, _I, _P, _I64, _P], None),
            "orc_sample3_real": ([_P, _I, _I, _I, _P, _I64, _P], None),
            "orc_advect4d_real": ([_P, _P, _I, _I, _I, _I, _D, _D], None),
            "orc_advect4d_vec4": ([_P, _P, _I, _I, _I, _I, _D, _D], None),
            "orc_advect_cfl4d_real": ([_P, _P, _I, _I, _I, _I, _D, _D, _D],
                                      None),
            "orc_interpolate_grid4d": ([_P, _I, _I, _I, _I, _P, _I, _I, _I, _I,
                                        _I, _P, _P, _P], None),
            "orc_pyramid_resample": ([_P, _I, _I, _I, _I, _P, _I, _I, _I, _I,
                                      _I], None),
            "orc_set_bound_neumann_real": ([_P, _I, _I, _I, _I, _I], None),
            "orc_set_bound_real": ([_P, _I, _I, _I, _I, _D, _I], None),
            "orc_ls_diff": ([_P, _P, _P, _I, _I, _I, _I, _D, _I], _D),
            "orc_gaussian_blur": ([_P, _I, _I, _I, _I, _I, _D, _I], None),
            "orc_optical_flow4d": ([_P, _P, _P, _P, _I, _I, _I, _I, _P, _D, _P,
                                    _P, _P, _P], None),
            "orc_multiscale4d": ([_P, _P, _P, _P, _I, _I, _I, _I, _P, _D, _P,
                                  _P], _D),
            "orc_corr_vels_of4d": ([_P, _P, _P, _P, _I, _I, _I, _I, _D, _D, _D,
                                    _I, _D], None),
            "orc_place_frames": ([_P, _I, _I, _I, _I, _P, _I, _I, _I, _P, _P,
                                  _I, _I, _I, _D, _P, _D, _I, _D, _I, _D, _D],
                                 None),
            "orc_warp_slice": ([_P, _I, _I, _I, _I, _P, _I, _I, _I, _P, _I, _I,
                                _I, _I, _D, _D, _D, _P, _P, _P, _P, _D, _I, _I,
                                _D, _I], None),
            "orc_warp_slice_vols": ([_P, _I, _I, _I, _I, _I, _P, _I, _I, _I, _P,
                                     _I, _I, _I, _I, _D, _D, _D, _P, _P, _P, _P,
                                     _D, _I, _I, _D, _I, _D, _D, _D], _I),
            "orc_drop_sdf": ([_P, _I, _I, _I, _I, _I, C.c_uint64, _I, _D, _D,
                              _D], None),
            "orc_repeat_frame4d": ([_P, _I, _I, _I, _I, _D, _D, _I], None),
            "orc_extrap4d_ls_simple": ([_P, _I, _I, _I, _I, _I, _I], None),
        }
        for name, (args, res) in sig.items():
            f = getattr(L, name + suf)
            f.argtypes = args
            f.restype = res


def set_sum_mode(mode):
    """0: FP64 sums in the reference's serial order (default); 1: correctly
    rounded sums (the GPU library's order-independent summation)."""
    lib().orc_set_sum_mode(int(mode))


def round_bf16(x):
    """The oracle's float -> bfloat16 -> float rounding (search-direction
    storage precision 1)."""
    return lib().orc_round_bf16(float(x))


def round_f16(x):
    """The oracle's float -> binary16 -> float rounding (the gradients of the
    mixed-precision solve)."""
    return lib().orc_round_f16(float(x))


def get_sum_mode():
    return lib().orc_get_sum_mode()


class sum_mode:
    """Context manager selecting the oracle's FP64 summation mode."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.prev = get_sum_mode()
        set_sum_mode(self.mode)
        return self

    def __exit__(self, *a):
        set_sum_mode(self.prev)


def _suf(a):
    if a.dtype == np.float32:
        return "_f32"
    if a.dtype == np.float64:
        return "_f64"
    raise TypeError("oracle arrays must be float32 or float64")


def _ptr(a):
    assert a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(C.c_void_p)


def _dims(a):
    """(X, Y, Z, T) of a (T,Z,Y,X[,4]) array."""
    return a.shape[3], a.shape[2], a.shape[1], a.shape[0]


def _v4(v):
    return (C.c_double * 4)(*[float(x) for x in v])


def _fn(name, a):
    return getattr(lib(), name + _suf(a))


# ---------------------------------------------------------------------------
def sample4(data, pos):
    """interp4 of a Real grid at positions pos (n, 4) given as (x, y, z, t)."""
    data = np.ascontiguousarray(data)
    pos = np.ascontiguousarray(pos, dtype=data.dtype)
    out = np.empty(pos.shape[0], dtype=data.dtype)
    _fn("orc_sample4_real", data)(_ptr(data), *_dims(data), _ptr(pos),
                                  pos.shape[0], _ptr(out))
    return out


def sample3(data, pos):
    """interp3 of a 3D Real grid (Z, Y, X) at positions pos (n, 3)."""
    data = np.ascontiguousarray(data)
    pos = np.ascontiguousarray(pos, dtype=data.dtype)
    out = np.empty(pos.shape[0], dtype=data.dtype)
    Z, Y, X = data.shape
    _fn("orc_sample3_real", data)(_ptr(data), X, Y, Z, _ptr(pos),
                                  pos.shape[0], _ptr(out))
    return out


def advect4d(vel, grid, dt=1.0, dtFac=1.0):
    """advect4d (optflow4d.cpp:1292): returns the advected copy of grid."""
    vel = np.ascontiguousarray(vel, dtype=grid.dtype)
    g = np.array(grid, copy=True, order="C")
    if g.ndim == 5:
        _fn("orc_advect4d_vec4", g)(_ptr(vel), _ptr(g), *_dims(g), dt, dtFac)
    else:
        _fn("orc_advect4d_real", g)(_ptr(vel), _ptr(g), *_dims(g), dt, dtFac)
    return g


def advect_cfl4d(vel, grid, cfl=999.0, velFactor=1.0, dt=1.0):
    vel = np.ascontiguousarray(vel, dtype=grid.dtype)
    g = np.array(grid, copy=True, order="C")
    _fn("orc_advect_cfl4d_real", g)(_ptr(vel), _ptr(g), *_dims(g), cfl,
                                    velFactor, dt)
    return g


def interpolate_grid4d(target_shape, source, offset=(0, 0, 0, 0),
                       scale=(1, 1, 1, 1), size=(-1, -1, -1, -1)):
    """interpolateGrid4d[Vec] (grid4d.cpp:539-557) into a new target array of
    spatial shape target_shape = (T, Z, Y, X)."""
    source = np.ascontiguousarray(source)
    vec = source.ndim == 5
    shp = tuple(target_shape) + ((4,) if vec else ())
    tgt = np.zeros(shp, dtype=source.dtype)
    _fn("orc_interpolate_grid4d", source)(
        _ptr(tgt), *_dims(tgt), _ptr(source), *_dims(source), int(vec),
        _v4(offset), _v4(scale), _v4(size))
    return tgt


def pyramid_resample(target_shape, source):
    """interpolGridTempl (optflow4d.cpp:40-57)."""
    source = np.ascontiguousarray(source)
    vec = source.ndim == 5
    tgt = np.zeros(tuple(target_shape) + ((4,) if vec else ()),
                   dtype=source.dtype)
    _fn("orc_pyramid_resample", source)(_ptr(tgt), *_dims(tgt), _ptr(source),
                                        *_dims(source), int(vec))
    return tgt


def set_bound_neumann(grid, w=1):
    g = np.array(grid, copy=True, order="C")
    _fn("orc_set_bound_neumann_real", g)(_ptr(g), *_dims(g), w)
    return g


def set_bound(grid, value, w=1):
    g = np.array(grid, copy=True, order="C")
    _fn("orc_set_bound_real", g)(_ptr(g), *_dims(g), float(value), w)
    return g


def ls_diff(a, b, correction=1.0, bnd=0, want_out=False):
    """calcLsDiff4d (optflow4d.cpp:2132); returns the error (and the per-cell
    output grid when want_out)."""
    a = np.ascontiguousarray(a)
    b = np.ascontiguousarray(b, dtype=a.dtype)
    out = np.zeros_like(a) if want_out else None
    r = _fn("orc_ls_diff", a)(_ptr(a), _ptr(b),
                              _ptr(out) if want_out else None, *_dims(a),
                              correction, bnd)
    return (r, out) if want_out else r


def gaussian_blur(grid, sigma, it=1):
    """gaussianBlurGeneric (optflow4d.cpp:160-174)."""
    g = np.array(grid, copy=True, order="C")
    _fn("orc_gaussian_blur", g)(_ptr(g), *_dims(g), int(g.ndim == 5), sigma,
                                it)
    return g


def optical_flow4d(vel, i0, i1, params=None, dt=1.0, warm=None,
                   want_rhs=False):
    """opticalFlow4d (optflow4d.cpp:2110): returns (vel, iters, residual,
    status[, rhsT])."""
    p = params if params is not None else default_params()
    v = np.array(vel, copy=True, order="C")
    i0 = np.ascontiguousarray(i0, dtype=v.dtype)
    i1 = np.ascontiguousarray(i1, dtype=v.dtype)
    rhs = np.zeros_like(i0) if want_rhs else None
    w = np.ascontiguousarray(warm, dtype=v.dtype) if warm is not None else None
    it, res, st = C.c_int32(), C.c_double(), C.c_int32()
    _fn("orc_optical_flow4d", v)(
        _ptr(v), _ptr(i0), _ptr(i1), _ptr(rhs) if want_rhs else None,
        *_dims(i0), C.byref(p), dt, _ptr(w) if w is not None else None,
        C.byref(it), C.byref(res), C.byref(st))
    out = (v, it.value, res.value, st.value)
    return out + (rhs,) if want_rhs else out


def multiscale4d(vel, i0, i1, params=None, dt=1.0, warm=None):
    """opticalFlowMultiscale4d (optflow4d.cpp:2182): returns
    (vel, err, SolveInfo)."""
    p = params if params is not None else default_params()
    v = np.array(vel, copy=True, order="C")
    i0 = np.ascontiguousarray(i0, dtype=v.dtype)
    i1 = np.ascontiguousarray(i1, dtype=v.dtype)
    w = np.ascontiguousarray(warm, dtype=v.dtype) if warm is not None else None
    info = SolveInfo()
    err = _fn("orc_multiscale4d", v)(_ptr(v), _ptr(i0), _ptr(i1), None,
                                     *_dims(i0), C.byref(p), dt,
                                     _ptr(w) if w is not None else None,
                                     C.byref(info))
    return v, err, info


def corr_vels_of4d(vel, phiOrg, phiTarget, threshPhi=1e10, postVelBlur=0.0,
                   resetBndWidth=-1.0, maxIter=100, dt=1.0):
    """corrVelsOf4d (optflow4d.cpp:2121): returns (dst, vel)."""
    v = np.array(vel, copy=True, order="C")
    dst = np.zeros_like(v)
    a = np.ascontiguousarray(phiOrg, dtype=v.dtype)
    b = np.ascontiguousarray(phiTarget, dtype=v.dtype)
    _fn("orc_corr_vels_of4d", v)(_ptr(dst), _ptr(v), _ptr(a), _ptr(b),
                                 *_dims(a), threshPhi, postVelBlur,
                                 resetBndWidth, maxIter, dt)
    return dst, v


def place_frames(phi_shape, frames, offset=(0, 0, 0, 0), scale=(1, 1, 1, 1),
                 fileIdxStart=-1, fileIdxEnd=-1, debugSkipLoad=999999,
                 spread=1.0, overrideSize=(-1, -1, -1, -1),
                 overrideTimeOff=0.0, overrideGoodRegion=0, loadTimeScale=1.0,
                 rescaleSdfValues=False, sdfIsoOff=0.0, repeatStartFrame=0.0,
                 phi=None):
    """loadPlaceGrid4d (optflow4d.cpp:1464) with frames (F, sz, sy, sx) in
    memory; frame f is file index max(fileIdxStart,0) + f."""
    frames = np.ascontiguousarray(frames)
    if phi is None:
        phi = np.zeros(phi_shape, dtype=frames.dtype)
    else:
        phi = np.array(phi, copy=True, order="C")
    F, sz, sy, sx = frames.shape
    _fn("orc_place_frames", phi)(
        _ptr(phi), *_dims(phi), _ptr(frames), sx, sy, sz, _v4(offset),
        _v4(scale), fileIdxStart, fileIdxEnd, debugSkipLoad, spread,
        _v4(overrideSize), overrideTimeOff, overrideGoodRegion, loadTimeScale,
        int(rescaleSdfValues), sdfIsoOff, repeatStartFrame)
    return phi


def warp_slice(defo, phi, dst, time, blendAlpha=1.0, loadTimeScale=1.0,
               defoOffset=(0, 0, 0, 0), defoScale=(1, 1, 1, 1),
               defoFactor=(1, 1, 1, 1), overrideSize=(-1, -1, -1, -1),
               overrideTimeOff=0.0, zeroVel=False, bordSkip=1,
               defoAniFac=1.0, optimized=True):
    """loadAdvectTimeSlice_OptRun (optimized=True, optflow4d.cpp:1951) or
    loadAdvectTimeSlice (optimized=False, :1671) with the 4D defo in memory.
    dst (Z, Y, X) is updated in a copy that is returned."""
    defo = np.ascontiguousarray(defo)
    phi = np.ascontiguousarray(phi, dtype=defo.dtype)
    d = np.array(dst, copy=True, order="C", dtype=defo.dtype)
    dz, dy, dx = d.shape
    _fn("orc_warp_slice", defo)(
        _ptr(defo), *_dims(defo), _ptr(d), dx, dy, dz, _ptr(phi), *_dims(phi),
        time, blendAlpha, loadTimeScale, _v4(defoOffset), _v4(defoScale),
        _v4(defoFactor), _v4(overrideSize), overrideTimeOff, int(zeroVel),
        bordSkip, defoAniFac, int(optimized))
    return d


def warp_slice_vols(vols, phi, dst, time, blendAlpha=1.0, loadTimeScale=1.0,
                    defoOffset=(0, 0, 0, 0), defoScale=(1, 1, 1, 1),
                    defoFactor=(1, 1, 1, 1), overrideSize=(-1, -1, -1, -1),
                    overrideTimeOff=0.0, zeroVel=False, bordSkip=1, defoAniFac=1.0,
                    doAligned=False, partialLoadFac=0.2, thirdAlpha=0.0,
                    fourthAlpha=0.0):
    """loadAdvectTimeSlice_OptRun with deformation volumes (useDefoVols,
    optflow4d.cpp:1786-1870, 2013-2085): vols is a list of 2 or 3 defos
    (T, Z, Y, X, 4) of one size; returns the updated copy of dst."""
    vols = [np.ascontiguousarray(v) for v in vols]
    dtype = vols[0].dtype
    phi = np.ascontiguousarray(phi, dtype=dtype)
    d = np.array(dst, copy=True, order="C", dtype=dtype)
    dz, dy, dx = d.shape
    arr = (C.c_void_p * len(vols))(*[v.ctypes.data for v in vols])
    rc = _fn("orc_warp_slice_vols", vols[0])(
        arr, len(vols), *_dims(vols[0]), _ptr(d), dx, dy, dz, _ptr(phi), *_dims(phi),
        time, blendAlpha, loadTimeScale, _v4(defoOffset), _v4(defoScale),
        _v4(defoFactor), _v4(overrideSize), overrideTimeOff, int(zeroVel), bordSkip,
        defoAniFac, int(doAligned), partialLoadFac, thirdAlpha, fourthAlpha)
    if rc != 0:
        raise RuntimeError("warp_slice_vols: 2 or 3 deformation volumes")
    return d


def drop_sdf(X, Y, Z, T, seed, nspheres=1, rmin=0.15, rmax=0.15,
             clampCells=4.0, scale=1, dtype=np.float32):
    """Seeded drop SDF: (T, Z*s, Y*s, X*s) for scale s."""
    out = np.empty((T, Z * scale, Y * scale, X * scale), dtype=dtype)
    _fn("orc_drop_sdf", out)(_ptr(out), X, Y, Z, T, scale, seed, nspheres,
                             rmin, rmax, clampCells)
    return out


def repeat_frame4d(phi, srct, range_=0.0, bnd=0):
    """repeatFrame4d (optflow4d.cpp:1254-1266); returns the updated copy."""
    g = np.array(phi, copy=True, order="C")
    _fn("orc_repeat_frame4d", g)(_ptr(g), *_dims(g), srct, range_, bnd)
    return g


def extrap4d_ls_simple(phi, distance=4, inside=False):
    """extrap4dLsSimple (optflow4d.cpp:1361-1406); returns the updated copy."""
    g = np.array(phi, copy=True, order="C")
    _fn("orc_extrap4d_ls_simple", g)(_ptr(g), *_dims(g), distance, int(inside))
    return g


def vertex_normals(pos, tris):
    """Mesh::computeVertexNormals (mesh.cpp:460-479) of nodes pos (V, 3) and
    triangles tris (F, 3): the normals (V, 3), float32."""
    pos = np.ascontiguousarray(pos, dtype=np.float32)
    tris = np.ascontiguousarray(tris, dtype=np.int32)
    out = np.empty_like(pos)
    lib().orc_vertex_normals(pos.ctypes.data, len(pos), tris.ctypes.data, len(tris),
                              out.ctypes.data)
    return out


def create_mesh(phi):
    """LevelsetGrid::createMesh (levelset.cpp:266-365) of a 3D level set
    (Z, Y, X) float32, with the vertex normals of Mesh::computeVertexNormals
    (mesh.cpp:460-479): returns (positions (V, 3), normals (V, 3),
    triangles (F, 3))."""
    phi = np.ascontiguousarray(phi, dtype=np.float32)
    Z, Y, X = phi.shape
    nv, nt = C.c_int64(), C.c_int64()
    cap_v, cap_t = 1 << 16, 1 << 17
    while True:
        pos = np.empty((cap_v, 3), np.float32)
        nrm = np.empty((cap_v, 3), np.float32)
        tri = np.empty((cap_t, 3), np.int32)
        rc = lib().orc_create_mesh(_ptr(phi), X, Y, Z, _ptr(pos), _ptr(nrm), _ptr(tri),
                                   cap_v, cap_t, C.byref(nv), C.byref(nt))
        if rc == 0:
            return pos[:nv.value].copy(), nrm[:nv.value].copy(), tri[:nt.value].copy()
        cap_v, cap_t = max(cap_v, nv.value), max(cap_t, nt.value)


P_PLAIN = dict(wSmooth=0.1, wEnergy=1e-4, cgAccuracy=1e-5, multiStep=1,
               postVelBlur=0.0, doFinalProject=0, resetBndWidth=-1.0,
               cfl=999.0)
