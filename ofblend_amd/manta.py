"""Python mirror of the mantaflow plugin surface that the FlOF path uses
(scenes/flof.py, tools/tests/test_0032_grid4dop.py, test_0042_interpol4d.py),
backed by libflof_hip through its C-ABI.

Names, keyword arguments and defaults follow the reference's PYTHON()
functions in source/plugin/optflow4d.cpp and source/grid4d.{h,cpp}; failures
raise RuntimeError (FlofError), as the reference's generated wrappers do
(source/pwrapper/pclass.cpp:50-54).  Grids live in GPU memory; to_numpy() /
from_numpy() copy them to and from host arrays of shape (T, Z, Y, X[, 4]).

    s  = Solver(gridSize=vec3(64, 64, 64), dim=3, fourthDim=64)
    i0 = s.create(Grid4Real); i1 = s.create(Grid4Real)
    vel = s.create(Grid4Vec4)
    opticalFlowMultiscale4d(vel=vel, i0=i0, i1=i1, wSmooth=0.1, wEnergy=1e-4,
                            cgAccuracy=1e-5, minGridSize=16)
"""
import ctypes as C

import numpy as np

from . import _lib
from .abi import (FLOF_F32, FLOF_FRAME_ADD_SCALED, FLOF_FRAME_COPY, FLOF_FRAME_JOIN,
                  FLOF_FRAME_MULT_CONST, FLOF_FRAME_SET_CONST, FLOF_INT, FLOF_REAL,
                  FLOF_VEC4, OfParams, SolveInfo, default_params)
from ._lib import FlofError, check, floats


class vec3(tuple):
    def __new__(cls, x=0, y=0, z=0):
        return tuple.__new__(cls, (x, y, z))

    x = property(lambda s: s[0])
    y = property(lambda s: s[1])
    z = property(lambda s: s[2])


class Vec4(tuple):
    def __new__(cls, x=0, y=None, z=None, t=None):
        if y is None:
            y = z = t = x
        return tuple.__new__(cls, (x, y, z, t))

    x = property(lambda s: s[0])
    y = property(lambda s: s[1])
    z = property(lambda s: s[2])
    t = property(lambda s: s[3])


vec4 = Vec4

_CTX = None


def context(device=0):
    """The process-wide device context (one per process, as one solver
    process of the reference)."""
    global _CTX
    if _CTX is None:
        p = C.c_void_p()
        rc = _lib.lib().flof_ctx_create(C.byref(p), device)
        if rc != 0:
            raise FlofError(rc, "could not create a device context (is a GPU "
                                "visible?)")
        _CTX = p
    return _CTX


class Solver:
    """FluidSolver (source/fluidsolver.h:31): owns the 3D grid size, the 4th
    dimension and the timestep (dt = 1 by default, fluidsolver.cpp:130)."""

    def __init__(self, name="solver", gridSize=vec3(32, 32, 32), dim=3,
                 fourthDim=-1, device=0):
        if dim != 3:
            raise FlofError(-2, "4D grids need a three-dimensional solver")
        if fourthDim is None or fourthDim < 1:
            raise FlofError(-2, "Solver does not support 4d grids")
        self.name = name
        self.gridSize = tuple(int(v) for v in gridSize)
        self.fourthDim = int(fourthDim)
        self.timestep = 1.0
        self.ctx = context(device)

    def create(self, cls, name=""):
        return cls(self)

    def getDt(self):
        return self.timestep

    def step(self):
        pass


class Grid4dBase:
    ELEM = FLOF_REAL
    NP_SHAPE_TAIL = ()
    NP_DTYPE = np.float32

    def __init__(self, parent, size=None):
        self.parent = parent
        self.ctx = parent.ctx
        X, Y, Z = parent.gridSize if size is None else size[:3]
        T = parent.fourthDim if size is None else size[3]
        self.size = (int(X), int(Y), int(Z), int(T))
        h = C.c_void_p()
        check(_lib.lib().flof_grid4d_create(self.ctx, X, Y, Z, T, self.ELEM,
                                            FLOF_F32, C.byref(h)), self.ctx)
        self.h = h

    def __del__(self):
        h = getattr(self, "h", None)
        if h is not None and h.value:
            try:
                _lib.lib().flof_grid4d_destroy(h)
            except Exception:
                pass
            self.h = None

    # sizes (grid4d.h:38-56)
    def getSizeX(self):
        return self.size[0]

    def getSizeY(self):
        return self.size[1]

    def getSizeZ(self):
        return self.size[2]

    def getSizeT(self):
        return self.size[3]

    def getSize(self):
        return Vec4(*self.size)

    def np_shape(self):
        X, Y, Z, T = self.size
        return (T, Z, Y, X) + self.NP_SHAPE_TAIL

    def nbytes(self):
        return int(np.prod(self.np_shape())) * 4

    # host transfer
    def from_numpy(self, a):
        a = np.ascontiguousarray(a, dtype=self.NP_DTYPE)
        if a.shape != self.np_shape():
            raise FlofError(-2, "shape %s != %s" % (a.shape, self.np_shape()))
        check(_lib.lib().flof_grid4d_upload(self.h, a.ctypes.data_as(C.c_void_p),
                                            a.nbytes), self.ctx)
        return self

    def to_numpy(self):
        a = np.empty(self.np_shape(), dtype=self.NP_DTYPE)
        check(_lib.lib().flof_grid4d_download(self.h, a.ctypes.data_as(C.c_void_p),
                                              a.nbytes), self.ctx)
        return a

    def _val(self, v):
        return floats(v)

    # Grid4d methods (grid4d.h:130-247)
    def clear(self):
        check(_lib.lib().flof_grid4d_clear(self.h), self.ctx)

    def copyFrom(self, a, copyType=True):
        check(_lib.lib().flof_grid4d_copy_from(self.h, a.h), self.ctx)
        return self

    def swap(self, other):
        check(_lib.lib().flof_grid4d_swap(self.h, other.h), self.ctx)

    def add(self, a):
        check(_lib.lib().flof_grid4d_add(self.h, a.h), self.ctx)

    def sub(self, a):
        check(_lib.lib().flof_grid4d_sub(self.h, a.h), self.ctx)

    def mult(self, a):
        check(_lib.lib().flof_grid4d_mult(self.h, a.h), self.ctx)

    def setConst(self, s):
        check(_lib.lib().flof_grid4d_set_const(self.h, self._val(s)), self.ctx)

    def addConst(self, s):
        check(_lib.lib().flof_grid4d_add_const(self.h, self._val(s)), self.ctx)

    def multConst(self, s):
        check(_lib.lib().flof_grid4d_mult_const(self.h, self._val(s)), self.ctx)

    def addScaled(self, a, factor):
        check(_lib.lib().flof_grid4d_add_scaled(self.h, a.h, self._val(factor)),
              self.ctx)

    def clamp(self, min, max):
        check(_lib.lib().flof_grid4d_clamp(self.h, min, max), self.ctx)

    def _red(self, fn):
        out = C.c_float()
        check(fn(self.h, C.byref(out)), self.ctx)
        return out.value

    def getMax(self):
        return self._red(_lib.lib().flof_grid4d_get_max)

    def getMin(self):
        return self._red(_lib.lib().flof_grid4d_get_min)

    def getMaxAbs(self):
        return self._red(_lib.lib().flof_grid4d_get_max_abs)

    getMaxValue = getMax
    getMinValue = getMin
    getMaxAbsValue = getMaxAbs

    def setBound(self, value, boundaryWidth=1):
        check(_lib.lib().flof_grid4d_set_bound(self.h, self._val(value),
                                               boundaryWidth), self.ctx)

    def setBoundNeumann(self, boundaryWidth=1):
        check(_lib.lib().flof_grid4d_set_bound_neumann(self.h, boundaryWidth),
              self.ctx)

    # .uni files (grid4d.cpp:122-138, fileio.cpp:834-999)
    def save(self, name):
        check(_lib.lib().flof_grid4d_save(self.h, str(name).encode()), self.ctx)

    def load(self, name):
        check(_lib.lib().flof_grid4d_load(self.h, str(name).encode()), self.ctx)


class Grid4Real(Grid4dBase):
    ELEM = FLOF_REAL


class Grid4Int(Grid4dBase):
    ELEM = FLOF_INT
    NP_DTYPE = np.int32


class Grid4Vec4(Grid4dBase):
    ELEM = FLOF_VEC4
    NP_SHAPE_TAIL = (4,)

    def _val(self, v):
        if np.isscalar(v):
            v = (v, v, v, v)
        return floats(*v)


_LAST_INFO = SolveInfo()


def getLastSolveInfo():
    """Convergence record of the last opticalFlow[Multiscale]4d call."""
    return _LAST_INFO


def _params(**kw):
    return default_params(**kw)


# ---------------------------------------------------------------------------
# plugin functions (source/plugin/optflow4d.cpp, source/grid4d.cpp)

def opticalFlowMultiscale4d(vel, i0, i1, rhsT=None, wSmooth=0., wEnergy=0.,
                            level=0, postVelBlur=0., cgAccuracy=1e-04,
                            blurType=1, cfl=999, orderTime=1, orderSpace=1,
                            resetBndWidth=-1., multiStep=1, projSizeThresh=9999,
                            minGridSize=10, doFinalProject=False, warm=None,
                            cgMaxIter=1000, precision=0):
    """optflow4d.cpp:2182-2195.  Extensions: `warm` is an initial guess of
    the total flow for the CG of every level; `precision` the storage
    precisions, a combination of FLOF_PSTORE_BF16 (the CG search direction in
    bfloat16) and FLOF_PREC_IMG_BF16 (the images in bfloat16 and the
    gradients of i1 rounded to binary16); 0 is the reference's FP32
    arithmetic."""
    if i0.parent.getDt() != 1.0:
        raise FlofError(-6, "Invalid, only dt 1 for now!")
    p = _params(wSmooth=wSmooth, wEnergy=wEnergy, level=level,
                postVelBlur=postVelBlur, cgAccuracy=cgAccuracy,
                blurType=blurType, cfl=cfl, orderTime=orderTime,
                orderSpace=orderSpace, resetBndWidth=resetBndWidth,
                multiStep=multiStep, projSizeThresh=projSizeThresh,
                minGridSize=minGridSize, doFinalProject=int(doFinalProject),
                cgMaxIter=cgMaxIter, warmStart=int(warm is not None),
                precision=precision)
    err = C.c_float()
    check(_lib.lib().flof_optical_flow_multiscale4d(
        vel.h, i0.h, i1.h, rhsT.h if rhsT is not None else None, C.byref(p),
        warm.h if warm is not None else None, C.byref(_LAST_INFO),
        C.byref(err)), vel.ctx)
    return err.value


def opticalFlowMultiscale4dBatch(vels, i0s, i1s, concurrency=4, **kw):
    """Extension: independent opticalFlowMultiscale4d solves (sizes may
    differ) run concurrently on worker streams; returns (errs, infos), each
    result identical to a single call with the same keyword arguments."""
    n = len(vels)
    if not (len(i0s) == n and len(i1s) == n):
        raise FlofError(-1, "opticalFlowMultiscale4dBatch: list lengths differ")
    if kw.pop("warm", None) is not None:
        raise FlofError(-1, "opticalFlowMultiscale4dBatch: no warm start")
    p = _params(**{k: (int(v) if k == "doFinalProject" else v) for k, v in kw.items()})
    ctx = vels[0].ctx if n else context()
    arr = lambda gs: (C.c_void_p * max(1, n))(*[g.h.value for g in gs])
    infos = (SolveInfo * max(1, n))()
    errs = (C.c_float * max(1, n))()
    check(_lib.lib().flof_optical_flow_multiscale4d_batch(
        ctx, n, arr(vels), arr(i0s), arr(i1s), C.byref(p), concurrency, infos, errs), ctx)
    return [errs[j] for j in range(n)], [infos[j] for j in range(n)]


def opticalFlowMultiscale4dFused(vels, i0s, i1s, **kw):
    """Extension: independent opticalFlowMultiscale4d solves (sizes may
    differ; one plain solve per level) whose PCG iterations share their
    launches level by level; returns (errs, infos), each result identical
    to a single call with the same keyword arguments."""
    n = len(vels)
    if not (len(i0s) == n and len(i1s) == n):
        raise FlofError(-1, "opticalFlowMultiscale4dFused: list lengths differ")
    p = _params(**{k: (int(v) if k == "doFinalProject" else v) for k, v in kw.items()})
    ctx = vels[0].ctx if n else context()
    arr = lambda gs: (C.c_void_p * max(1, n))(*[g.h.value for g in gs])
    infos = (SolveInfo * max(1, n))()
    errs = (C.c_float * max(1, n))()
    check(_lib.lib().flof_optical_flow_multiscale4d_fused(
        ctx, n, arr(vels), arr(i0s), arr(i1s), C.byref(p), infos, errs), ctx)
    return [errs[j] for j in range(n)], [infos[j] for j in range(n)]


def opticalFlow4d(vel, i0, i1, rhsT=None, wSmooth=0., wEnergy=0.,
                  postVelBlur=0., cgAccuracy=1e-04, blurType=1,
                  resetBndWidth=-1., cgMaxIter=1000, precision=0):
    """optflow4d.cpp:2110-2119 (extensions: cgMaxIter, the reference's fixed
    cap of 1000 at optflow4d.cpp:502; precision as for the multiscale
    solve)."""
    p = _params(wSmooth=wSmooth, wEnergy=wEnergy, postVelBlur=postVelBlur,
                cgAccuracy=cgAccuracy, blurType=blurType,
                resetBndWidth=resetBndWidth, cgMaxIter=cgMaxIter, precision=precision)
    check(_lib.lib().flof_optical_flow4d(
        vel.h, i0.h, i1.h, rhsT.h if rhsT is not None else None, C.byref(p),
        C.byref(_LAST_INFO)), vel.ctx)


def corrVelsOf4d(dst, vel, phiOrg, phiCurr, phiTarget, threshPhi=1e10,
                 threshNorm=1e10, postVelBlur=0., resetBndWidth=-1.,
                 maxIter=100):
    """optflow4d.cpp:2121-2130."""
    check(_lib.lib().flof_corr_vels_of4d(dst.h, vel.h, phiOrg.h, phiCurr.h,
                                         phiTarget.h, threshPhi, threshNorm,
                                         postVelBlur, resetBndWidth, maxIter),
          vel.ctx)


def advect4d(vel, grid, dtFac=1.):
    """optflow4d.cpp:1292-1306 (the grid's solver timestep times dtFac)."""
    check(_lib.lib().flof_advect4d(vel.h, grid.h, grid.parent.getDt(), dtFac),
          vel.ctx)


def calcLsDiff4d(i0, i1, out=None, correction=1., bnd=0):
    """optflow4d.cpp:2132-2137."""
    r = C.c_float()
    check(_lib.lib().flof_calc_ls_diff4d(i0.h, i1.h,
                                         out.h if out is not None else None,
                                         correction, bnd, C.byref(r)), i0.ctx)
    return r.value


def interpolateGrid4d(target, source, offset=Vec4(0.), scale=Vec4(1.),
                      size=Vec4(-1.)):
    """grid4d.cpp:539-547."""
    check(_lib.lib().flof_interpolate_grid4d(target.h, source.h,
                                             floats(*offset), floats(*scale),
                                             floats(*size)), target.ctx)


interpolateGrid4dVec = interpolateGrid4d


def gaussianBlur4d(grid, sigma=1., iter=1):
    """gaussianBlurGeneric (optflow4d.cpp:160-174) on a 4D grid."""
    check(_lib.lib().flof_gaussian_blur4d(grid.h, sigma, iter), grid.ctx)


def setRegion4d(dst, start, end, value):
    """grid4d.cpp:476-483."""
    check(_lib.lib().flof_set_region4d(dst.h, floats(*start), floats(*end),
                                       floats(value)), dst.ctx)


def setRegion4dVec4(dst, start, end, value):
    check(_lib.lib().flof_set_region4d(dst.h, floats(*start), floats(*end),
                                       floats(*value)), dst.ctx)


def grid4dMaxDiff(g1, g2):
    """grid4d.cpp:419-428 (Vec4 grids: grid4dMaxDiffVec4, :451-464)."""
    r = C.c_double()
    check(_lib.lib().flof_grid4d_max_diff(g1.h, g2.h, C.byref(r)), g1.ctx)
    return r.value


grid4dMaxDiffVec4 = grid4dMaxDiff


class DeviceArray:
    """A float32 device buffer for 3D grids / frame stacks."""

    def __init__(self, shape, ctx=None):
        self.ctx = ctx if ctx is not None else context()
        self.shape = tuple(int(s) for s in shape)
        self.nbytes = int(np.prod(self.shape)) * 4
        p = C.c_void_p()
        check(_lib.lib().flof_dev_alloc(self.ctx, self.nbytes, C.byref(p)),
              self.ctx)
        self.ptr = p

    @classmethod
    def from_numpy(cls, a, ctx=None):
        a = np.ascontiguousarray(a, dtype=np.float32)
        d = cls(a.shape, ctx)
        check(_lib.lib().flof_dev_upload(d.ctx, d.ptr,
                                         a.ctypes.data_as(C.c_void_p), a.nbytes),
              d.ctx)
        return d

    def to_numpy(self):
        a = np.empty(self.shape, dtype=np.float32)
        check(_lib.lib().flof_dev_download(self.ctx, a.ctypes.data_as(C.c_void_p),
                                           self.ptr, self.nbytes), self.ctx)
        return a

    # Grid<Real> / LevelsetGrid methods of a (Z, Y, X) frame (Z == 1: 2D)
    def _op(self, op, other, s):
        if other is not None and other.shape != self.shape:
            raise FlofError(-1, "frame shapes differ: %s vs %s" % (self.shape, other.shape))
        n = int(np.prod(self.shape))
        check(_lib.lib().flof_frame_op(self.ctx, op, self.ptr,
                                       other.ptr if other is not None else None, n,
                                       float(s)), self.ctx)
        return self

    def copyFrom(self, other):
        """source/grid.cpp:198-207"""
        return self._op(FLOF_FRAME_COPY, other, 0.)

    def setConst(self, value):
        """source/grid.cpp:230-233"""
        return self._op(FLOF_FRAME_SET_CONST, None, value)

    def multConst(self, s):
        """gridMultScalar, source/grid.h:735"""
        return self._op(FLOF_FRAME_MULT_CONST, None, s)

    def addScaled(self, other, factor):
        """gridScaledAdd, source/grid.h:740-743"""
        return self._op(FLOF_FRAME_ADD_SCALED, other, factor)

    def join(self, other):
        """LevelsetGrid::join, source/levelset.cpp:114-118"""
        return self._op(FLOF_FRAME_JOIN, other, 0.)

    def setBound(self, value=0., boundaryWidth=1):
        """Grid::setBound, source/grid.cpp:462-473"""
        z, y, x = self.shape
        check(_lib.lib().flof_frame_set_bound(self.ctx, self.ptr, x, y, z, float(value),
                                              int(boundaryWidth)), self.ctx)
        return self

    def createMesh(self, mesh, notiming=False):
        """LevelsetGrid::createMesh (source/levelset.cpp:266-365) of this 3D
        level set into `mesh` (a Mesh), on the GPU."""
        z, y, x = self.shape
        h = C.c_void_p()
        check(_lib.lib().flof_create_mesh(self.ctx, self.ptr, x, y, z, C.byref(h)), self.ctx)
        mesh._set(h, self.ctx)
        return mesh

    def __del__(self):
        if getattr(self, "ptr", None) is not None and self.ptr.value:
            try:
                _lib.lib().flof_dev_free(self.ctx, self.ptr)
            except Exception:
                pass
            self.ptr = None


class Mesh:
    """A triangle mesh made by createMesh (source/mesh.h): nodes, their file
    normals and triangles, kept on the GPU; save() writes .bobj.gz files
    (writeBobjFile, source/fileio.cpp:121-160)."""

    def __init__(self, parent=None):
        self.h = None
        self.ctx = None

    def _set(self, h, ctx):
        self._free()
        self.h, self.ctx = h, ctx

    def _free(self):
        if self.h is not None and self.h.value:
            try:
                _lib.lib().flof_mesh_destroy(self.h)
            except Exception:
                pass
        self.h = None

    def __del__(self):
        self._free()

    def _size(self):
        if self.h is None:
            return 0, 0
        nv, nt = C.c_int64(), C.c_int64()
        _lib.lib().flof_mesh_size(self.h, C.byref(nv), C.byref(nt))
        return nv.value, nt.value

    def numNodes(self):
        return self._size()[0]

    def numTris(self):
        return self._size()[1]

    def to_numpy(self):
        """(positions (V, 3), normals (V, 3), triangles (F, 3))."""
        nv, nt = self._size()
        pos = np.empty((nv, 3), np.float32)
        nrm = np.empty((nv, 3), np.float32)
        tri = np.empty((nt, 3), np.int32)
        if self.h is not None:
            check(_lib.lib().flof_mesh_download(self.h, pos.ctypes.data_as(C.c_void_p),
                                                nrm.ctypes.data_as(C.c_void_p),
                                                tri.ctypes.data_as(C.c_void_p)), self.ctx)
        return pos, nrm, tri

    def save(self, name, notiming=False):
        """Mesh::save for .bobj.gz files."""
        if not str(name).endswith(".bobj.gz"):
            raise FlofError(-1, "mesh files: only .bobj.gz is written")
        if self.h is None:
            raise FlofError(-1, "mesh: empty (no createMesh yet)")
        check(_lib.lib().flof_mesh_save_bobj(self.h, str(name).encode()), self.ctx)


def getSliceFrom4d(src, srct, dst):
    """grid4d.cpp:486-499; dst is a DeviceArray of shape (Z, Y, X)."""
    dz, dy, dx = dst.shape
    check(_lib.lib().flof_get_slice_from4d(src.h, srct, dst.ptr, dx, dy, dz),
          src.ctx)


def placeGrid3d(src, dst, dstt):
    """grid4d.cpp:516-524; src is a DeviceArray of shape (Z, Y, X)."""
    sz, sy, sx = src.shape
    check(_lib.lib().flof_place_grid3d(src.ptr, sx, sy, sz, dst.h, dstt),
          dst.ctx)


def loadPlaceGrid4dFrames(frames, phi, offset, scale, fileIdxStart=-1,
                          fileIdxEnd=-1, debugSkipLoad=999999, spread=1.,
                          overrideSize=Vec4(-1.), overrideTimeOff=0.,
                          overrideGoodRegion=0, loadTimeScale=1.,
                          rescaleSdfValues=False, sdfIsoOff=0.,
                          repeatStartFrame=0.):
    """loadPlaceGrid4d (optflow4d.cpp:1464-1595) with the frame files
    already decoded: `frames` is a DeviceArray (F, sz, sy, sx), frame f being
    file index max(fileIdxStart, 0) + f."""
    F, sz, sy, sx = frames.shape
    check(_lib.lib().flof_load_place_grid4d(
        frames.ptr, sx, sy, sz, phi.h, floats(*offset), floats(*scale),
        fileIdxStart, fileIdxEnd, debugSkipLoad, spread, floats(*overrideSize),
        overrideTimeOff, overrideGoodRegion, loadTimeScale,
        int(rescaleSdfValues), sdfIsoOff, repeatStartFrame), phi.ctx)


def loadAdvectTimeSliceFromDefo(defo, dst, phi, time, blendAlpha,
                                loadTimeScale=1., defoOffset=Vec4(0.),
                                defoScale=Vec4(1.), defoFactor=Vec4(1.),
                                overrideSize=Vec4(-1.), overrideTimeOff=0.,
                                zeroVel=False, bordSkip=1, defoAniFac=1., cached=True):
    """loadAdvectTimeSlice_OptRun (optflow4d.cpp:1951-2105; cached=False:
    loadAdvectTimeSlice, :1671-1760) for one deformation volume held in the
    Grid4Vec4 `defo`; dst is a DeviceArray (Z, Y, X) updated in place."""
    dz, dy, dx = dst.shape
    L = _lib.lib()
    if cached:
        rc = L.flof_advect_time_slice_opt(
            defo.h, dst.ptr, dx, dy, dz, phi.h, time, blendAlpha, loadTimeScale,
            floats(*defoOffset), floats(*defoScale), floats(*defoFactor),
            floats(*overrideSize), overrideTimeOff, int(zeroVel), bordSkip, defoAniFac)
    else:
        rc = L.flof_advect_time_slice(
            defo.h, dst.ptr, dx, dy, dz, phi.h, time, blendAlpha, loadTimeScale,
            floats(*defoOffset), floats(*defoScale), floats(*defoFactor),
            floats(*overrideSize), overrideTimeOff, int(zeroVel), defoAniFac)
    check(rc, phi.ctx)


def repeatFrame4d(phi, srct, range=0., bnd=0):
    """optflow4d.cpp:1254-1266."""
    check(_lib.lib().flof_repeat_frame4d(phi.h, srct, range, bnd), phi.ctx)


def extrap4dLsSimple(phi, distance=4, inside=False):
    """optflow4d.cpp:1361-1406."""
    check(_lib.lib().flof_extrap4d_ls_simple(phi.h, distance, int(inside)),
          phi.ctx)


def shiftForwGrid4d(phi, overrideGoodRegion=0):
    """optflow4d.cpp:1598-1616."""
    check(_lib.lib().flof_shift_forw_grid4d(phi.h, overrideGoodRegion), phi.ctx)


def getUniFileSize(name):
    """fileio.cpp:572-599: (x, y, z, t) of a .uni file (t = 1 for 3D)."""
    d = (C.c_int * 4)()
    rc = _lib.lib().flof_uni_file_size(str(name).encode(), d)
    if rc != 0:
        raise FlofError(rc, "can't read header of %s" % name)
    return tuple(d)


def readUni(name, shape4, elem=FLOF_REAL):
    """A .uni file into a host array ((T,Z,Y,X) or (T,Z,Y,X,4))."""
    X, Y, Z, T = shape4
    shape = (T, Z, Y, X) + ((4,) if elem == FLOF_VEC4 else ())
    a = np.empty(shape, dtype=np.int32 if elem == FLOF_INT else np.float32)
    rc = _lib.lib().flof_uni_read(str(name).encode(), a.ctypes.data_as(C.c_void_p),
                                  X, Y, Z, T, elem)
    if rc != 0:
        raise FlofError(rc, "can't read %s" % name)
    return a


def writeUni4d(name, a, elem=None):
    """A host array (T,Z,Y,X[,4]) as an M4T2 .uni file."""
    a = np.ascontiguousarray(a)
    if elem is None:
        elem = FLOF_VEC4 if a.ndim == 5 else (FLOF_INT if a.dtype == np.int32 else FLOF_REAL)
    a = np.ascontiguousarray(a, dtype=np.int32 if elem == FLOF_INT else np.float32)
    T, Z, Y, X = a.shape[:4]
    rc = _lib.lib().flof_uni_write4d(str(name).encode(), a.ctypes.data_as(C.c_void_p),
                                     X, Y, Z, T, elem)
    if rc != 0:
        raise FlofError(rc, "can't write %s" % name)


def loadPlaceGrid4d(fname, phi, offset, scale, fileIdxStart=-1, fileIdxEnd=-1,
                    debugSkipLoad=999999, spread=1., overrideSize=Vec4(-1.),
                    overrideTimeOff=0., overrideGoodRegion=0, loadTimeScale=1.,
                    rescaleSdfValues=False, sdfIsoOff=0., repeatStartFrame=0.):
    """optflow4d.cpp:1464-1595: frames from the .uni files `fname % index`."""
    check(_lib.lib().flof_load_place_grid4d_uni(
        str(fname).encode(), phi.h, floats(*offset), floats(*scale), fileIdxStart,
        fileIdxEnd, debugSkipLoad, spread, floats(*overrideSize), overrideTimeOff,
        overrideGoodRegion, loadTimeScale, int(rescaleSdfValues), sdfIsoOff,
        repeatStartFrame), phi.ctx)


_LATS = {}


def loadAdvectTimeSlice_OptInit(ID, fname1, useDefoVols=False, doAligned=False,
                                partialLoadFac=0.2):
    """optflow4d.cpp:1871-1913: the deformation is loaded to the GPU once."""
    loadAdvectTimeSlice_Finish(ID)
    h = C.c_void_p()
    ctx = context()
    check(_lib.lib().flof_warp_state_open(ctx, str(fname1).encode(), int(useDefoVols),
                                          int(doAligned), partialLoadFac, C.byref(h)), ctx)
    _LATS[ID] = h


def loadAdvectTimeSlice_OptAdd(ID, fname):
    """optflow4d.cpp:1914-1928: another deformation volume."""
    if ID not in _LATS:
        return
    ctx = context()
    check(_lib.lib().flof_warp_state_add(_LATS[ID], str(fname).encode()), ctx)


def loadAdvectTimeSlice_OptRun(ID, fname, dst, phi, time, blendAlpha, loadTimeScale,
                               defoOffset, defoScale, defoFactor, overrideSize=Vec4(-1.),
                               overrideTimeOff=0., debugVel=None, debugVelT=None,
                               zeroVel=False, thirdAlpha=0., bordSkip=1, fourthAlpha=0.,
                               defoAniFac=1.):
    """optflow4d.cpp:1951-2105 (dst: DeviceArray (Z, Y, X), updated in place)."""
    if ID not in _LATS:
        raise FlofError(-6, "Load-advect data id %d not initialized!" % ID)
    dz, dy, dx = dst.shape
    check(_lib.lib().flof_warp_state_run(
        _LATS[ID], dst.ptr, dx, dy, dz, phi.h, time, blendAlpha, loadTimeScale,
        floats(*defoOffset), floats(*defoScale), floats(*defoFactor),
        floats(*overrideSize), overrideTimeOff, int(zeroVel), thirdAlpha, bordSkip,
        fourthAlpha, defoAniFac), phi.ctx)


def loadAdvectTimeSliceFromDefoVols(vols, dst, phi, time, blendAlpha, loadTimeScale=1.,
                                    defoOffset=Vec4(0.), defoScale=Vec4(1.),
                                    defoFactor=Vec4(1.), overrideSize=Vec4(-1.),
                                    overrideTimeOff=0., zeroVel=False, thirdAlpha=0.,
                                    bordSkip=1, fourthAlpha=0., defoAniFac=1.,
                                    doAligned=False, partialLoadFac=0.2):
    """loadAdvectTimeSlice_OptRun with deformation volumes already on the GPU
    (a list of 2 or 3 Grid4Vec4, optflow4d.cpp:2013-2085)."""
    dz, dy, dx = dst.shape
    arr = (C.c_void_p * len(vols))(*[v.h.value for v in vols])
    check(_lib.lib().flof_advect_time_slice_vols(
        arr, len(vols), dst.ptr, dx, dy, dz, phi.h, time, blendAlpha, loadTimeScale,
        floats(*defoOffset), floats(*defoScale), floats(*defoFactor),
        floats(*overrideSize), overrideTimeOff, int(zeroVel), bordSkip, defoAniFac,
        int(doAligned), partialLoadFac, thirdAlpha, fourthAlpha), phi.ctx)


def loadAdvectTimeSlice(dummyID, fname, dst, phi, time, blendAlpha, loadTimeScale,
                        defoOffset, defoScale, defoFactor, overrideSize=Vec4(-1.),
                        overrideTimeOff=0., debugVel=None, debugVelT=None, zeroVel=False,
                        thirdAlpha=0., bordSkip=1, fourthAlpha=0., defoAniFac=1.):
    """optflow4d.cpp:1671-1760: the deformation is read from the M4T2 file
    `fname` on every call; dst is a DeviceArray (Z, Y, X) updated in place."""
    X, Y, Z, T = getUniFileSize(fname)
    defo = Grid4Vec4(phi.parent, (X, Y, Z, T))
    defo.load(fname)
    loadAdvectTimeSliceFromDefo(defo, dst, phi, time, blendAlpha, loadTimeScale,
                                defoOffset, defoScale, defoFactor, overrideSize,
                                overrideTimeOff, zeroVel, bordSkip, defoAniFac,
                                cached=False)


def loadAdvectTimeSlice_Finish(ID):
    """optflow4d.cpp:1930-1950."""
    h = _LATS.pop(ID, None)
    if h is not None:
        _lib.lib().flof_warp_state_close(h)


def dropSdf4d(phi, seed, nspheres=1, rmin=0.15, rmax=0.15, clampCells=4.):
    """Seeded falling-drop SDF (no reference counterpart)."""
    check(_lib.lib().flof_drop_sdf4d(phi.h, seed, nspheres, rmin, rmax,
                                     clampCells), phi.ctx)


def dropFrames3d(frames, X, Y, Z, T, scale, seed, nspheres=1, rmin=0.15,
                 rmax=0.15, clampCells=4.):
    """The same drops as T frames at `scale` x resolution into a DeviceArray
    (T, scale*Z, scale*Y, scale*X)."""
    check(_lib.lib().flof_drop_frames3d(frames.ctx, frames.ptr, X, Y, Z, T,
                                        scale, seed, nspheres, rmin, rmax,
                                        clampCells), frames.ctx)


def simpleBlurSpecial(a, iter=1, thresh=0., bord=0):
    """source/test.cpp:121-127 on a (Z, Y, X) DeviceArray, in place."""
    z, y, x = a.shape
    check(_lib.lib().flof_simple_blur_special(a.ctx, a.ptr, x, y, z, int(iter),
                                              float(thresh), int(bord)), a.ctx)


# --- blend weights of the union blends (scenes/flof.py:576-585,
# scenes/ofHelpers.py:128-245).  vec3 components are single precision as in
# the reference's Python binding; the rest is Python (double) arithmetic.

def blendW(t, l):
    """flof.py:576-585: weights (w0, wu, w1) of the two-way union blend."""
    w0 = (1.0 + l - 2.0 * t) / (1.0 + l)
    w1 = (-1.0 + l + 2.0 * t) / (1.0 + l)
    if w0 < 0.:
        w0 = 0.
    if w1 < 0.:
        w1 = 0.
    if w0 > 1.:
        w0 = 1.
    if w1 > 1.:
        w1 = 1.
    wu = 1. - w0 - w1
    return (w0, wu, w1)


def _v3(*c):
    return tuple(float(np.float32(x)) for x in c)


def _v3sub(a, b):
    return tuple(float(np.float32(np.float32(x) - np.float32(y))) for x, y in zip(a, b))


def world2uvw(p, r1, r2, r3):
    """ofHelpers.py:128-167: barycentric (u, v, w) of p (x, y) in the
    triangle r1, r2, r3; raises for a degenerate triangle or a point
    outside it (the reference exits)."""
    r13 = _v3sub(r1, r3)
    r23 = _v3sub(r2, r3)
    ma, mc = r13[0], r13[1]
    mb, md = r23[0], r23[1]
    if abs(ma * md - mb * mc) < 1e-05:
        raise FlofError(-1, "world2uvw: invalid coords")
    mdet = 1. / (ma * md - mb * mc)
    ina = mdet * md
    inb = mdet * -mb
    inc = mdet * -mc
    ind = mdet * ma
    pr3 = _v3sub(p, r3)
    triu = ina * pr3[0] + inb * pr3[1]
    triv = inc * pr3[0] + ind * pr3[1]
    triw = 1. - triu - triv
    eps = 1e-03
    if (triu < -eps or triv < -eps or triw < -eps or triu > 1. + eps or triv > 1. + eps
            or triw > 1. + eps):
        raise FlofError(-1, "world2uvw: invalid UVW")
    # the reference's clamp01(triu), clamp01(triv) discard their results
    triw = 1. - triu - triv
    return _v3(triu, triv, triw)


def weights2dim(uvw):
    """ofHelpers.py:174-245: weights (w0, w1, w2, w_uv, w_vw, w_uw, w_uvw) of
    the three-way union blend at barycentric position uvw."""
    u, v, w = _v3(*uvw)
    w0 = w1 = w2 = w_uv = w_vw = w_uw = w_uvw = 0.
    p = _v3(u, v, w)
    third = 1. / 3.
    if u >= 0.5:
        w0, w_uv, w_uw = world2uvw(p, _v3(1., 0., 0.), _v3(.5, .5, 0.), _v3(.5, 0., .5))
    elif v >= 0.5:
        w1, w_uv, w_vw = world2uvw(p, _v3(0., 1., 0.), _v3(.5, .5, 0.), _v3(0., .5, .5))
    elif w >= 0.5:
        w2, w_uw, w_vw = world2uvw(p, _v3(0., 0., 1.), _v3(.5, 0., .5), _v3(0., .5, .5))
    elif u >= v and u >= w:
        w_uv, w_uw, w_uvw = world2uvw(p, _v3(.5, .5, 0.), _v3(.5, 0., .5),
                                      _v3(third, third, third))
    elif v >= u and v >= w:
        w_uv, w_vw, w_uvw = world2uvw(p, _v3(.5, .5, 0.), _v3(0., .5, .5),
                                      _v3(third, third, third))
    elif w >= u and w >= v:
        w_uw, w_vw, w_uvw = world2uvw(p, _v3(.5, 0., .5), _v3(0., .5, .5),
                                      _v3(third, third, third))
    return (w0, w1, w2, w_uv, w_vw, w_uw, w_uvw)


# --- ofHelpers.py glue (scenes/ofHelpers.py:9-100).  The reference's vec3 /
# vec4 store single precision, so every assigned component is rounded to
# float32 here; the arithmetic between assignments is Python (double).

def _f32(x):
    return float(np.float32(x))


def vec4ToVec3(vec4d):
    """ofHelpers.py:9-14."""
    return vec3(_f32(vec4d[0]), _f32(vec4d[1]), _f32(vec4d[2]))


def vec4ToVec3i(vec4d):
    """ofHelpers.py:16-21: components truncated to int."""
    return vec3(int(vec4d[0]), int(vec4d[1]), int(vec4d[2]))


def vec4intString(vec4d):
    """ofHelpers.py:23-24."""
    return "[%d,%d,%d,%d]" % (int(vec4d[0]), int(vec4d[1]), int(vec4d[2]),
                              int(vec4d[3]))


def initRes(res, yFactor, fourdFactor, loadTimeScale):
    """ofHelpers.py:26-33: (integer 3D grid size, 4D size) of a resolution."""
    ret = Vec4(_f32(res * 1), _f32(res * yFactor), _f32(res * 1),
               _f32(res * fourdFactor * loadTimeScale))
    return (vec4ToVec3i(ret), ret)


def getParam(name, default, paramUsed, argv=None):
    """ofHelpers.py:38-45: the value after `name` on the command line (argv
    defaults to sys.argv); marks both entries as used."""
    import sys
    argv = sys.argv if argv is None else argv
    while len(paramUsed) < len(argv):
        paramUsed.append(0)
    for it in range(1, len(argv)):
        if argv[it].lower() == name and it + 1 < len(paramUsed):
            paramUsed[it] = paramUsed[it + 1] = 1
            return argv[it + 1]
    return default


def checkUnusedParam(paramUsed, argv=None):
    """ofHelpers.py:47-54: exits with status 1 if a parameter was not used."""
    import sys
    argv = sys.argv if argv is None else argv
    err = False
    for it in range(1, len(argv)):
        if paramUsed[it] == 0:
            print("Error: param %d '%s' not used!" % (it, argv[it]))
            err = True
    if err:
        sys.exit(1)


def initOffsets(gs4d, gs4dDefo, autoBorder, repeatStartDist, loadTimeScale):
    """ofHelpers.py:59-86: (loadOffset, loadScale, defoOffset, defoScale,
    defoFactor) of the phi load (a border of autoBorder on every side, the
    start shifted by repeatStartDist in t) and of the deformation (rescaled
    by the ratio of the two 4D sizes)."""
    lt = _f32(gs4d[3] * autoBorder)
    loadOffset = Vec4(_f32(gs4d[0] * autoBorder), _f32(gs4d[1] * autoBorder),
                      _f32(gs4d[2] * autoBorder), _f32(lt + repeatStartDist))
    ls = _f32(1. - 2.0 * autoBorder)
    loadScale = Vec4(ls, ls, ls, _f32(ls * loadTimeScale))
    defoOffset = Vec4(0., 0., 0., 0.)
    defoScale = Vec4(1., 1., 1., _f32(loadTimeScale))
    defoFactor = Vec4(*[_f32(_f32(gs4d[c]) / _f32(gs4dDefo[c])) for c in range(4)])
    return (loadOffset, loadScale, defoOffset, defoScale, defoFactor)


def loadVelRescaled(velgrid, prefix, defoSolver, defoOffset, defoScale, defoFactor):
    """ofHelpers.py:91-100: "<prefix>_vel.uni" (of the deformation solver's
    size) resampled into velgrid and scaled by defoFactor, on the GPU."""
    size = (C.c_int * 4)(*(tuple(defoSolver.gridSize) + (defoSolver.fourthDim,)))
    check(_lib.lib().flof_load_vel_rescaled(
        velgrid.h, ("%s_vel.uni" % prefix).encode(), size, floats(*defoOffset),
        floats(*defoScale), floats(*defoFactor)), velgrid.ctx)


def sync():
    check(_lib.lib().flof_sync(context()), context())
