"""ctypes mirrors of the POD structs and constants of include/flof.h.

The layout must match the C header field for field; the library exports
``flof_abi_struct_size`` so the tests can check the sizes agree.
"""
import ctypes as C

FLOF_OK = 0
FLOF_REAL = 1
FLOF_INT = 2
FLOF_VEC4 = 8
FLOF_F32 = 0
FLOF_BF16 = 1
FLOF_F16 = 2

FLOF_MAX_LEVELS = 16
FLOF_MAX_STEPS = 10

FLOF_PSTORE_F32 = 0
FLOF_PSTORE_BF16 = 1
FLOF_PREC_IMG_BF16 = 2  # bfloat16 images, binary16 gradients (mixed precision)

FLOF_STENCIL_AUTO = 0
FLOF_STENCIL_TMARCH = 1
FLOF_STENCIL_LDS = 2
FLOF_STENCIL_QARRAY = 3

FLOF_MIX_UPDATE = 0
FLOF_MIX_DIR = 1
FLOF_MIX_COPY = 2
FLOF_MIX_DIR_X = 3

FLOF_FRAME_COPY = 0
FLOF_FRAME_SET_CONST = 1
FLOF_FRAME_MULT_CONST = 2
FLOF_FRAME_ADD_SCALED = 3
FLOF_FRAME_JOIN = 4

FLOF_KT_CG_APPLY = 0
FLOF_KT_CG_UPDATE = 1
FLOF_KT_CG_DIR = 2
FLOF_KT_WARP = 3
FLOF_KT_ADVECT = 4


class OfParams(C.Structure):
    """flof_of_params -- keyword arguments of opticalFlowMultiscale4d
    (source/plugin/optflow4d.cpp:2182-2189)."""
    _fields_ = [
        ("wSmooth", C.c_double),
        ("wEnergy", C.c_double),
        ("level", C.c_int32),
        ("postVelBlur", C.c_double),
        ("cgAccuracy", C.c_double),
        ("blurType", C.c_int32),
        ("cfl", C.c_double),
        ("orderTime", C.c_int32),
        ("orderSpace", C.c_int32),
        ("resetBndWidth", C.c_double),
        ("multiStep", C.c_int32),
        ("projSizeThresh", C.c_int32),
        ("minGridSize", C.c_int32),
        ("doFinalProject", C.c_int32),
        ("cgMaxIter", C.c_int32),
        ("warmStart", C.c_int32),
        ("precision", C.c_int32),
    ]


def default_params(**kw):
    """The reference defaults (optflow4d.cpp:2182-2189, cap 1000 at :502),
    overridden by ``kw``."""
    p = OfParams(wSmooth=0.0, wEnergy=0.0, level=0, postVelBlur=0.0,
                 cgAccuracy=1e-4, blurType=1, cfl=999.0, orderTime=1,
                 orderSpace=1, resetBndWidth=-1.0, multiStep=1,
                 projSizeThresh=9999, minGridSize=10, doFinalProject=0,
                 cgMaxIter=1000, warmStart=0, precision=0)
    for k, v in kw.items():
        if not hasattr(p, k):
            raise TypeError("unknown parameter %r" % k)
        if isinstance(v, bool):
            v = int(v)
        setattr(p, k, v)
    return p


class SolveInfo(C.Structure):
    """flof_solve_info -- per-level convergence record (level 0 finest)."""
    _fields_ = [
        ("levels", C.c_int32),
        ("dims", (C.c_int32 * 4) * FLOF_MAX_LEVELS),
        ("steps", C.c_int32 * FLOF_MAX_LEVELS),
        ("iters", (C.c_int32 * FLOF_MAX_STEPS) * FLOF_MAX_LEVELS),
        ("residual", (C.c_double * FLOF_MAX_STEPS) * FLOF_MAX_LEVELS),
        ("status", (C.c_int32 * FLOF_MAX_STEPS) * FLOF_MAX_LEVELS),
        ("err_level", C.c_double * FLOF_MAX_LEVELS),
        ("err_pre", C.c_double),
        ("err_final", C.c_double),
        ("warm_used", C.c_int32 * FLOF_MAX_LEVELS),
        ("pad_", C.c_int32),
        ("total_iters", C.c_int64),
        ("residual_l2", (C.c_double * FLOF_MAX_STEPS) * FLOF_MAX_LEVELS),
    ]

    def summary(self):
        """Plain-dict view: per level dims, iterations, residuals, status."""
        lv = []
        for l in range(self.levels):
            ns = self.steps[l]
            lv.append({
                "dims": list(self.dims[l]),
                "iters": [self.iters[l][s] for s in range(ns)],
                "residual": [self.residual[l][s] for s in range(ns)],
                "residual_l2": [self.residual_l2[l][s] for s in range(ns)],
                "status": [self.status[l][s] for s in range(ns)],
                "err": self.err_level[l],
                "warm_used": self.warm_used[l],
            })
        return {"levels": lv, "err_pre": self.err_pre,
                "err_final": self.err_final,
                "total_iters": self.total_iters}
