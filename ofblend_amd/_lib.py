"""Loader of libflof_hip.so with ctypes signatures for every C-ABI entry point
of include/flof.h.

There is no fallback: if the library is not built, or no GPU is visible, the
calls fail loudly (FlofError).
"""
import ctypes as C
import os

from .abi import OfParams, SolveInfo

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)),
                        "libflof_hip.so")


class FlofError(RuntimeError):
    """A failed C-ABI call (the reference raises RuntimeError from its
    wrappers, pwrapper/pclass.cpp:50-54)."""

    def __init__(self, code, msg):
        super().__init__("flof error %d: %s" % (code, msg))
        self.code = code


_P = C.c_void_p
_I = C.c_int
_F = C.c_float
_D = C.c_double
_SZ = C.c_size_t
_U64 = C.c_uint64
_FP = C.POINTER(C.c_float)
_DP = C.POINTER(C.c_double)
_I64P = C.POINTER(C.c_int64)
_IP = C.POINTER(C.c_int)

SIGNATURES = {
    "flof_of_params_default": ([C.POINTER(OfParams)], None),
    "flof_abi_struct_size": ([_I], _SZ),
    "flof_ctx_create": ([C.POINTER(_P), _I], _I),
    "flof_ctx_destroy": ([_P], _I),
    "flof_last_error": ([_P], C.c_char_p),
    "flof_sync": ([_P], _I),
    "flof_ctx_stream": ([_P], _P),
    "flof_ctx_trim": ([_P], _I),
    "flof_ctx_enable_timing": ([_P, _I, _I], _I),
    "flof_ctx_timing": ([_P, _I, _I64P, _DP], _I),
    "flof_ctx_timing_tag": ([_P, _I, C.c_int64, _I64P, _DP], _I),
    "flof_ctx_set_autotune": ([_P, _I], _I),
    "flof_ctx_tuned_blocks": ([_P, _I, _I, _I, _I], _I),
    "flof_ctx_tuned_shape": ([_P, _I, _I, _I, _I, _IP, _IP, _IP], _I),
    "flof_ctx_set_stencil": ([_P, _I], _I),
    "flof_stream_bandwidth": ([_P, _I, C.c_int64, _I, _DP], _I),
    "flof_check_reciprocal": ([_P, C.c_uint32, C.c_uint32, C.POINTER(_U64)], _I),
    "flof_create_mesh": ([_P, _P, _I, _I, _I, C.POINTER(_P)], _I),
    "flof_mesh_size": ([_P, _I64P, _I64P], _I),
    "flof_mesh_download": ([_P, _P, _P, _P], _I),
    "flof_mesh_save_bobj": ([_P, C.c_char_p], _I),
    "flof_mesh_destroy": ([_P], _I),
    "flof_frame_op": ([_P, _I, _P, _P, C.c_int64, _F], _I),
    "flof_simple_blur_special": ([_P, _P, _I, _I, _I, _I, _F, _I], _I),
    "flof_frame_set_bound": ([_P, _P, _I, _I, _I, _F, _I], _I),
    "flof_grid4d_create": ([_P, _I, _I, _I, _I, _I, _I, C.POINTER(_P)], _I),
    "flof_grid4d_destroy": ([_P], _I),
    "flof_grid4d_dims": ([_P, C.POINTER(C.c_int), C.POINTER(C.c_int),
                          C.POINTER(C.c_int)], _I),
    "flof_grid4d_device_ptr": ([_P], _P),
    "flof_grid4d_upload": ([_P, _P, _SZ], _I),
    "flof_grid4d_download": ([_P, _P, _SZ], _I),
    "flof_grid4d_clear": ([_P], _I),
    "flof_grid4d_copy_from": ([_P, _P], _I),
    "flof_grid4d_swap": ([_P, _P], _I),
    "flof_grid4d_set_const": ([_P, _FP], _I),
    "flof_grid4d_add_const": ([_P, _FP], _I),
    "flof_grid4d_mult_const": ([_P, _FP], _I),
    "flof_grid4d_add": ([_P, _P], _I),
    "flof_grid4d_sub": ([_P, _P], _I),
    "flof_grid4d_mult": ([_P, _P], _I),
    "flof_grid4d_add_scaled": ([_P, _P, _FP], _I),
    "flof_grid4d_clamp": ([_P, _F, _F], _I),
    "flof_grid4d_get_max": ([_P, _FP], _I),
    "flof_grid4d_get_min": ([_P, _FP], _I),
    "flof_grid4d_get_max_abs": ([_P, _FP], _I),
    "flof_grid4d_set_bound": ([_P, _FP, _I], _I),
    "flof_grid4d_set_bound_neumann": ([_P, _I], _I),
    "flof_grid4d_max_diff": ([_P, _P, _DP], _I),
    "flof_set_region4d": ([_P, _FP, _FP, _FP], _I),
    "flof_get_slice_from4d": ([_P, _I, _P, _I, _I, _I], _I),
    "flof_place_grid3d": ([_P, _I, _I, _I, _P, _I], _I),
    "flof_interpolate_grid4d": ([_P, _P, _FP, _FP, _FP], _I),
    "flof_advect4d": ([_P, _P, _F, _F], _I),
    "flof_advect_cfl4d": ([_P, _P, _F, _F, _F], _I),
    "flof_calc_ls_diff4d": ([_P, _P, _P, _F, _I, _FP], _I),
    "flof_calc_ls_diff4d_noalloc": ([_P, _P, _P, _I, _I, _I, _I, _F, _I, _P, _P], _I),
    "flof_gaussian_blur4d": ([_P, _F, _I], _I),
    "flof_optical_flow4d": ([_P, _P, _P, _P, C.POINTER(OfParams),
                             C.POINTER(SolveInfo)], _I),
    "flof_optical_flow_multiscale4d": ([_P, _P, _P, _P, C.POINTER(OfParams), _P,
                                        C.POINTER(SolveInfo), _FP], _I),
    "flof_optical_flow_multiscale4d_net": ([_P, _P, _P, C.POINTER(OfParams), _P, _I,
                                            C.POINTER(SolveInfo), _FP], _I),
    "flof_optical_flow_multiscale4d_capture": ([_P, _P, _P, C.POINTER(OfParams), _I, _P, _P,
                                                C.POINTER(SolveInfo), _FP], _I),
    "flof_optical_flow_multiscale4d_batch": ([_P, _I, C.POINTER(_P), C.POINTER(_P),
                                              C.POINTER(_P), C.POINTER(OfParams), _I,
                                              C.POINTER(SolveInfo), _FP], _I),
    "flof_optical_flow_multiscale4d_fused": ([_P, _I, C.POINTER(_P), C.POINTER(_P),
                                              C.POINTER(_P), C.POINTER(OfParams),
                                              C.POINTER(SolveInfo), _FP], _I),
    "flof_corr_vels_of4d": ([_P, _P, _P, _P, _P, _F, _F, _F, _F, _I], _I),
    "flof_load_place_grid4d": ([_P, _I, _I, _I, _P, _FP, _FP, _I, _I, _I, _F,
                                _FP, _F, _I, _F, _I, _F, _F], _I),
    "flof_advect_time_slice_opt": ([_P, _P, _I, _I, _I, _P, _F, _F, _F, _FP,
                                    _FP, _FP, _FP, _F, _I, _I, _F], _I),
    "flof_advect_time_slice": ([_P, _P, _I, _I, _I, _P, _F, _F, _F, _FP, _FP, _FP,
                                _FP, _F, _I, _F], _I),
    "flof_advect_time_slice_vols": ([C.POINTER(_P), _I, _P, _I, _I, _I, _P, _F, _F, _F,
                                     _FP, _FP, _FP, _FP, _F, _I, _I, _F, _I, _F, _F, _F],
                                    _I),
    "flof_advect_time_slices_opt": ([_P, _P, _I, _I, _I, _I, _F, _F, _P, _F,
                                     _F, _FP, _FP, _FP, _I, _F], _I),
    "flof_drop_sdf4d": ([_P, _U64, _I, _D, _D, _D], _I),
    "flof_drop_frames3d": ([_P, _P, _I, _I, _I, _I, _I, _U64, _I, _D, _D, _D],
                           _I),
    "flof_repeat_frame4d": ([_P, _F, _F, _I], _I),
    "flof_extrap4d_ls_simple": ([_P, _I, _I], _I),
    "flof_shift_forw_grid4d": ([_P, _I], _I),
    "flof_uni_file_size": ([C.c_char_p, C.POINTER(C.c_int)], _I),
    "flof_uni_write4d": ([C.c_char_p, _P, _I, _I, _I, _I, _I], _I),
    "flof_uni_read": ([C.c_char_p, _P, _I, _I, _I, _I, _I], _I),
    "flof_grid4d_save": ([_P, C.c_char_p], _I),
    "flof_grid4d_load": ([_P, C.c_char_p], _I),
    "flof_load_vel_rescaled": ([_P, C.c_char_p, C.POINTER(C.c_int), _FP, _FP, _FP], _I),
    "flof_load_place_grid4d_uni": ([C.c_char_p, _P, _FP, _FP, _I, _I, _I, _F, _FP,
                                    _F, _I, _F, _I, _F, _F], _I),
    "flof_warp_state_open": ([_P, C.c_char_p, _I, _I, _F, C.POINTER(_P)], _I),
    "flof_warp_state_add": ([_P, C.c_char_p], _I),
    "flof_warp_state_run": ([_P, _P, _I, _I, _I, _P, _F, _F, _F, _FP, _FP, _FP, _FP,
                             _F, _I, _F, _I, _F, _F], _I),
    "flof_warp_state_close": ([_P], _I),
    "flof_warmnet_num_params": ([], _I),
    "flof_warmnet_create": ([_P, _U64, C.POINTER(_P)], _I),
    "flof_warmnet_destroy": ([_P], _I),
    "flof_warmnet_get_params": ([_P, _FP, _I], _I),
    "flof_warmnet_set_params": ([_P, _FP, _I], _I),
    "flof_warmnet_get_grads": ([_P, _FP, _I], _I),
    "flof_warmnet_forward": ([_P, _P, _P, _P], _I),
    "flof_warmnet_loss_grad": ([_P, _P, _P, _P, _DP], _I),
    "flof_warmnet_adam": ([_P, _F, _F, _F, _F], _I),
    "flof_warmnet_warm_start": ([_P, _P, _P, _I, _P], _I),
    "flof_dev_alloc": ([_P, _SZ, C.POINTER(_P)], _I),
    "flof_dev_free": ([_P, _P], _I),
    "flof_dev_upload": ([_P, _P, _P, _SZ], _I),
    "flof_dev_download": ([_P, _P, _P, _SZ], _I),
}

_lib = None


def lib():
    """The loaded library (raises if it is not built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise FlofError(-4, "libflof_hip.so is not built (run "
                                "`python -m ofblend_amd.build`)")
        L = C.CDLL(LIB_PATH)
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        _lib = L
    return _lib


def floats(*v):
    return (C.c_float * len(v))(*[float(x) for x in v])


def check(rc, ctx=None):
    if rc != 0:
        msg = ""
        if ctx is not None:
            m = lib().flof_last_error(ctx)
            msg = m.decode() if m else ""
        raise FlofError(rc, msg)
