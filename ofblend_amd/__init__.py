"""MI355X-native FlOF 4D deformation path (ofblend's optflow4d plugin surface)
on hand-written gfx950 HIP kernels behind a C-ABI (include/flof.h).

    from ofblend_amd import manta        # reference-style Python surface
    from ofblend_amd import _lib         # raw ctypes bindings of the C-ABI
"""
from .abi import OfParams, SolveInfo, default_params  # noqa: F401

__all__ = ["OfParams", "SolveInfo", "default_params"]
