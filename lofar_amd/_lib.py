"""ctypes binding of liblofar_hip.so (the C ABI declared in include/lofar_hip.h).

The library is the only compute path: importing a function that needs it raises
``LibraryMissing`` when the shared object has not been built, and every device
entry point raises ``LofarError`` on a non-zero status.  There is no CPU
fallback.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liblofar_hip.so")

OK, EARG, ECOORD, ESIZE, EHIP, ENOMEM, ENODEV, ESCRATCH = 0, -1, -2, -3, -4, -5, -6, -7
STATIC_SCRATCH_BYTES = 256 << 20  # LOFAR_STATIC_SCRATCH_BYTES
FP64, FP32C = 0, 1
NSTAGES = 10

_dp = C.POINTER(C.c_double)
_fp = C.POINTER(C.c_float)
_vp = C.c_void_p
_i, _l, _d = C.c_int, C.c_long, C.c_double


class LibraryMissing(RuntimeError):
    pass


class LofarError(RuntimeError):
    def __init__(self, where, code):
        self.code = code
        super().__init__("%s: %s (%d)" % (where, error_string(code), code))


# name -> (restype, argtypes); mirrors include/lofar_hip.h
SIGNATURES = {
    "wstack": (None, [_i, _l, _l, _l, _dp, _dp, _dp, _fp, _fp, _fp, _d, _d, _i, _i, _i, _dp, _i]),
    "phase_correction": (None, [_dp, _dp, _dp, _i, _i, _i, _i, _i, _d, _d, _d, _i]),
    "test": (_i, [_i]),
    "lofar_fft_planes": (_i, [_dp, _dp, _i, _i, _i]),
    "lofar_last_error": (_i, []),
    "lofar_error_string": (C.c_char_p, [_i]),
    "lofar_version": (_i, []),
    "lofar_profile_enable": (_i, [_i]),
    "lofar_profile_read": (_i, [_dp, C.POINTER(C.c_int)]),
    "lofar_stage_name": (C.c_char_p, [_i]),
    "lofar_grid_dev": (_i, [_i, _i, _i, _l, _l, _l, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _i, _i,
                            _i, _vp, _vp]),
    "lofar_fft_dev": (_i, [_i, _i, _vp, _vp, _i, _i, _i, _d, _vp]),
    "lofar_phase_dev": (_i, [_i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _d, _d, _d, _i, _vp]),
    "lofar_image_dev": (_i, [_i, _i, _l, _l, _l, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _i, _i, _i,
                             _d, _d, _d, _vp, _vp, _vp, _i, _vp]),
    "lofar_image_dev_static": (_i, [_i, _i, _l, _l, _l, _vp, _vp, _vp, _vp, _vp, _vp, _d, _d, _i,
                                    _i, _i, _d, _d, _d, _vp, _vp, _vp, _i, _vp]),
    "lofar_phase_adjoint_dev": (_i, [_i, _vp, _vp, _vp, _i, _i, _i, _i, _i, _d, _d, _d, _vp]),
    "lofar_degrid_cube_dev": (_i, [_i, _i, _l, _l, _l, _vp, _vp, _vp, _vp, _d, _d, _i, _i, _i, _vp,
                                   _vp, _vp, _vp]),
    "lofar_degrid_dev": (_i, [_i, _i, _l, _l, _l, _vp, _vp, _vp, _vp, _d, _d, _i, _i, _i, _d, _d,
                              _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lofar_plan_create": (_i, [_i, _i, _l, _l, _l, _vp, _vp, _vp, _vp, _d, _d, _i, _i, _i, _vp,
                               C.POINTER(_vp)]),
    "lofar_plan_image": (_i, [_vp, _vp, _vp, _vp, _d, _d, _d, _vp, _vp, _vp, _i, _vp]),
    "lofar_plan_degrid": (_i, [_vp, _d, _d, _d, _vp, _vp, _vp, _vp, _vp, _vp]),
    "lofar_plan_destroy": (_i, [_vp]),
}

_lib = None


def lib():
    """Load liblofar_hip.so once; raise LibraryMissing if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryMissing("%s is not built: run __graft_entry__.build() or "
                                 "make -C lofar_amd/csrc" % LIB_PATH)
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def error_string(code):
    try:
        return lib().lofar_error_string(code).decode()
    except LibraryMissing:
        return "error %d" % code


def check(where, code):
    if code != OK:
        raise LofarError(where, code)
    return code
