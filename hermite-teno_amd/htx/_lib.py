"""ctypes binding of libhtx.so (include/htx.h).

The library is loaded from this package directory only.  If it is missing the
import fails loudly: there is no CPU fallback for the product path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libhtx.so")

HTX_OK = 0
ERRORS = {-1: "HTX_EINVAL", -2: "HTX_EDEVICE", -3: "HTX_ENOMEM", -4: "HTX_ESTATE", -5: "HTX_ENOTSUP", -6: "HTX_EIO",
          -7: "HTX_ENONFINITE"}

SCHEMES = {"teno5": 0, "hteno": 1, "teno5i": 2, "weno5js": 3, "weno5z": 4, "weno5n": 5, "teno5n": 6, "teno5a": 7,
           "weno5is": 8, "weno5is_ext": 9, "weno5ismod": 10, "weno5z_diff": 11, "teno5_diff": 12, "linear": 13,
           "hweno": 14, "htenoi": 15, "hlinear": 16}
# the schemes of the Hermite solver (12 fields: u and the x, y derivative moments)
HERMITE_SCHEMES = ("hteno", "hweno", "htenoi", "hlinear")
FLUXES = {"rusanov": 0, "rusanov_char": 0, "hllc": 1, "cu": 2}
BCS = {"zero_grad": 0, "zerograd": 0, "periodic": 1}
PRECISIONS = {"fp64": 0, "fp32_mixed": 1, "learned": 2, "fp64_fast": 3, "fp32": 4, "learned_fp32": 5}


class HtxParams(ctypes.Structure):
    _fields_ = [
        ("struct_size", ctypes.c_int),
        ("dim", ctypes.c_int),
        ("xl", ctypes.c_double),
        ("xr", ctypes.c_double),
        ("nx", ctypes.c_int),
        ("yb", ctypes.c_double),
        ("yt", ctypes.c_double),
        ("ny", ctypes.c_int),
        ("cfl", ctypes.c_double),
        ("tmax", ctypes.c_double),
        ("tinit", ctypes.c_double),
        ("gamma", ctypes.c_double),
        ("scheme", ctypes.c_int),
        ("flux", ctypes.c_int),
        ("bc", ctypes.c_int),
        ("precision", ctypes.c_int),
        ("ref_quirks", ctypes.c_int),
        ("tile_x", ctypes.c_int),
        ("tile_y", ctypes.c_int),
    ]


class HtxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_S = ctypes.c_size_t
_DP = ctypes.POINTER(ctypes.c_double)
_IP = ctypes.POINTER(ctypes.c_int)
_FP = ctypes.POINTER(ctypes.c_float)

# name -> (restype, argtypes); every symbol include/htx.h declares
SIGNATURES = {
    "htx_params_default": (_I, [ctypes.POINTER(HtxParams)]),
    "htx_create": (_I, [ctypes.POINTER(HtxParams), ctypes.POINTER(_P)]),
    "htx_destroy": (None, [_P]),
    "htx_set_state": (_I, [_P, _DP, _S]),
    "htx_get_state": (_I, [_P, _DP, _S]),
    "htx_init_vortex": (_I, [_P, _D, _D, _DP, _I]),
    "htx_init_quadrant": (_I, [_P, _D, _D, _DP]),
    "htx_init_shock_blobs": (_I, [_P, _D, _D, _D, _D, _DP, _DP, _DP, _I]),
    "htx_rhs": (_I, [_P, _DP, _S]),
    "htx_compute_dt": (_I, [_P, _DP]),
    "htx_stage": (_I, [_P, _I, _D]),
    "htx_get_stage_state": (_I, [_P, _I, _DP, _S]),
    "htx_step": (_I, [_P, _DP]),
    "htx_steps_async": (_I, [_P, _I]),
    "htx_advance": (_I, [_P, _D, _IP]),
    "htx_get_time": (_I, [_P, _DP, _IP]),
    "htx_sync": (_I, [_P]),
    "htx_save": (_I, [_P, ctypes.c_char_p]),
    "htx_load": (_I, [_P, ctypes.c_char_p]),
    "htx_timing_enable": (_I, [_P, _I]),
    "htx_timing_read": (_I, [_P, _DP, _IP]),
    "htx_trainer_create": (_I, [ctypes.c_longlong, ctypes.POINTER(_P)]),
    "htx_trainer_destroy": (None, [_P]),
    "htx_trainer_add_state": (_I, [_P, _P]),
    "htx_trainer_size": (_I, [_P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong)]),
    "htx_trainer_read": (_I, [_P, ctypes.c_longlong, ctypes.c_longlong, _FP, ctypes.POINTER(ctypes.c_ubyte)]),
    "htx_trainer_init": (_I, [_P, ctypes.c_ulonglong]),
    "htx_trainer_set_params": (_I, [_P, _FP]),
    "htx_trainer_get_params": (_I, [_P, _FP]),
    "htx_trainer_train": (_I, [_P, _I, _I, _D, _D, ctypes.c_ulonglong, _DP]),
    "htx_trainer_eval": (_I, [_P, _D, ctypes.POINTER(ctypes.c_longlong), _FP]),
    "htx_trainer_eval_inference": (_I, [_P, _D, ctypes.POINTER(ctypes.c_longlong), _FP]),
    "htx_trainer_gradient": (_I, [_P, _D, _FP, _DP]),
    "htx_load_classifier": (_I, [_P, _FP, _D]),
    "htx_classify": (_I, [_P, _I, _FP, _FP]),
    "htx_classifier_flags": (_I, [_P, ctypes.c_void_p, ctypes.c_void_p]),
    "htx_kxrcf_flags": (_I, [_P, ctypes.c_void_p]),
    "htx_classifier_stats": (_I, [_P, ctypes.POINTER(ctypes.c_longlong), ctypes.POINTER(ctypes.c_longlong),
                                  ctypes.POINTER(ctypes.c_longlong), _I]),
    "htx_batch_create": (_I, [ctypes.POINTER(HtxParams), _I, ctypes.POINTER(_P)]),
    "htx_batch_destroy": (None, [_P]),
    "htx_batch_set_state": (_I, [_P, _I, _DP, _S]),
    "htx_batch_get_state": (_I, [_P, _I, _DP, _S]),
    "htx_batch_init_quadrant": (_I, [_P, _I, _D, _D, _DP]),
    "htx_batch_load_classifier": (_I, [_P, _FP, _D]),
    "htx_batch_steps": (_I, [_P, _I]),
    "htx_batch_advance": (_I, [_P, _I, _IP]),
    "htx_batch_get_time": (_I, [_P, _I, _DP, _IP]),
    "htx_batch_sync": (_I, [_P]),
    "htx_batch_use_graph": (_I, [_P, _I]),
    "htx_batch_set_tile": (_I, [_P, _I, _I]),
    "htx_batch_tune": (_I, [_P, _I, _IP, _IP, _DP]),
    "htx_step_noalloc": (_I, [ctypes.POINTER(HtxParams), _P, _D, _P, _P]),
    "htx_init_entropy_wave": (_I, [_P, _D, _D, _D, _D, _D]),
    "htx_l1_density_error": (_I, [_P, _D, _DP]),
    "htx_lmpe_flags": (_I, [_P, ctypes.c_void_p]),
    "htx_last_error": (ctypes.c_char_p, []),
    "htx_abi_version": (_I, []),
}

_lib = None


def load():
    """Load libhtx.so once; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: build it with __graft_entry__.build()")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc):
    if rc != HTX_OK:
        msg = load().htx_last_error().decode(errors="replace")
        raise HtxError(rc, msg)
    return rc


def dptr(a):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_DP)


def fptr(a):
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_FP)


CLS_NPARAM = 5057
