"""Flow and image files through libbriefmatch_cpp.so (include/bm_io.h):
Middlebury .flo (flowIO.cpp:35-133), raw .bin, PNG / PGM / PPM input, PNG
output and the Middlebury colour coding (colorcode.cpp:26-74)."""
import ctypes as C
import os

import numpy as np

from . import BriefMatchError, BM_E_STATE

_HERE = os.path.dirname(os.path.abspath(__file__))
CPP_LIB_PATH = os.path.join(_HERE, "libbriefmatch_cpp.so")
CLI_PATH = os.path.join(_HERE, "briefmatch")
_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(CPP_LIB_PATH):
            raise BriefMatchError(BM_E_STATE, "C++ library %s is not built; run `python -m briefmatch_amd.build`"
                                  % CPP_LIB_PATH)
        L = C.CDLL(CPP_LIB_PATH)
        V, I, P = C.c_void_p, C.c_int, C.c_char_p
        IP = C.POINTER(C.c_int)
        for name, res, args in [("bm_io_last_error", P, []), ("bm_write_flo", I, [P, I, I, V, V]),
                                ("bm_read_flo", I, [P, IP, IP, V, V]), ("bm_write_raw", I, [P, I, I, V, V]),
                                ("bm_flow_to_rgb", I, [I, I, V, V, C.c_float, V]),
                                ("bm_read_image_rgb", I, [P, IP, IP, V]), ("bm_write_png_rgb", I, [P, I, I, V])]:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def _check(status):
    if status != 0:
        raise BriefMatchError(status, lib().bm_io_last_error().decode(errors="replace"))


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def write_flo(path, U, V):
    U = np.ascontiguousarray(U, np.float32)
    V = np.ascontiguousarray(V, np.float32)
    _check(lib().bm_write_flo(path.encode(), U.shape[1], U.shape[0], _p(U), _p(V)))


def read_flo(path):
    W, H = C.c_int(), C.c_int()
    _check(lib().bm_read_flo(path.encode(), C.byref(W), C.byref(H), None, None))
    U = np.empty((H.value, W.value), np.float32)
    V = np.empty_like(U)
    _check(lib().bm_read_flo(path.encode(), C.byref(W), C.byref(H), _p(U), _p(V)))
    return U, V


def write_raw(path, U, V):
    U = np.ascontiguousarray(U, np.float32)
    V = np.ascontiguousarray(V, np.float32)
    _check(lib().bm_write_raw(path.encode(), U.shape[1], U.shape[0], _p(U), _p(V)))


def flow_to_rgb(U, V, maxmot=0.0):
    U = np.ascontiguousarray(U, np.float32)
    V = np.ascontiguousarray(V, np.float32)
    out = np.empty(U.shape + (3,), np.uint8)
    _check(lib().bm_flow_to_rgb(U.shape[1], U.shape[0], _p(U), _p(V), maxmot, _p(out)))
    return out


def read_image(path):
    """(H, W, 3) uint8 RGB of a PNG / PGM / PPM file (gray replicated)."""
    W, H = C.c_int(), C.c_int()
    _check(lib().bm_read_image_rgb(path.encode(), C.byref(W), C.byref(H), None))
    out = np.empty((H.value, W.value, 3), np.uint8)
    _check(lib().bm_read_image_rgb(path.encode(), C.byref(W), C.byref(H), _p(out)))
    return out


def write_png(path, rgb):
    rgb = np.ascontiguousarray(rgb, np.uint8)
    _check(lib().bm_write_png_rgb(path.encode(), rgb.shape[1], rgb.shape[0], _p(rgb)))
