"""ctypes wrapper of the CPU oracle (liboracle_bm.so) -- TEST INFRASTRUCTURE ONLY.

The oracle is the checker for the HIP flow path: tests, the smoke check and the
CPU-baseline leg of the benchmark call it; the product never does.  Arrays are
numpy; the layouts are those of oracle_bm.h.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle_bm.so")
_lib = None


class Params(C.Structure):
    _fields_ = [(n, C.c_uint) for n in ("sx", "sy", "sx_in", "sy_in", "N", "iterations",
                                        "jfMax", "P", "fsize")] + \
               [(n, C.c_float) for n in ("patchArea", "sigma", "sEPE", "sI", "sS", "maxmot")] + \
               [(n, C.c_int) for n in ("doubleProp", "filter", "tInit")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


def _p(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
        _lib = C.CDLL(_LIB_PATH)
        V = C.c_void_p
        _lib.orc_derive_params.argtypes = [C.c_uint, C.c_uint] + [C.c_float] * 6 + [C.POINTER(Params)]
        _lib.orc_make_pattern.argtypes = [C.c_uint64, C.c_uint, C.c_float, V]
        _lib.orc_gauss_kernel.argtypes = [C.c_float, V, C.POINTER(C.c_uint)]
        _lib.orc_rs_sigma.argtypes = [C.c_uint]
        _lib.orc_rs_sigma.restype = C.c_float
        _lib.orc_rs_sigma_q16.argtypes = [C.c_uint]
        _lib.orc_rs_sigma_q16.restype = C.c_int64
        _lib.orc_philox4x32_10.argtypes = [V, V, V]
        _lib.orc_gray.argtypes = [V, C.c_int, C.c_int, V]
        _lib.orc_upsample.argtypes = [V, C.c_int, C.c_int, C.c_int, C.c_int, V]
        _lib.orc_upsample_lanczos.argtypes = [V, C.c_int, C.c_int, C.c_int, C.c_int, V]
        _lib.orc_lanczos_table.argtypes = [C.c_int, C.c_int, V, V]
        _lib.orc_flow_pair_up.argtypes = [C.POINTER(Params), C.c_uint64, C.c_uint32, V, V, C.c_int, V, V, V, V]
        _lib.orc_brief.argtypes = [V, C.c_int, C.c_int, V, C.c_int, V]
        _lib.orc_init.argtypes = [V, V, C.c_int, C.c_int, C.c_int, C.c_int, V, V, V]
        _lib.orc_propagate_level.argtypes = [V, V, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                             C.c_int, V, V, V, V]
        _lib.orc_propagate.argtypes = [V, V, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, V, V]
        _lib.orc_random_search.argtypes = [V, V, C.c_int, C.c_int, C.c_int, C.c_int64, C.c_uint32,
                                           C.c_uint32, C.c_uint64, V, V]
        _lib.orc_match.argtypes = [V, V, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int,
                                   V, C.c_uint32, C.c_uint64, V, V]
        _lib.orc_downsample.argtypes = [V, C.c_int, C.c_int, C.c_int, C.c_int, V, C.c_int, V, V]
        _lib.orc_median1d.argtypes = [V, V, C.c_int, C.c_int, C.c_int, C.c_int]
        _lib.orc_trilateral.argtypes = [V, V, V, V, V, C.c_int, C.c_int, C.c_float, C.c_float,
                                        C.c_float, C.c_int, C.c_int, V, V]
        _lib.orc_flow_pair.argtypes = [C.POINTER(Params), C.c_uint64, C.c_uint32, V, V, V, V, V, V]
        _lib.orc_set_num_threads.argtypes = [C.c_int]
        _lib.orc_set_num_threads.restype = None
        _lib.orc_epe.argtypes = [V, V, V, V, V, C.c_int64, C.POINTER(C.c_double), C.POINTER(C.c_double)]
    return _lib


def num_threads():
    return lib().orc_num_threads()


def set_num_threads(n):
    """OpenMP threads of the oracle's parallel regions from now on."""
    lib().orc_set_num_threads(int(n))


def philox(ctr, key):
    c = np.asarray(ctr, np.uint32)
    k = np.asarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().orc_philox4x32_10(_p(c), _p(k), _p(out))
    return out


def derive_params(W, H, ux=0.0, uy=0.0, patch_radius=3.0, sigma_ds=0.025, median_size=1.0,
                  range_spatial=0.85):
    p = Params()
    lib().orc_derive_params(W, H, ux, uy, patch_radius, sigma_ds, median_size, range_spatial, C.byref(p))
    return p


def make_pattern(seed, N, patch_area):
    out = np.zeros(4 * N, np.int16)
    lib().orc_make_pattern(seed, N, patch_area, _p(out))
    return out


def gauss_kernel(sigma):
    k = np.zeros(128, np.float32)
    P = C.c_uint()
    if lib().orc_gauss_kernel(sigma, _p(k), C.byref(P)) != 0:
        raise ValueError("Gaussian kernel too large")
    return k[:P.value].copy()


def rs_sigma_q16(sx):
    return lib().orc_rs_sigma_q16(sx)


def gray(rgb):
    rgb = np.ascontiguousarray(rgb, np.uint8)
    H, W = rgb.shape[:2]
    out = np.zeros((H, W), np.uint8)
    lib().orc_gray(_p(rgb), W, H, _p(out))
    return out


def upsample(g, sx, sy):
    g = np.ascontiguousarray(g, np.uint8)
    H, W = g.shape
    out = np.zeros((sy, sx), np.float32)
    lib().orc_upsample(_p(g), W, H, sx, sy, _p(out))
    return out


def upsample_lanczos(g, sx, sy):
    g = np.ascontiguousarray(g, np.uint8)
    H, W = g.shape
    out = np.zeros((sy, sx), np.float32)
    lib().orc_upsample_lanczos(_p(g), W, H, sx, sy, _p(out))
    return out


def lanczos_table(n_in, n_out):
    """(first source index [n_out], weights [n_out, 6]) of one axis."""
    first = np.zeros(n_out, np.int32)
    w = np.zeros((n_out, 6), np.float32)
    lib().orc_lanczos_table(n_in, n_out, _p(first), _p(w))
    return first, w


def brief(L, pattern, N):
    L = np.ascontiguousarray(L, np.float32)
    sy, sx = L.shape
    NS = (N + 31) // 32
    F = np.zeros((sy, sx, NS), np.uint32)
    lib().orc_brief(_p(L), sx, sy, _p(np.ascontiguousarray(pattern, np.int16)), N, _p(F))
    return F


def init(F1, F2, first=True, prev_off=None):
    sy, sx, NS = F1.shape
    off = np.zeros((sy, sx), np.uint32)
    cost = np.zeros((sy, sx), np.uint16)
    lib().orc_init(_p(F1), _p(F2), sx, sy, NS, int(first), _p(prev_off), _p(off), _p(cost))
    return off, cost


def propagate_level(F1, F2, p, pd, double_prop, off, cost):
    sy, sx, NS = F1.shape
    o2 = np.zeros_like(off)
    c2 = np.zeros_like(cost)
    lib().orc_propagate_level(_p(F1), _p(F2), sx, sy, NS, p, pd, int(double_prop), _p(off), _p(cost),
                              _p(o2), _p(c2))
    return o2, c2


def propagate(F1, F2, pd, jf_max, double_prop, off, cost):
    sy, sx, NS = F1.shape
    off = off.copy()
    cost = cost.copy()
    lib().orc_propagate(_p(F1), _p(F2), sx, sy, NS, pd, jf_max, int(double_prop), _p(off), _p(cost))
    return off, cost


def random_search(F1, F2, iteration, pair, seed, off, cost):
    sy, sx, NS = F1.shape
    off = off.copy()
    cost = cost.copy()
    lib().orc_random_search(_p(F1), _p(F2), sx, sy, NS, rs_sigma_q16(sx), iteration, pair, seed,
                            _p(off), _p(cost))
    return off, cost


def match(F1, F2, iterations=2, jf_max=256, double_prop=True, first=True, prev_off=None, pair=0, seed=0):
    sy, sx, NS = F1.shape
    off = np.zeros((sy, sx), np.uint32)
    cost = np.zeros((sy, sx), np.uint16)
    lib().orc_match(_p(F1), _p(F2), sx, sy, NS, iterations, jf_max, int(double_prop), int(first),
                    _p(prev_off), pair, seed, _p(off), _p(cost))
    return off, cost


def downsample(off, W, H, kernel):
    sy, sx = off.shape
    U = np.zeros((H, W), np.float32)
    V = np.zeros((H, W), np.float32)
    k = np.ascontiguousarray(kernel, np.float32)
    lib().orc_downsample(_p(np.ascontiguousarray(off)), sx, sy, W, H, _p(k), len(k), _p(U), _p(V))
    return U, V


def median1d(a, direction, fsize):
    a = np.ascontiguousarray(a, np.float32)
    H, W = a.shape
    out = np.zeros_like(a)
    lib().orc_median1d(_p(a), _p(out), W, H, direction, fsize)
    return out


def trilateral(Ui, Vi, Um, Vm, I, sigma_epe2, sigma_i2, sigma_s2, row0=0, row1=None):
    H, W = Ui.shape
    row1 = H if row1 is None else row1
    Uo = np.zeros((H, W), np.float32)
    Vo = np.zeros((H, W), np.float32)
    args = [np.ascontiguousarray(a, np.float32) for a in (Ui, Vi, Um, Vm, I)]
    lib().orc_trilateral(*[_p(a) for a in args], W, H, sigma_epe2, sigma_i2, sigma_s2, row0, row1,
                         _p(Uo), _p(Vo))
    return Uo, Vo


def flow_pair(params, rgb1, rgb2, seed=0, pair=0, lanczos=False):
    """Whole path for one isolated pair; returns (off, cost, U, V)."""
    sx, sy, W, H = params.sx, params.sy, params.sx_in, params.sy_in
    off = np.zeros((sy, sx), np.uint32)
    cost = np.zeros((sy, sx), np.uint16)
    U = np.zeros((H, W), np.float32)
    V = np.zeros((H, W), np.float32)
    r = lib().orc_flow_pair_up(C.byref(params), seed, pair, _p(np.ascontiguousarray(rgb1, np.uint8)),
                               _p(np.ascontiguousarray(rgb2, np.uint8)), int(lanczos), _p(off), _p(cost), _p(U),
                               _p(V))
    if r != 0:
        raise ValueError("oracle rejected the parameters")
    return off, cost, U, V


def epe(U, V, gt, occ=None):
    U = np.ascontiguousarray(U, np.float32)
    V = np.ascontiguousarray(V, np.float32)
    gu = np.ascontiguousarray(gt[..., 0], np.float32)
    gv = np.ascontiguousarray(gt[..., 1], np.float32)
    o = None if occ is None else np.ascontiguousarray(occ, np.uint8)
    a, n = C.c_double(), C.c_double()
    lib().orc_epe(_p(U), _p(V), _p(gu), _p(gv), _p(o), U.size, C.byref(a), C.byref(n))
    return a.value, n.value


def unpack_off(off):
    """(dx, dy) int16 planes of a packed offset field."""
    off = np.asarray(off, np.uint32)
    return (off & 0xFFFF).astype(np.uint16).view(np.int16), (off >> 16).astype(np.uint16).view(np.int16)
