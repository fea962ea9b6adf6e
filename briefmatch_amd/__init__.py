"""BriefMatch optical flow on AMD MI355X (gfx950): Python host interface.

A thin ctypes layer over the C ABI in include/briefmatch_amd.h
(libbriefmatch_amd.so, built by ``python -m briefmatch_amd.build``).  It mirrors
the reference's C++ interface: ``Params`` is BMParams (briefmatch.h:65-89),
``derive_params`` is the CLI derivation of OFPipeline::setParams
(of_pipeline.cpp:178-190), ``Context`` is BriefMatch (briefmatch.h:108-155) with
one method per stage, and ``BriefMatchError`` is BriefMatchException
(briefmatch.h:91-106).  There is no CPU fallback: if the HIP library is missing
or the GPU is unavailable, calls raise.
"""
import ctypes as C
import os

import numpy as np

__all__ = ["BriefMatchError", "Params", "derive_params", "default_params", "DeviceBuffer", "Context",
           "lib", "LIB_PATH", "philox4x32_10", "epe"]

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbriefmatch_amd.so")

BM_OK, BM_E_PARAM, BM_E_ALLOC, BM_E_LAUNCH, BM_E_IO, BM_E_STATE = range(6)
# stage ids of bm_profile_read (BM_STAGE_* in briefmatch_amd.h)
STAGES = ("gray", "upsample", "brief", "init", "level", "level_rs", "random_search", "downsample", "median",
          "trilateral", "copy", "proposal", "local_search")
H2D, D2H, D2D, H2H = range(4)


class BriefMatchError(RuntimeError):
    """A non-zero status from the C ABI; ``code`` is the BM_E_* value."""

    def __init__(self, code, msg):
        super().__init__(msg)
        self.code = code


class Tuning(C.Structure):
    """bm_tuning: launch-shape choices; descriptors and match fields do not
    depend on them, the flow only through the trilateral filter's summation
    order (tri_groups, tri_split)."""
    _fields_ = [(n, C.c_int) for n in ("level_px", "strip_rows", "brief_near", "tri_band", "tri_groups",
                                       "brief_frame", "tri_split")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class BatchSchedule(C.Structure):
    """bm_batch_schedule: the result-invariant schedule of a batch."""
    _fields_ = [(n, C.c_int) for n in ("mode", "branches", "order", "persist")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}


class Params(C.Structure):
    """bm_params: the fields of BMParams (briefmatch.h:65-89)."""
    _fields_ = [(n, C.c_uint) for n in ("sx", "sy", "sx_in", "sy_in", "N", "iterations", "jfMax", "P",
                                        "fsize")] + \
               [(n, C.c_float) for n in ("patchArea", "sigma", "sEPE", "sI", "sS", "maxmot")] + \
               [(n, C.c_int) for n in ("doubleProp", "filter", "tInit")]

    def as_dict(self):
        return {n: getattr(self, n) for n, _ in self._fields_}

    def copy(self):
        p = Params()
        C.memmove(C.byref(p), C.byref(self), C.sizeof(Params))
        return p


_lib = None


def lib():
    """The loaded HIP library; raises BriefMatchError when it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise BriefMatchError(BM_E_STATE, "HIP library %s is not built; run `python -m briefmatch_amd.build`"
                              % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    V, I, U32, U64, F = C.c_void_p, C.c_int, C.c_uint32, C.c_uint64, C.c_float
    sig = {
        "bm_params_default": (None, [C.POINTER(Params)]),
        "bm_params_derive": (I, [C.c_uint, C.c_uint, F, F, F, F, F, F, C.POINTER(Params)]),
        "bm_ctx_create": (I, [C.POINTER(Params), U64, C.POINTER(V)]),
        "bm_ctx_destroy": (None, [V]),
        "bm_ctx_get_params": (I, [V, C.POINTER(Params)]),
        "bm_ctx_stream": (V, [V]),
        "bm_last_error": (C.c_char_p, [V]),
        "bm_sync": (I, [V]),
        "bm_philox4x32_10": (I, [V, V, V]),
        "bm_descriptor_stride": (C.c_uint, [V]),
        "bm_get_pattern": (I, [V, V]),
        "bm_set_pattern": (I, [V, V]),
        "bm_get_gauss_kernel": (I, [V, V, C.POINTER(C.c_uint)]),
        "bm_ctx_field": (I, [V, C.POINTER(V), C.POINTER(V)]),
        "bm_malloc": (I, [C.POINTER(V), C.c_size_t]),
        "bm_free": (I, [V]),
        "bm_memcpy": (I, [V, V, C.c_size_t, I]),
        "bm_memset": (I, [V, I, C.c_size_t]),
        "bm_gray": (I, [V, V, V, V]),
        "bm_upsample": (I, [V, V, V]),
        "bm_upsample_lanczos": (I, [V, V, V]),
        "bm_set_upsampler": (I, [V, I]),
        "bm_get_upsampler": (I, [V, C.POINTER(I)]),
        "bm_brief": (I, [V, V, V]),
        "bm_brief_gray": (I, [V, V, V]),
        "bm_init": (I, [V, V, V, I, V, V, V]),
        "bm_propagate_level": (I, [V, V, V, V, V, V, V, I, I]),
        "bm_propagate": (I, [V, V, V, V, V, I]),
        "bm_random_search": (I, [V, V, V, V, V, U32, U32]),
        "bm_match": (I, [V, V, V, I, V, U32, V, V]),
        "bm_downsample": (I, [V, V, V, V]),
        "bm_median1d": (I, [V, V, V, I]),
        "bm_trilateral": (I, [V, V, V, V, V, V, V, V]),
        "bm_refine": (I, [V, V, V, V, V, V]),
        "bm_flow_pair": (I, [V, V, V, U32, V, V]),
        "bm_flow_pair_static": (I, [C.POINTER(Params), U64, V, V, U32, V, V, V, V, V]),
        "bm_stream_frame": (I, [V, V, V, V, C.POINTER(I)]),
        "bm_stream_reset": (I, [V]),
        "bm_set_pipelining": (I, [V, I]),
        "bm_timer_start": (I, [V]),
        "bm_timer_stop": (I, [V, C.POINTER(F)]),
        "bm_epe": (I, [V, V, V, V, V, C.c_size_t, C.POINTER(C.c_double), C.POINTER(C.c_double)]),
        "bm_profile_enable": (I, [V, I]),
        "bm_profile_read": (I, [V, V, V]),
        "bm_device_copy": (I, [V, V, V, C.c_size_t]),
        "bm_get_tuning": (I, [V, C.POINTER(Tuning)]),
        "bm_set_tuning": (I, [V, C.POINTER(Tuning)]),
        "bm_autotune": (I, [V, V, V, C.c_char_p, C.POINTER(Tuning)]),
        "bm_batch_create": (I, [I, C.POINTER(Params), U64, I, C.POINTER(V)]),
        "bm_batch_destroy": (None, [V]),
        "bm_batch_ctx": (I, [V, I, C.POINTER(V)]),
        "bm_batch_flow": (I, [V, V, V, V, V, I]),
        "bm_batch_sync": (I, [V]),
        "bm_batch_timer_start": (I, [V]),
        "bm_batch_timer_stop": (I, [V, C.POINTER(F)]),
        "bm_batch_last_error": (C.c_char_p, [V]),
        "bm_batch_get_schedule": (I, [V, C.POINTER(BatchSchedule)]),
        "bm_batch_set_schedule": (I, [V, C.POINTER(BatchSchedule)]),
        "bm_batch_autotune": (I, [V, V, V, V, V, C.POINTER(BatchSchedule)]),
        "bm_net_create": (I, [C.POINTER(V)]),
        "bm_net_destroy": (None, [V]),
        "bm_net_last_error": (C.c_char_p, [V]),
        "bm_net_set_weights": (I, [V, V, C.c_size_t]),
        "bm_net_get_weights": (I, [V, V, C.c_size_t]),
        "bm_net_load": (I, [V, C.c_char_p]),
        "bm_net_save": (I, [V, C.c_char_p]),
        "bm_net_set_optimizer": (I, [V, F, F, F]),
        "bm_net_forward": (I, [V, V, C.c_uint, C.c_uint, C.c_uint, V]),
        "bm_net_gradients": (I, [V, V, V, C.c_uint, C.c_uint, C.c_uint, C.POINTER(F), V]),
        "bm_net_train_step": (I, [V, V, V, C.c_uint, V, C.c_uint, C.c_uint, C.c_uint, F, C.POINTER(F)]),
        "bm_net_sync": (I, [V]),
        "bm_proposal_dims": (I, [V, C.POINTER(C.c_uint), C.POINTER(C.c_uint)]),
        "bm_ctx_frames": (I, [V, V, V]),
        "bm_proposal_features": (I, [V, V, V, V]),
        "bm_coarse_flow": (I, [V, V, V, V]),
        "bm_proposal_far": (I, [V, V, V, V]),
        "bm_proposal_offsets": (I, [V, V, V, V, V]),
        "bm_init_proposal": (I, [V, V, V, I, V, V, V, V]),
        "bm_set_local_search": (I, [V, I, I]),
        "bm_get_local_search": (I, [V, C.POINTER(C.c_int), C.POINTER(C.c_int)]),
        "bm_local_search": (I, [V, V, V, V, V, U32, U32]),
        "bm_flow_pair_learned": (I, [V, V, V, V, U32, V, V]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype = res
        fn.argtypes = args
    _lib = L
    return L


def _ptr(a):
    if a is None:
        return None
    if isinstance(a, DeviceBuffer):
        return a.ptr
    if isinstance(a, np.ndarray):
        return a.ctypes.data_as(C.c_void_p)
    return a


def _check(status, ctx=None):
    if status != BM_OK:
        msg = lib().bm_last_error(ctx)
        raise BriefMatchError(status, (msg or b"").decode(errors="replace"))


def default_params():
    p = Params()
    lib().bm_params_default(C.byref(p))
    return p


def derive_params(W, H, ux=0.0, uy=0.0, patch_radius=3.0, sigma_ds=0.025, median_size=1.0, range_spatial=0.85,
                  **overrides):
    """OFPipeline::setParams derivation (of_pipeline.cpp:178-190); keyword
    overrides set BMParams fields afterwards (e.g. N=32, iterations=1)."""
    p = Params()
    _check(lib().bm_params_derive(W, H, ux, uy, patch_radius, sigma_ds, median_size, range_spatial, C.byref(p)))
    for k, v in overrides.items():
        setattr(p, k, v)
    return p


STATIC_SCRATCH_BYTES = 64 * 1024 * 1024  # BM_STATIC_SCRATCH_BYTES


def flow_pair_static(params, seed, rgb1, rgb2, U, V, pair=0, off=None, cost=None, stream=None):
    """bm_flow_pair_static: one isolated pair with no device allocation (its
    buffers live in the library's static device array); device buffers in and
    out, queued on `stream` (a hipStream_t as int, None = the null stream)."""
    _check(lib().bm_flow_pair_static(C.byref(params), seed, _ptr(rgb1), _ptr(rgb2), pair, _ptr(U), _ptr(V),
                                     _ptr(off) if off is not None else None,
                                     _ptr(cost) if cost is not None else None, stream))


def philox4x32_10(ctr, key):
    c = np.ascontiguousarray(ctr, np.uint32)
    k = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    _check(lib().bm_philox4x32_10(_ptr(c), _ptr(k), _ptr(out)))
    return out


def epe(U, V, gt, occ=None):
    """(EPE-all, EPE-noc) of a flow against ground truth (BriefMatch::error)."""
    U = np.ascontiguousarray(U, np.float32)
    V = np.ascontiguousarray(V, np.float32)
    gu = np.ascontiguousarray(gt[..., 0], np.float32)
    gv = np.ascontiguousarray(gt[..., 1], np.float32)
    o = None if occ is None else np.ascontiguousarray(occ, np.uint8)
    a, n = C.c_double(), C.c_double()
    _check(lib().bm_epe(_ptr(U), _ptr(V), _ptr(gu), _ptr(gv), _ptr(o), U.size, C.byref(a), C.byref(n)))
    return a.value, n.value


class DeviceBuffer:
    """Device allocation (CUDA_constructor / CUDA_destructor, util_cuda.cu:206-216)."""

    def __init__(self, nbytes=None, array=None):
        if array is not None:
            array = np.ascontiguousarray(array)
            nbytes = array.nbytes
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        _check(lib().bm_malloc(C.byref(p), max(self.nbytes, 1)))
        self.ptr = p
        if array is not None:
            self.upload(array)

    @classmethod
    def like(cls, shape, dtype):
        return cls(int(np.prod(shape)) * np.dtype(dtype).itemsize)

    def upload(self, array):
        array = np.ascontiguousarray(array)
        if array.nbytes > self.nbytes:
            raise BriefMatchError(BM_E_PARAM, "upload larger than the buffer")
        _check(lib().bm_memcpy(self.ptr, _ptr(array), array.nbytes, H2D))

    def download(self, shape, dtype):
        out = np.empty(shape, dtype)
        if out.nbytes > self.nbytes:
            raise BriefMatchError(BM_E_PARAM, "download larger than the buffer")
        _check(lib().bm_memcpy(_ptr(out), self.ptr, out.nbytes, D2H))
        return out

    def zero(self):
        _check(lib().bm_memset(self.ptr, 0, self.nbytes))

    def free(self):
        if self.ptr:
            lib().bm_free(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


def _download_raw(ptr, shape, dtype):
    out = np.empty(shape, dtype)
    _check(lib().bm_memcpy(_ptr(out), ptr, out.nbytes, D2H))
    return out


class Context:
    """One configuration (BriefMatch after setup(), briefmatch.cpp:72-214)."""

    def __init__(self, params, seed=0):
        self.params = params.copy()
        h = C.c_void_p()
        _check(lib().bm_ctx_create(C.byref(self.params), seed, C.byref(h)))
        self.h = h
        got = Params()
        _check(lib().bm_ctx_get_params(self.h, C.byref(got)), self.h)
        self.params = got
        self.seed = seed
        p = self.params
        self.W, self.H, self.sx, self.sy = p.sx_in, p.sy_in, p.sx, p.sy
        self.N = p.N
        self.NS = (p.N + 31) // 32
        self.NSP = lib().bm_descriptor_stride(self.h)

    # -- plumbing ---------------------------------------------------------
    def _c(self, status):
        _check(status, self.h)

    def close(self):
        if getattr(self, "h", None):
            lib().bm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        self._c(lib().bm_sync(self.h))

    def pattern(self):
        out = np.zeros(4 * self.N, np.int16)
        self._c(lib().bm_get_pattern(self.h, _ptr(out)))
        return out

    def set_pattern(self, pattern):
        pat = np.ascontiguousarray(pattern, np.int16)
        self._c(lib().bm_set_pattern(self.h, _ptr(pat)))

    def gauss_kernel(self):
        out = np.zeros(128, np.float32)
        P = C.c_uint()
        self._c(lib().bm_get_gauss_kernel(self.h, _ptr(out), C.byref(P)))
        return out[:P.value].copy()

    def field(self):
        """The context's last match field as numpy (off uint32, cost uint16)."""
        o, c = C.c_void_p(), C.c_void_p()
        self._c(lib().bm_ctx_field(self.h, C.byref(o), C.byref(c)))
        self.sync()
        return (_download_raw(o, (self.sy, self.sx), np.uint32), _download_raw(c, (self.sy, self.sx), np.uint16))

    # buffer helpers
    def rgb_buffer(self):
        return DeviceBuffer(self.W * self.H * 3)

    def image_buffer(self):
        return DeviceBuffer(self.W * self.H * 4)

    def desc_buffer(self):
        return DeviceBuffer(self.sx * self.sy * self.NSP * 4)

    def field_buffers(self):
        return DeviceBuffer(self.sx * self.sy * 4), DeviceBuffer(self.sx * self.sy * 2)

    def download_desc(self, buf):
        return buf.download((self.sy, self.sx, self.NSP), np.uint32)

    # -- stages -------------------------------------------------------------
    def gray(self, rgb, g8=None, g32=None):
        self._c(lib().bm_gray(self.h, _ptr(rgb), _ptr(g8), _ptr(g32)))

    def upsample(self, g8, L):
        self._c(lib().bm_upsample(self.h, _ptr(g8), _ptr(L)))

    def upsample_lanczos(self, g8, L):
        self._c(lib().bm_upsample_lanczos(self.h, _ptr(g8), _ptr(L)))

    def set_upsampler(self, mode):
        """UPSAMPLE_BILINEAR (default) or UPSAMPLE_LANCZOS3 for the whole-path calls."""
        self._c(lib().bm_set_upsampler(self.h, mode))

    def upsampler(self):
        m = C.c_int()
        self._c(lib().bm_get_upsampler(self.h, C.byref(m)))
        return m.value

    def brief(self, L, F):
        self._c(lib().bm_brief(self.h, _ptr(L), _ptr(F)))

    def brief_gray(self, g8, F):
        self._c(lib().bm_brief_gray(self.h, _ptr(g8), _ptr(F)))

    def init(self, F1, F2, off, cost, first=True, prev_off=None):
        self._c(lib().bm_init(self.h, _ptr(F1), _ptr(F2), int(first), _ptr(prev_off), _ptr(off), _ptr(cost)))

    def propagate_level(self, F1, F2, off_in, cost_in, off_out, cost_out, p, pd):
        self._c(lib().bm_propagate_level(self.h, _ptr(F1), _ptr(F2), _ptr(off_in), _ptr(cost_in), _ptr(off_out),
                                         _ptr(cost_out), p, pd))

    def propagate(self, F1, F2, off, cost, pd):
        self._c(lib().bm_propagate(self.h, _ptr(F1), _ptr(F2), _ptr(off), _ptr(cost), pd))

    def random_search(self, F1, F2, off, cost, iteration, pair=0):
        self._c(lib().bm_random_search(self.h, _ptr(F1), _ptr(F2), _ptr(off), _ptr(cost), iteration, pair))

    def match(self, F1, F2, off, cost, first=True, prev_off=None, pair=0):
        self._c(lib().bm_match(self.h, _ptr(F1), _ptr(F2), int(first), _ptr(prev_off), pair, _ptr(off),
                               _ptr(cost)))

    def downsample(self, off, U, V):
        self._c(lib().bm_downsample(self.h, _ptr(off), _ptr(U), _ptr(V)))

    def median1d(self, src, dst, direction):
        self._c(lib().bm_median1d(self.h, _ptr(src), _ptr(dst), direction))

    def trilateral(self, Ui, Vi, Um, Vm, I, Uo, Vo):
        self._c(lib().bm_trilateral(self.h, *[_ptr(a) for a in (Ui, Vi, Um, Vm, I, Uo, Vo)]))

    def refine(self, U, V, I, Uo, Vo):
        self._c(lib().bm_refine(self.h, *[_ptr(a) for a in (U, V, I, Uo, Vo)]))

    def flow_pair(self, rgb1, rgb2, U, V, pair=0):
        self._c(lib().bm_flow_pair(self.h, _ptr(rgb1), _ptr(rgb2), pair, _ptr(U), _ptr(V)))

    def stream_frame(self, rgb, U, V):
        has = C.c_int()
        self._c(lib().bm_stream_frame(self.h, _ptr(rgb), _ptr(U), _ptr(V), C.byref(has)))
        return bool(has.value)

    def stream_reset(self):
        self._c(lib().bm_stream_reset(self.h))

    def set_pipelining(self, on=True):
        """bm_set_pipelining: the refinement tail of stream_frame on a second
        stream, overlapping the next frame; results complete after sync()."""
        self._c(lib().bm_set_pipelining(self.h, int(on)))

    def profile_enable(self, on=True):
        self._c(lib().bm_profile_enable(self.h, int(on)))

    def profile_read(self):
        """{stage: (ms, launches)} since the last read (waits for the stream)."""
        ms = np.zeros(len(STAGES), np.float64)
        n = np.zeros(len(STAGES), np.uint32)
        self._c(lib().bm_profile_read(self.h, _ptr(ms), _ptr(n)))
        return {name: (float(ms[i]), int(n[i])) for i, name in enumerate(STAGES)}

    def device_copy(self, dst, src, nbytes):
        self._c(lib().bm_device_copy(self.h, _ptr(dst), _ptr(src), nbytes))

    def tuning(self):
        t = Tuning()
        self._c(lib().bm_get_tuning(self.h, C.byref(t)))
        return t

    def set_tuning(self, **kw):
        t = self.tuning()
        for k, v in kw.items():
            setattr(t, k, v)
        self._c(lib().bm_set_tuning(self.h, C.byref(t)))
        return t

    def autotune(self, rgb1, rgb2, cache_path=None):
        t = Tuning()
        self._c(lib().bm_autotune(self.h, _ptr(rgb1), _ptr(rgb2), cache_path.encode() if cache_path else None,
                                  C.byref(t)))
        return t

    def timer_start(self):
        self._c(lib().bm_timer_start(self.h))

    def timer_stop(self):
        ms = C.c_float()
        self._c(lib().bm_timer_stop(self.h, C.byref(ms)))
        return ms.value

    # -- learned proposal ---------------------------------------------------
    def proposal_dims(self):
        """(cw, ch): the coarse grid, W / 8 x H / 8."""
        cw, ch = C.c_uint(), C.c_uint()
        self._c(lib().bm_proposal_dims(self.h, C.byref(cw), C.byref(ch)))
        return cw.value, ch.value

    def frames(self):
        """Device pointers (g8[2], F[2]) of the context's own frame buffers."""
        g8 = (C.c_void_p * 2)()
        F = (C.c_void_p * 2)()
        self._c(lib().bm_ctx_frames(self.h, g8, F))
        return [C.c_void_p(g8[0]), C.c_void_p(g8[1])], [C.c_void_p(F[0]), C.c_void_p(F[1])]

    def proposal_features(self, g1, g2, feat):
        self._c(lib().bm_proposal_features(self.h, _ptr(g1), _ptr(g2), _ptr(feat)))

    def coarse_flow(self, U, V, label):
        self._c(lib().bm_coarse_flow(self.h, _ptr(U), _ptr(V), _ptr(label)))

    def proposal_far(self, g1, g2, far):
        """Far block match [H / 16][W / 16][2] int32 of two gray frames."""
        self._c(lib().bm_proposal_far(self.h, _ptr(g1), _ptr(g2), _ptr(far)))

    def proposal_offsets(self, net_out, feat, coff, far=None):
        """Per-cell proposals [ch][cw][4] int32 (network u, v and argmin u, v in Q8
        matched pixels); with a far match (proposal_far) it replaces the network's
        flow where it reaches ring 3."""
        self._c(lib().bm_proposal_offsets(self.h, _ptr(net_out), _ptr(feat), _ptr(far), _ptr(coff)))

    def init_proposal(self, F1, F2, coff, off, cost, first=True, prev_off=None):
        self._c(lib().bm_init_proposal(self.h, _ptr(F1), _ptr(F2), 1 if first else 0, _ptr(prev_off), _ptr(coff),
                                       _ptr(off), _ptr(cost)))

    def set_local_search(self, radius, samples):
        """Local search around the seeds of the learned path: `samples`
        candidates per matched pixel around the running best offset, the
        square's half side halving from `radius` (0 samples: off)."""
        self._c(lib().bm_set_local_search(self.h, int(radius), int(samples)))

    def local_search_config(self):
        r, n = C.c_int(), C.c_int()
        self._c(lib().bm_get_local_search(self.h, C.byref(r), C.byref(n)))
        return r.value, n.value

    def local_search(self, F1, F2, off, cost, iteration, pair=0):
        self._c(lib().bm_local_search(self.h, _ptr(F1), _ptr(F2), _ptr(off), _ptr(cost), iteration, pair))

    def flow_pair_learned(self, net, rgb1, rgb2, U, V, pair=0):
        self._c(lib().bm_flow_pair_learned(self.h, net.h, _ptr(rgb1), _ptr(rgb2), pair, _ptr(U), _ptr(V)))

    # -- convenience ---------------------------------------------------------
    def flow_numpy(self, frame1, frame2, pair=0):
        """Host RGB uint8 frames -> host (U, V) FP32 via bm_flow_pair."""
        r1 = DeviceBuffer(array=np.ascontiguousarray(frame1, np.uint8))
        r2 = DeviceBuffer(array=np.ascontiguousarray(frame2, np.uint8))
        U, V = self.image_buffer(), self.image_buffer()
        self.flow_pair(r1, r2, U, V, pair=pair)
        self.sync()
        return U.download((self.H, self.W), np.float32), V.download((self.H, self.W), np.float32)

    def flow_numpy_learned(self, net, frame1, frame2, pair=0):
        """Host RGB uint8 frames -> host (U, V) FP32 via bm_flow_pair_learned."""
        r1 = DeviceBuffer(array=np.ascontiguousarray(frame1, np.uint8))
        r2 = DeviceBuffer(array=np.ascontiguousarray(frame2, np.uint8))
        U, V = self.image_buffer(), self.image_buffer()
        self.flow_pair_learned(net, r1, r2, U, V, pair=pair)
        self.sync()
        return U.download((self.H, self.W), np.float32), V.download((self.H, self.W), np.float32)


class _BorrowedContext(Context):
    """A context owned by a Batch (not destroyed on its own)."""

    def __init__(self, handle):
        self.h = handle
        got = Params()
        _check(lib().bm_ctx_get_params(self.h, C.byref(got)), self.h)
        self.params = got
        p = got
        self.W, self.H, self.sx, self.sy = p.sx_in, p.sy_in, p.sx, p.sy
        self.N = p.N
        self.NS = (p.N + 31) // 32
        self.NSP = lib().bm_descriptor_stride(self.h)

    def close(self):
        self.h = None


BATCH_PER_PAIR, BATCH_GRAPH, BATCH_MULTI, BATCH_AUTO = 0, 1, 2, 3
UPSAMPLE_BILINEAR, UPSAMPLE_LANCZOS3 = 0, 1


class Batch:
    """Pairs of uneven size / up-sampling / feature length run together
    (bm_batch_*): per-pair launches, one captured hipGraph of them, or one
    launch per stage for all pairs (captured as a hipGraph)."""

    def __init__(self, params_list, seed=0, streams=8):
        arr = (Params * len(params_list))(*[p.copy() for p in params_list])
        h = C.c_void_p()
        _check(lib().bm_batch_create(len(params_list), arr, seed, streams, C.byref(h)))
        self.h = h
        self.n = len(params_list)
        self.ctx = []
        for i in range(self.n):
            c = C.c_void_p()
            _check(lib().bm_batch_ctx(self.h, i, C.byref(c)))
            self.ctx.append(_BorrowedContext(c))

    def _c(self, status):
        if status != BM_OK:
            raise BriefMatchError(status, (lib().bm_batch_last_error(self.h) or b"").decode(errors="replace"))

    def _arrays(self, rgb1, rgb2, U, V):
        n = self.n
        return [(C.c_void_p * n)(*[_ptr(b) for b in lst]) for lst in (rgb1, rgb2, U, V)]

    def flow(self, rgb1, rgb2, U, V, mode=BATCH_GRAPH):
        self._c(lib().bm_batch_flow(self.h, *self._arrays(rgb1, rgb2, U, V), mode))

    def schedule(self):
        s = BatchSchedule()
        self._c(lib().bm_batch_get_schedule(self.h, C.byref(s)))
        return s

    def set_schedule(self, **kw):
        s = self.schedule()
        for k, v in kw.items():
            setattr(s, k, v)
        self._c(lib().bm_batch_set_schedule(self.h, C.byref(s)))

    def autotune(self, rgb1, rgb2, U, V):
        s = BatchSchedule()
        self._c(lib().bm_batch_autotune(self.h, *self._arrays(rgb1, rgb2, U, V), C.byref(s)))
        return s

    def sync(self):
        self._c(lib().bm_batch_sync(self.h))

    def timer_start(self):
        self._c(lib().bm_batch_timer_start(self.h))

    def timer_stop(self):
        ms = C.c_float()
        self._c(lib().bm_batch_timer_stop(self.h, C.byref(ms)))
        return ms.value

    def close(self):
        if getattr(self, "h", None):
            for c in self.ctx:
                c.h = None
            lib().bm_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


NET_FEATURES, NET_OUTPUTS, NET_PARAMS = 128, 16, 193808


class Net:
    """The learned-proposal network (bm_net_*): BF16-on-MFMA conv net trained
    with Adam on device mini-batches."""

    def __init__(self, weights=None, path=None):
        h = C.c_void_p()
        st = lib().bm_net_create(C.byref(h))
        if st != BM_OK:
            raise BriefMatchError(st, (lib().bm_net_last_error(None) or b"").decode(errors="replace"))
        self.h = h
        if path is not None:
            self.load(path)
        elif weights is not None:
            self.set_weights(weights)

    def _c(self, status):
        if status != BM_OK:
            raise BriefMatchError(status, (lib().bm_net_last_error(self.h) or b"").decode(errors="replace"))

    def set_weights(self, w):
        w = np.ascontiguousarray(w, np.float32)
        self._c(lib().bm_net_set_weights(self.h, _ptr(w), w.size))

    def weights(self):
        w = np.empty(NET_PARAMS, np.float32)
        self._c(lib().bm_net_get_weights(self.h, _ptr(w), w.size))
        return w

    def load(self, path):
        self._c(lib().bm_net_load(self.h, os.fsencode(path)))

    def save(self, path):
        self._c(lib().bm_net_save(self.h, os.fsencode(path)))

    def set_optimizer(self, beta1=0.9, beta2=0.999, eps=1e-8):
        self._c(lib().bm_net_set_optimizer(self.h, beta1, beta2, eps))

    def forward(self, feat, n, ch, cw, out):
        self._c(lib().bm_net_forward(self.h, _ptr(feat), n, ch, cw, _ptr(out)))

    def gradients(self, feat, label, n, ch, cw):
        """(loss, gradient vector) of the mean absolute error; no update."""
        loss = C.c_float()
        g = np.empty(NET_PARAMS, np.float32)
        self._c(lib().bm_net_gradients(self.h, _ptr(feat), _ptr(label), n, ch, cw, C.byref(loss), _ptr(g)))
        return loss.value, g

    def train_step(self, feat_set, label_set, rows, idx, ch, cw, lr, want_loss=False):
        idx = np.ascontiguousarray(idx, np.int32)
        loss = C.c_float()
        self._c(lib().bm_net_train_step(self.h, _ptr(feat_set), _ptr(label_set), rows, _ptr(idx), idx.size, ch, cw,
                                        lr, C.byref(loss) if want_loss else None))
        return loss.value if want_loss else None

    def sync(self):
        self._c(lib().bm_net_sync(self.h))

    def close(self):
        if getattr(self, "h", None):
            lib().bm_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
