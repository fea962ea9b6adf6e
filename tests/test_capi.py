"""The C-ABI library on the CPU: it loads, exports every entry point the public
header declares, and its host-side logic (parameter derivation, limits, Philox,
EPE) agrees with the oracle.  No call here touches the GPU."""
import ctypes as C
import os
import re

import numpy as np
import pytest

import briefmatch_amd as bm
import oracle

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    txt = open(os.path.join(ROOT, "include", header)).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(bm_\w+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    L = bm.lib()
    names = _declared("briefmatch_amd.h")
    assert len(names) > 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_io_library_exports_every_declared_symbol():
    from briefmatch_amd import io as bio
    L = bio.lib()
    missing = [n for n in _declared("bm_io.h") if not hasattr(L, n)]
    assert not missing, missing


def test_cpp_api_library_exports():
    so = os.path.join(ROOT, "briefmatch_amd", "libbriefmatch_cpp.so")
    if not os.path.exists(so):
        pytest.skip("C++ API library not built")
    out = os.popen("nm -D --defined-only '%s'" % so).read()
    for sym in ("BriefMatch5setupEv", "BriefMatch5matchEPfS0_", "BriefMatch10downsampleEv",
                "BriefMatch7processEPf", "OFPipeline9setParamsEiPPc", "OFPipeline12processFrameEv"):
        assert sym in out, sym


@pytest.mark.parametrize("W,H,u", [(320, 240, 1.5), (1920, 1080, 2.9), (1280, 720, 4.9), (584, 388, 0.0),
                                   (641, 479, 2.37), (17, 9, 1.0)])
def test_params_derive_matches_oracle(W, H, u):
    p = bm.derive_params(W, H, u, u)
    q = oracle.derive_params(W, H, u, u)
    assert p.as_dict() == q.as_dict()


def test_params_default_is_bmparams():
    p = bm.default_params()
    assert (p.sx, p.iterations, p.jfMax, p.N, p.fsize, p.P) == (64, 2, 256, 128, 13, 11)
    assert (p.doubleProp, p.filter, p.tInit) == (1, 1, 1)
    assert abs(p.sEPE - 0.4) < 1e-7 and abs(p.sI - 0.1) < 1e-7 and abs(p.sS - 6.0) < 1e-7


def test_philox_host_matches_oracle_and_kat():
    assert ["%08x" % v for v in bm.philox4x32_10([0] * 4, [0] * 2)] == ["6627e8d5", "e169c58d", "bc57ac4c",
                                                                        "9b00dbd8"]
    rng = np.random.default_rng(0)
    for _ in range(50):
        c = rng.integers(0, 1 << 32, 4, dtype=np.uint64).astype(np.uint32)
        k = rng.integers(0, 1 << 32, 2, dtype=np.uint64).astype(np.uint32)
        assert np.array_equal(bm.philox4x32_10(c, k), oracle.philox(c, k))


@pytest.mark.parametrize("field,value,fragment", [
    ("N", 513, "Feature length exceeds maximum value"),
    ("N", 1, "Feature length exceeds maximum value"),
    ("fsize", 49, "Median filter size exceeds maximum value"),
    ("fsize", 24, "must be odd"),
    ("sigma", 40.0, "Gaussian down-sampling filter"),
    ("sx", 40000, "32767"),
    ("sEPE", 0.0, "positive"),
])
def test_ctx_create_rejects_limits(field, value, fragment):
    """The limits of BriefMatch::setup (briefmatch.cpp:78-96) are checked before
    any device work, so this runs without a GPU."""
    p = bm.derive_params(320, 240, 1.5, 1.5)
    setattr(p, field, value)
    with pytest.raises(bm.BriefMatchError) as e:
        bm.Context(p)
    assert e.value.code == bm.BM_E_PARAM
    assert fragment in str(e.value)


def test_null_arguments_rejected():
    L = bm.lib()
    assert L.bm_params_derive(0, 10, 1.0, 1.0, 3.0, 0.025, 1.0, 0.85, None) == bm.BM_E_PARAM
    assert L.bm_sync(None) == bm.BM_E_PARAM
    assert L.bm_gray(None, None, None, None) == bm.BM_E_PARAM
    assert L.bm_flow_pair(None, None, None, 0, None, None) == bm.BM_E_PARAM
    assert L.bm_memcpy(None, None, 16, bm.H2D) == bm.BM_E_PARAM


def test_epe_matches_oracle():
    rng = np.random.default_rng(3)
    gt = rng.normal(size=(20, 30, 2)).astype(np.float32)
    gt[3, 4] = (2e9, 0)
    U = rng.normal(size=(20, 30)).astype(np.float32)
    V = rng.normal(size=(20, 30)).astype(np.float32)
    occ = (rng.random((20, 30)) < 0.2).astype(np.uint8)
    a, n = bm.epe(U, V, gt, occ)
    oa, on = oracle.epe(U, V, gt, occ)
    assert abs(a - oa) < 1e-12 and abs(n - on) < 1e-12
