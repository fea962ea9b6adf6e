"""Builds the HIP library (libbriefmatch_amd.so, gfx950), the C++ API library
(libbriefmatch_cpp.so), the `briefmatch` command-line tool and the CPU oracle.

    python -m briefmatch_amd.build            # everything
    python -m briefmatch_amd.build --lib      # only the HIP library

Every HIP translation unit is compiled with packed FP32 instructions disabled
(-packed-fp32-ops); host code is compiled without FP contraction.
"""
import argparse
import concurrent.futures
import os
import shutil
import subprocess
import sys

PKG = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB = os.path.join(PKG, "libbriefmatch_amd.so")
CPPLIB = os.path.join(PKG, "libbriefmatch_cpp.so")
CLI = os.path.join(PKG, "briefmatch")
ARCH = os.environ.get("BM_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
HIP_FLAGS = ["-O3", "-std=c++17", "--offload-arch=" + ARCH, "-fPIC", "-Wall",
             "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
             "-Xarch_host", "-ffp-contract=off", "-I" + INCLUDE]


# The trilateral filter's MFMA accumulators stay in VGPRs: with the default
# AGPR form the register allocator swaps the row groups' accumulators through
# accvgpr moves on every tap.
EXTRA_FLAGS = {"bm_kernels.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form"]}


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    out = (r.stdout or "") + (r.stderr or "")
    # the host pass does not know the gfx feature flag; that warning is harmless
    lines = [l for l in out.splitlines() if "packed-fp32-ops" not in l and "not a recognized feature" not in l]
    if r.returncode != 0:
        sys.stderr.write("\n".join(lines) + "\n")
        raise RuntimeError("build step failed: " + " ".join(cmd[:3]) + " ...")
    if verbose and lines:
        print("\n".join(lines))


def build_lib(verbose=False):
    os.makedirs(os.path.join(PKG, "_build"), exist_ok=True)
    cmds, objs = [], []
    for src in ("bm_kernels.hip", "bm_net.hip", "bm_capi.cpp", "bm_net_capi.cpp"):
        obj = os.path.join(PKG, "_build", src + ".o")
        lang = ["-x", "hip"] if src.endswith(".cpp") else []
        cmds.append([HIPCC] + HIP_FLAGS + lang + EXTRA_FLAGS.get(src, []) + ["-c", os.path.join(CSRC, src), "-o", obj])
        objs.append(obj)
    # the translation units compile independently
    with concurrent.futures.ThreadPoolExecutor(max_workers=len(cmds)) as pool:
        for f in [pool.submit(_run, c, verbose) for c in cmds]:
            f.result()
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs, verbose)
    os.replace(tmp, LIB)
    return LIB


def build_cpp(verbose=False):
    """C++ API (BriefMatch / OFPipeline) over the C ABI, and the CLI."""
    srcs = [os.path.join(CSRC, f) for f in ("briefmatch.cpp", "of_pipeline.cpp", "bm_io.cpp")]
    if not all(os.path.exists(s) for s in srcs):
        return None
    rpath = "-Wl,-rpath,$ORIGIN"
    flags = ["-O2", "-std=c++17", "-fPIC", "-Wall", "-Wextra", "-ffp-contract=off", "-I" + INCLUDE]
    tmp = CPPLIB + ".tmp"
    _run(["g++"] + flags + ["-shared", "-o", tmp] + srcs + ["-L" + PKG, "-lbriefmatch_amd", "-lz", rpath], verbose)
    os.replace(tmp, CPPLIB)
    main = os.path.join(CSRC, "main.cpp")
    if os.path.exists(main):
        tmp = CLI + ".tmp"
        _run(["g++"] + flags + ["-o", tmp, main, "-L" + PKG, "-lbriefmatch_cpp", "-lbriefmatch_amd", "-lz", rpath],
             verbose)
        os.replace(tmp, CLI)
    return CPPLIB


def build_oracle(verbose=False):
    _run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], verbose)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib", action="store_true", help="only the HIP library")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args()
    build_lib(a.verbose)
    if not a.lib:
        build_cpp(a.verbose)
        build_oracle(a.verbose)


if __name__ == "__main__":
    main()
