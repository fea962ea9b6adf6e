"""Time one stage of a workload under chosen launch shapes.

    python -m briefmatch_amd.stagebench trilateral --workload c2_1920x1080_u2.9_n128 \\
        --groups 0 2 4 --bands 8 --split 1 2 4 --reps 3
    python -m briefmatch_amd.stagebench brief --workload c2_1920x1080_u2.9_n128 --near 32 40 --reps 5

trilateral: synthetic device buffers of the workload's input resolution (flow
planes N(0, 3) px, an integer gray guide 0..255), per (kernel, band, split)
setting.
brief: the fused up-sampling + BRIEF kernel on the gray first frame of the
workload's seed-0 pair, per near halo.  Each setting runs --reps untimed
launches, then --reps timed ones, and prints ms per launch.  It is a profiling
harness for rocprofv3 (kernel trace and counters), not a benchmark of the path:
bench.py times the whole pair.
"""
import argparse
import time

import numpy as np

from . import Context, DeviceBuffer, derive_params
from . import synth


def trilateral(args):
    wl = synth.WORKLOADS[args.workload]
    W, H = wl["W"], wl["H"]
    p = derive_params(W, H, wl["up"], wl["up"], N=wl["N"])
    ctx = Context(p, seed=1)
    rng = np.random.default_rng(args.seed)
    U = rng.normal(0, 3, (H, W)).astype(np.float32)
    V = rng.normal(0, 3, (H, W)).astype(np.float32)
    I = rng.integers(0, 256, (H, W)).astype(np.float32)
    ins = [DeviceBuffer(array=a) for a in (U, V, U, V, I)]
    Uo, Vo = ctx.image_buffer(), ctx.image_buffer()
    bands = args.bands or [ctx.tuning().tri_band]
    for groups in args.groups:
        for band in bands:
            for split in args.split:
                ctx.set_tuning(tri_groups=groups, tri_band=band, tri_split=split)
                for _ in range(args.reps):
                    ctx.trilateral(*ins, Uo, Vo)
                ctx.sync()
                t = time.perf_counter()
                for _ in range(args.reps):
                    ctx.trilateral(*ins, Uo, Vo)
                ctx.sync()
                ms = (time.perf_counter() - t) / args.reps * 1e3
                print("trilateral %s groups %d band %2d split %d  %.3f ms" % (args.workload, groups, band, split, ms),
                      flush=True)


def brief(args):
    wl = synth.WORKLOADS[args.workload]
    W, H = wl["W"], wl["H"]
    p = derive_params(W, H, wl["up"], wl["up"], N=wl["N"])
    ctx = Context(p, seed=1)
    pr = synth.workload_pair(args.workload, seed=0)
    rgb = DeviceBuffer(array=pr.frame1)
    g8 = DeviceBuffer(W * H)
    ctx.gray(rgb, g8=g8)
    F = ctx.desc_buffer()
    for near in args.near or [ctx.tuning().brief_near]:
        ctx.set_tuning(brief_near=near)
        for _ in range(args.reps):
            ctx.brief_gray(g8, F)
        ctx.sync()
        t = time.perf_counter()
        for _ in range(args.reps):
            ctx.brief_gray(g8, F)
        ctx.sync()
        ms = (time.perf_counter() - t) / args.reps * 1e3
        print("brief %s near %2d  %.3f ms" % (args.workload, near, ms), flush=True)


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("stage", choices=["trilateral", "brief"])
    ap.add_argument("--workload", default="c2_1920x1080_u2.9_n128", choices=sorted(synth.WORKLOADS))
    ap.add_argument("--groups", type=int, nargs="+", default=[0, 2, 4], help="tri_groups values")
    ap.add_argument("--split", type=int, nargs="+", default=[1], help="tri_split values")
    ap.add_argument("--bands", type=int, nargs="*", default=None, help="tri_band values (default: the context's)")
    ap.add_argument("--near", type=int, nargs="*", default=None, help="brief_near values (default: the context's)")
    ap.add_argument("--reps", type=int, default=3)
    ap.add_argument("--seed", type=int, default=0)
    args = ap.parse_args()
    {"trilateral": trilateral, "brief": brief}[args.stage](args)


if __name__ == "__main__":
    main()
