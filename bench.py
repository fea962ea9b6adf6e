"""Benchmark of the BriefMatch flow path on one MI355X.

A step is one isolated frame pair through the whole HIP path (bm_flow_pair):
device-resident interleaved RGB uint8 1920x1080 pair -> gray -> bilinear x2.9
(5568x3132 matched pixels) fused into 128-bit BRIEF for both frames -> match
(init + 2 x {9 jump-flooding levels + random search}) -> Gaussian/NN
down-sampling -> separable median (23) -> cross-trilateral (94x94 taps) -> FP32
flow.  value = output flow pixels (1920*1080) per second / 1e6.

Prints one JSON line.  Besides the driver's fields it carries:
  roofline      the dominant kernel (largest summed device time over the timed
                steps): algorithmic bytes (or flops) per launch / its average
                launch time from HIP events on the context stream; "pipeline":
                the sum over stages of work / peak against the measured step;
  stages        per-stage event times and their own roofline fractions;
  streaming     sequence mode (one BRIEF per frame, temporal init), serial and
                pipelined (each frame's refinement tail beside the next frame);
  learned       the same pair through the learned-proposal path with half the
                propagation iterations: throughput, per-stage times and EPE
                next to the plain method's (see --no-learned);
  parity        the GPU match field against the CPU oracle on the same pair;
  cpu_baseline  the CPU oracle on the box's cores (see --no-cpu).

Options: --gpus (only 1), --steps, --warmup, --workload, --no-cpu,
--copy-calibrate BYTES (runs only the plain copy kernel, for HBM counter scaling).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = "flow Mpix/s at 1080p 128-bit (+%roofline); learned-init vs plain at equal EPE"
HBM_PEAK_GBS = 8000.0          # MI355X HBM3E
FP32_PEAK_TFLOPS = 157.3       # MI355X FP32 (published)
TRI_FLOPS_PER_TAP = 15         # 3 sub + 3 fma + exp2 + add + 2 fma (fma = 2)


def stage_work(p, NSP, s_tri):
    """Algorithmic bytes or flops per launch of each stage (compulsory traffic:
    every array read once and written once)."""
    Win = p.sx_in * p.sy_in
    M = p.sx * p.sy
    b = p.N / 8.0
    taps = Win * (2 * s_tri) ** 2
    return {
        "gray": ("hbm", Win * (3 + 1 + 4) * 1.0),              # frame 1: RGB in, gray8 + gray32 out
        "brief": ("hbm", Win * 1.0 + M * NSP * 4.0),            # fused: gray8 in, descriptors out
        "init": ("hbm", M * (2 * b + 6)),
        "level": ("hbm", M * (2 * b + 12)),                     # F1, F2 once, field in + out
        "level_rs": ("hbm", M * (2 * b + 12)),
        "downsample": ("hbm", M * 4.0 + Win * 8.0),
        "median": ("hbm", Win * 16.0),                          # two components, in + out
        "trilateral": ("valu", taps * TRI_FLOPS_PER_TAP),
    }


def roofline_of(kind, work, avg_s):
    if kind == "hbm":
        a = work / avg_s / 1e9
        return {"bound": "hbm", "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS}
    a = work / avg_s / 1e12
    return {"bound": "valu", "achieved": a, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
            "frac": a / FP32_PEAK_TFLOPS}


def load_traffic(kernel_key):
    """HBM bytes per launch from the committed rocprofv3 --pmc summary, if any."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get(kernel_key)
    except (OSError, ValueError):
        return None


def cpu_baseline(pr, p, pattern, seed, gpu_off, gpu_cost, gpu_U, gpu_V, kernel, band_rows):
    """The oracle on the host cores: every stage on the full pair except the
    trilateral filter, which runs on three row bands and is scaled to the frame."""
    import oracle
    W, H = p.sx_in, p.sy_in
    t = {}
    t0 = time.perf_counter()
    g1, g2 = oracle.gray(pr.frame1), oracle.gray(pr.frame2)
    t["gray"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    L1 = oracle.upsample(g1, p.sx, p.sy)
    L2 = oracle.upsample(g2, p.sx, p.sy)
    t["upsample"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    F1 = oracle.brief(L1, pattern, p.N)
    F2 = oracle.brief(L2, pattern, p.N)
    t["brief"] = time.perf_counter() - t0
    del L1, L2
    t0 = time.perf_counter()
    off, cost = oracle.match(F1, F2, p.iterations, p.jfMax, bool(p.doubleProp), True, None, 0, seed)
    t["match"] = time.perf_counter() - t0
    del F1, F2
    t0 = time.perf_counter()
    U0, V0 = oracle.downsample(off, W, H, kernel)
    t["downsample"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    Um = oracle.median1d(oracle.median1d(U0, 1, p.fsize), 0, p.fsize)
    Vm = oracle.median1d(oracle.median1d(V0, 1, p.fsize), 0, p.fsize)
    t["median"] = time.perf_counter() - t0
    guide = g1.astype(np.float32)
    bands = [(0, band_rows // 3), (H // 2 - band_rows // 6, H // 2 - band_rows // 6 + band_rows // 3),
             (H - band_rows // 3, H)]
    rows = sum(b - a for a, b in bands)
    t0 = time.perf_counter()
    dmax = 0.0
    for a, b in bands:
        tU, tV = oracle.trilateral(U0, V0, Um, Vm, guide, p.sEPE ** 2, p.sI ** 2, p.sS ** 2, a, b)
        dmax = max(dmax, float(np.abs(tU[a:b] - gpu_U[a:b]).max()), float(np.abs(tV[a:b] - gpu_V[a:b]).max()))
    t_tri_band = time.perf_counter() - t0
    t["trilateral"] = t_tri_band * H / rows
    total = sum(t.values())
    parity = {"match_field_bit_exact": bool(np.array_equal(off, gpu_off) and np.array_equal(cost, gpu_cost)),
              "flow_max_abs_px_on_bands": dmax, "flow_tolerance_px": 1e-3, "rows_checked": rows}
    cb = {"value": W * H / total / 1e6, "unit": "Mpix/s", "cores": oracle.num_threads(), "kind": "port",
          "sample": "same %dx%d pair: gray, up-sampling, BRIEF, match, down-sampling, median on the full "
                    "frame; trilateral on %d of %d rows (3 bands), its time scaled x%.2f" % (W, H, rows, H, H / rows),
          "seconds_measured": sum(v for k, v in t.items() if k != "trilateral") + t_tri_band,
          "stage_seconds": {k: round(v, 3) for k, v in t.items()}}
    return cb, parity


def bench_batch(a):
    """Uneven pairs (sizes uniform in [320x240, 1920x1080], random aspect,
    32/128-bit, up-sampling 1.5-2.9): per-pair launches on one stream against
    one hipGraph launch of all pairs, results compared bit for bit."""
    import briefmatch_amd as bm
    from briefmatch_amd import synth
    rng = np.random.default_rng(a.seed)
    params, frames = [], []
    t0 = time.perf_counter()
    for i in range(a.pairs):
        W, H = int(rng.integers(320, 1921)), int(rng.integers(240, 1081))
        N = int(rng.choice([32, 128]))
        u = float(rng.uniform(1.5, 2.9))
        mm = 8.0 + 32.0 * (max(W, H) - 320) / 1600.0
        pr = synth.make_pair(W, H, motion="rects", n_rects=3 + (i % 6), max_motion=mm, seed=1000 + i)
        params.append(bm.derive_params(W, H, u, u, N=N))
        frames.append(pr)
    gen_s = time.perf_counter() - t0
    batch = bm.Batch(params, seed=1, streams=8)
    r1 = [bm.DeviceBuffer(array=f.frame1) for f in frames]
    r2 = [bm.DeviceBuffer(array=f.frame2) for f in frames]
    Us = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p in params]
    Vs = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p in params]
    tuned = 0
    if not a.no_autotune:
        for i, c in enumerate(batch.ctx):
            c.autotune(r1[i], r2[i])
            tuned += 1
    pix = sum(p.sx_in * p.sy_in for p in params)
    out = {}
    fields = {}
    for name, mode in (("per_pair", bm.BATCH_PER_PAIR), ("graph", bm.BATCH_GRAPH)):
        for _ in range(max(1, a.warmup)):
            batch.flow(r1, r2, Us, Vs, mode=mode)
        batch.sync()
        batch.timer_start()
        for _ in range(a.steps):
            batch.flow(r1, r2, Us, Vs, mode=mode)
        ms = batch.timer_stop() / a.steps
        out[name] = {"ms_per_batch": round(ms, 3), "value": round(pix / (ms * 1e3), 2), "unit": "Mpix/s"}
        fields[name] = [(c.field(), U.download((p.sy_in, p.sx_in), np.float32)) for c, U, p in
                        zip(batch.ctx, Us, params)]
    exact = all(np.array_equal(x[0][0], y[0][0]) and np.array_equal(x[0][1], y[0][1]) and np.array_equal(x[1], y[1])
                for x, y in zip(fields["per_pair"], fields["graph"]))
    res = {"metric": "batched flow Mpix/s (%d uneven pairs)" % a.pairs, "value": out["graph"]["value"],
           "unit": "Mpix/s", "n_gpus": 1, "steps": a.steps, "warmup": a.warmup,
           "ms_per_step": out["graph"]["ms_per_batch"], "higher_is_better": True, "scaling": "weak",
           "vs_baseline": None, "dtype": "u32+f32", "data": "synthetic",
           "config": {"workload": "batch%d" % a.pairs, "pairs": a.pairs, "input_mpix": round(pix / 1e6, 3),
                      "matched_mpix": round(sum(p.sx * p.sy for p in params) / 1e6, 3),
                      "feature_bits": sorted(set(int(p.N) for p in params)), "autotuned_pairs": tuned,
                      "streams": 8, "generator_seconds": round(gen_s, 1)},
           "per_pair": out["per_pair"], "graph": out["graph"],
           "graph_over_per_pair": round(out["graph"]["value"] / out["per_pair"]["value"], 3),
           "bit_exact_graph_vs_per_pair": bool(exact)}
    print(json.dumps(res))


def bench_learned(a, p, tuning, rgb1, rgb2, pr, W, H, plain_all, plain_noc):
    """The pair through bm_flow_pair_learned with half the propagation
    iterations (the committed network), timed like the plain path."""
    import briefmatch_amd as bm
    from briefmatch_amd import learned
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        return None
    net = bm.Net(path=learned.DEFAULT_WEIGHTS)
    pf = p.copy()
    pf.iterations = max(1, p.iterations // 2)
    ctx = bm.Context(pf, seed=1)
    ctx.set_tuning(**tuning)
    U, V = ctx.image_buffer(), ctx.image_buffer()
    for _ in range(a.warmup):
        ctx.flow_pair_learned(net, rgb1, rgb2, U, V)
    ctx.sync()
    ctx.profile_enable(True)
    ctx.profile_read()
    ctx.timer_start()
    for _ in range(a.steps):
        ctx.flow_pair_learned(net, rgb1, rgb2, U, V)
    total_ms = ctx.timer_stop()
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    ms = total_ms / a.steps
    e_all, e_noc = bm.epe(U.download((H, W), np.float32), V.download((H, W), np.float32), pr.gt, pr.occ)
    return {"iterations": pf.iterations, "plain_iterations": p.iterations,
            "ms_per_step": round(ms, 4), "value": round(W * H / (ms * 1e3), 3), "unit": "Mpix/s",
            "epe": {"all": round(e_all, 4), "noc": round(e_noc, 4)},
            "plain_epe": {"all": round(plain_all, 4), "noc": round(plain_noc, 4)},
            "epe_ratio_vs_plain": round(e_all / plain_all, 4),
            "stages_ms_per_step": {k: round(v[0] / a.steps, 4) for k, v in prof.items() if v[1]},
            "weights": os.path.relpath(learned.DEFAULT_WEIGHTS, ROOT)}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="c2_1920x1080_u2.9_n128")
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline and parity leg")
    ap.add_argument("--band-rows", type=int, default=540,
                    help="trilateral rows of the CPU sample (about 10 s of oracle work at 1080p)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--pairs", type=int, default=64, help="pairs of the batch workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (--gpus %d requested); the flow path shards by "
              "frame pair with no communication, so N GPUs run N independent copies of this benchmark."
              % a.gpus, file=sys.stderr)
        sys.exit(2)

    import briefmatch_amd as bm
    from briefmatch_amd import synth

    if a.copy_calibrate:
        p = bm.derive_params(64, 64, 1.0, 1.0, N=32)
        ctx = bm.Context(p)
        n = a.copy_calibrate // 16 * 16
        src, dst = bm.DeviceBuffer(n), bm.DeviceBuffer(n)
        src.zero()
        ctx.profile_enable(True)
        for _ in range(5):
            ctx.device_copy(dst, src, n)
        prof = ctx.profile_read()
        ms, cnt = prof["copy"]
        print(json.dumps({"copy_bytes": n, "launches": cnt, "avg_ms": ms / cnt,
                          "GBps_read_plus_write": 2 * n / (ms / cnt / 1e3) / 1e9}))
        return

    if a.workload.startswith("batch"):
        return bench_batch(a)

    wl = synth.WORKLOADS[a.workload]
    pr = synth.workload_pair(a.workload, seed=a.seed)
    W, H = wl["W"], wl["H"]
    p = bm.derive_params(W, H, wl["up"], wl["up"], N=wl["N"])
    ctx = bm.Context(p, seed=1)
    p = ctx.params
    rgb1, rgb2 = bm.DeviceBuffer(array=pr.frame1), bm.DeviceBuffer(array=pr.frame2)
    U, V = ctx.image_buffer(), ctx.image_buffer()

    tuning = ctx.tuning().as_dict() if a.no_autotune else ctx.autotune(rgb1, rgb2).as_dict()
    for _ in range(a.warmup):
        ctx.flow_pair(rgb1, rgb2, U, V)
    ctx.sync()
    # timed region: K steps, HIP events around every kernel on the context stream
    ctx.profile_enable(True)
    ctx.profile_read()
    ctx.timer_start()
    for _ in range(a.steps):
        ctx.flow_pair(rgb1, rgb2, U, V)
    total_ms = ctx.timer_stop()
    prof = ctx.profile_read()
    ctx.profile_enable(False)
    ms_per_step = total_ms / a.steps
    value = W * H / (ms_per_step * 1e3)

    s_tri = int(np.ceil(2.5 * float(np.sqrt(np.float32(p.sS) * np.float32(p.sS)))))
    work = stage_work(p, ctx.NSP, s_tri)
    stages = {}
    for name, (ms, n) in prof.items():
        if n == 0:
            continue
        st = {"ms_per_step": ms / a.steps, "launches_per_step": n / a.steps, "avg_launch_ms": ms / n}
        if name in work:
            kind, w = work[name]
            st["roofline"] = roofline_of(kind, w, ms / n / 1e3)
        stages[name] = st
    dom = max((k for k in stages if k in work), key=lambda k: stages[k]["ms_per_step"])
    kind, w = work[dom]
    roof = roofline_of(kind, w, stages[dom]["avg_launch_ms"] / 1e3)
    roof["kernel"] = dom
    roof["work_per_launch"] = w
    roof["traffic"] = load_traffic(dom)
    # whole pair: sum over stages of algorithmic work / peak against the step time
    t_roof = 0.0
    for name, st in stages.items():
        if name in work:
            k, wk = work[name]
            t_roof += wk * st["launches_per_step"] / (HBM_PEAK_GBS * 1e9 if k == "hbm" else FP32_PEAK_TFLOPS * 1e12)
    roof["pipeline"] = {"t_roof_ms": round(t_roof * 1e3, 4), "t_measured_ms": round(ms_per_step, 4),
                        "frac": round(t_roof * 1e3 / ms_per_step, 4)}

    # streaming: alternate the two frames through the sequence entry point
    has = ctx.stream_frame(rgb1, U, V)
    for _ in range(2):
        ctx.stream_frame(rgb2, U, V)
        ctx.stream_frame(rgb1, U, V)
    ctx.sync()
    ctx.timer_start()
    ns = max(a.steps, 4)
    for i in range(ns):
        ctx.stream_frame(rgb2 if i % 2 == 0 else rgb1, U, V)
    st_ms = ctx.timer_stop() / ns
    # the same sequence pipelined: each frame's medians and trilateral filter
    # overlap the next frame's descriptors and matching (identical results)
    ctx.set_pipelining(True)
    for _ in range(2):
        ctx.stream_frame(rgb2, U, V)
        ctx.stream_frame(rgb1, U, V)
    ctx.sync()
    ctx.timer_start()
    for i in range(ns):
        ctx.stream_frame(rgb2 if i % 2 == 0 else rgb1, U, V)
    sp_ms = ctx.timer_stop() / ns
    ctx.set_pipelining(False)

    # final flow of the pair for EPE and parity
    ctx.flow_pair(rgb1, rgb2, U, V)
    ctx.sync()
    gU, gV = U.download((H, W), np.float32), V.download((H, W), np.float32)
    off, cost = ctx.field()
    e_all, e_noc = bm.epe(gU, gV, pr.gt, pr.occ)

    out = {
        "metric": METRIC, "value": round(value, 3), "unit": "Mpix/s", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms_per_step, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32+f32", "data": "synthetic",
        "config": {"workload": a.workload, "input": "%dx%d RGB uint8" % (W, H),
                   "matched": "%dx%d" % (p.sx, p.sy), "feature_bits": p.N, "iterations": p.iterations,
                   "jfMax": p.jfMax, "median": p.fsize, "trilateral_window": 2 * s_tri,
                   "mode": "isolated pair, inputs resident in HBM", "tuning": tuning},
        "roofline": roof,
        "stages": stages,
        "streaming": {"ms_per_frame": round(st_ms, 4), "value": round(W * H / (st_ms * 1e3), 3),
                      "unit": "Mpix/s",
                      "pipelined": {"ms_per_frame": round(sp_ms, 4), "value": round(W * H / (sp_ms * 1e3), 3)}},
        "epe": {"all": round(e_all, 4), "noc": round(e_noc, 4)},
    }
    if not a.no_learned:
        out["learned"] = bench_learned(a, p, tuning, rgb1, rgb2, pr, W, H, e_all, e_noc)
    if not a.no_cpu:
        cb, parity = cpu_baseline(pr, p, ctx.pattern(), 1, off, cost, gU, gV, ctx.gauss_kernel(), a.band_rows)
        out["cpu_baseline"] = cb
        out["parity"] = parity
        out["gpu_over_cpu"] = round(value / cb["value"], 1)
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))


if __name__ == "__main__":
    main()
