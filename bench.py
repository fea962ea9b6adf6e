"""Benchmark of the BriefMatch flow path on one MI355X.

A step is one isolated frame pair through the whole HIP path (bm_flow_pair):
device-resident interleaved RGB uint8 1920x1080 pair -> gray -> bilinear x2.9
(5568x3132 matched pixels) fused into 128-bit BRIEF for both frames -> match
(init + 2 x {9 jump-flooding levels + random search}) -> Gaussian/NN
down-sampling -> separable median (23) -> cross-trilateral (94x94 taps) -> FP32
flow.  value = output flow pixels (1920*1080) per second / 1e6.

Prints one JSON line.  Besides the driver's fields it carries:
  roofline      the dominant kernel (largest summed device time over the timed
                steps): algorithmic work per launch / its average launch time
                from HIP events on the context stream, against its binding
                ceiling; traffic = its HBM bytes per launch from the committed
                rocprofv3 FETCH_SIZE / WRITE_SIZE summary of this workload;
                "pipeline": the sum over stages of work at peak against the
                measured step;
  stages        per-stage event times (a second pass of the K steps with HIP
                events around every kernel) and roofline fractions against
                each stage's own ceilings (HBM, LDS, VALU issue, FP32);
  graph_replay  the same pair captured once as a hipGraph and replayed,
                bit-exact against bm_flow_pair;
  streaming     sequence mode (one BRIEF per frame, temporal init), serial and
                pipelined (each frame's refinement tail beside the next frame);
  learned       the same pair through the learned-proposal path (seeded
                initialisation, a local search around the seeds, one
                propagation iteration over the jump range 8): throughput,
                executed propagation levels per pair for both paths, EPE next
                to the plain method's and to the zero-network ablation's; the
                same with the full jump range and with no propagation
                (see --no-learned);
  parity        the GPU match field against the CPU oracle on the same pair;
  cpu_baseline  the CPU oracle on the box's cores: one warm-up, the median of
                three timed runs, a 1-thread figure, the CPU model (--no-cpu);
  legs          the 320x240/32-bit and 1280x720/512-bit pairs and the
                64-pair uneven batch, measured the same way without the oracle
                (--no-legs).

Options: --gpus (only 1), --steps, --warmup, --workload, --no-cpu, --no-legs,
--no-learned, --tune-cache PATH, --copy-calibrate BYTES (runs only the plain
copy kernel, for HBM counter scaling).
"""
import argparse
import json
import os
import platform
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

from briefmatch_amd import roofline as rl  # noqa: E402  (no GPU work at import)

METRIC = "flow Mpix/s at 1080p 128-bit (+%roofline); learned-init vs plain at equal EPE"
MAIN_WORKLOAD = "c2_1920x1080_u2.9_n128"
LEG_WORKLOADS = ("c1_320x240_u1.5_n32", "c3_1280x720_u4.9_n512")


def load_traffic(workload, stage):
    """HBM bytes per launch of a stage at a workload, from the committed
    rocprofv3 --pmc summary (profiles/traffic.json), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            return json.load(f).get("workloads", {}).get(workload, {}).get(stage)
    except (OSError, ValueError):
        return None


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def _oracle_stages(pr, p, seed, pattern, kern, bands):
    """One run of the oracle's stages on a pair; the trilateral filter on the
    given row bands only.  Returns (seconds per stage, field, cost, U0, V0,
    Um, Vm, guide, band outputs)."""
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
    U0, V0 = oracle.downsample(off, W, H, kern)
    t["downsample"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    Um = oracle.median1d(oracle.median1d(U0, 1, p.fsize), 0, p.fsize)
    Vm = oracle.median1d(oracle.median1d(V0, 1, p.fsize), 0, p.fsize)
    t["median"] = time.perf_counter() - t0
    guide = g1.astype(np.float32)
    outs = []
    t0 = time.perf_counter()
    for a, b in bands:
        outs.append(oracle.trilateral(U0, V0, Um, Vm, guide, p.sEPE ** 2, p.sI ** 2, p.sS ** 2, a, b))
    t["trilateral_bands"] = time.perf_counter() - t0
    return t, off, cost, outs


def cpu_baseline(pr, p, seed, gpu_off, gpu_cost, gpu_U, gpu_V, tri_rows):
    """The oracle on the host cores (the CPU restatement of the reference's
    kernels): every stage on the full pair except the trilateral filter, which
    runs on three row bands of tri_rows rows in all and is scaled to the frame;
    one warm-up run, then the median of three timed runs.  A 1-thread figure
    comes from the whole oracle path on a 1/8 x 1/8 crop of the same pair with
    the full frame's per-pixel parameters, scaled by the pixel count."""
    import oracle
    W, H = p.sx_in, p.sy_in
    pattern = oracle.make_pattern(seed, p.N, p.patchArea)
    kern = oracle.gauss_kernel(p.sigma)
    threads = oracle.num_threads()
    # the oracle splits a band's rows over its threads: bands of a multiple of
    # the thread count keep every thread busy, so the scaled time is the frame's
    third = max(threads, -(-max(1, tri_rows // 3) // threads) * threads)
    third = min(third, max(1, H // 3))
    bands = [(0, third), (H // 2 - third // 2, H // 2 - third // 2 + third), (H - third, H)]
    rows = sum(b - a for a, b in bands)
    runs = []
    parity = None
    measured = 0.0                                  # seconds actually spent, unscaled
    for k in range(4):                              # run 0 is the warm-up
        t, off, cost, outs = _oracle_stages(pr, p, seed, pattern, kern, bands)
        measured += sum(t.values())
        t["trilateral"] = t.pop("trilateral_bands") * H / rows
        if k == 0:
            dmax = 0.0
            for (a, b), (tU, tV) in zip(bands, outs):
                dmax = max(dmax, float(np.abs(tU[a:b] - gpu_U[a:b]).max()), float(np.abs(tV[a:b] - gpu_V[a:b]).max()))
            parity = {"match_field_bit_exact": bool(np.array_equal(off, gpu_off) and np.array_equal(cost, gpu_cost)),
                      "flow_max_abs_px_on_bands": dmax, "flow_tolerance_px": 1e-3, "rows_checked": rows}
        else:
            runs.append(t)
        del off, cost, outs
    totals = [sum(t.values()) for t in runs]
    med = int(np.argsort(totals)[1])
    stage_med = {k: round(float(np.median([t[k] for t in runs])), 4) for k in runs[0]}
    # 1 thread: the whole oracle path on a crop, per-pixel parameters of the full frame
    cw, ch = max(16, W // 8), max(16, H // 8)
    x0, y0 = (W - cw) // 2, (H - ch) // 2
    cp = oracle.derive_params(cw, ch, p.sx / W, p.sy / H)
    cp.sx, cp.sy = int(round(p.sx * cw / W)), int(round(p.sy * ch / H))
    for f in ("N", "iterations", "jfMax", "P", "fsize", "patchArea", "sigma", "sEPE", "sI", "sS", "doubleProp",
              "filter", "tInit"):
        setattr(cp, f, getattr(p, f))
    c1 = np.ascontiguousarray(pr.frame1[y0:y0 + ch, x0:x0 + cw])
    c2 = np.ascontiguousarray(pr.frame2[y0:y0 + ch, x0:x0 + cw])
    oracle.set_num_threads(1)
    try:
        t0 = time.perf_counter()
        oracle.flow_pair(cp, c1, c2, seed=seed, pair=0)
        t1 = (time.perf_counter() - t0) * (W * H) / (cw * ch)
    finally:
        oracle.set_num_threads(threads)
    total = totals[med]
    cb = {"value": W * H / total / 1e6, "unit": "Mpix/s", "cores": threads, "kind": "port",
          "sample": "same %dx%d pair: gray, up-sampling, BRIEF, match, down-sampling, median on the full frame, "
                    "trilateral on %d of %d rows (3 bands, time scaled x%.1f); one warm-up, median of 3 runs"
                    % (W, H, rows, H, H / rows),
          "cpu_model": cpu_model(), "host_logical_cpus": os.cpu_count(),
          "runs_seconds": [round(x, 3) for x in totals], "stage_seconds_median": stage_med,
          "one_thread": {"value": W * H / t1 / 1e6, "unit": "Mpix/s", "cores": 1,
                         "sample": "whole oracle path on the central %dx%d crop (%dx%d matched) with the full "
                                   "frame's per-pixel parameters, time scaled x%.1f" % (cw, ch, cp.sx, cp.sy,
                                                                                       W * H / (cw * ch)),
                         "seconds_scaled": round(t1, 2)},
          "seconds_measured": round(measured + t1 * (cw * ch) / (W * H), 2)}
    return cb, parity


def measure_pair(a, bm, name, tune_cache=None):
    """One workload: autotuned launch shapes, warm-up, K timed steps with HIP
    events around every kernel; per-stage times and roofline fractions."""
    from briefmatch_amd import synth
    wl = synth.WORKLOADS[name]
    pr = synth.workload_pair(name, seed=a.seed)
    W, H = wl["W"], wl["H"]
    p = bm.derive_params(W, H, wl["up"], wl["up"], N=wl["N"])
    ctx = bm.Context(p, seed=1)
    p = ctx.params
    rgb1, rgb2 = bm.DeviceBuffer(array=pr.frame1), bm.DeviceBuffer(array=pr.frame2)
    U, V = ctx.image_buffer(), ctx.image_buffer()
    tuning = ctx.tuning().as_dict() if a.no_autotune else ctx.autotune(rgb1, rgb2, tune_cache).as_dict()
    for _ in range(a.warmup):
        ctx.flow_pair(rgb1, rgb2, U, V)
    ctx.sync()
    # timed region: K steps of the isolated pair
    ctx.timer_start()
    for _ in range(a.steps):
        ctx.flow_pair(rgb1, rgb2, U, V)
    total_ms = ctx.timer_stop()
    ms_per_step = total_ms / a.steps
    # stage times: the same K steps again with HIP events around every kernel
    # on the context stream (the events cost a few microseconds per launch,
    # which matters only for small frames)
    ctx.profile_enable(True)
    ctx.profile_read()
    ctx.timer_start()
    for _ in range(a.steps):
        ctx.flow_pair(rgb1, rgb2, U, V)
    prof_ms = ctx.timer_stop() / a.steps
    prof = ctx.profile_read()
    ctx.profile_enable(False)

    s_tri = int(np.ceil(2.5 * float(np.sqrt(np.float32(p.sS) * np.float32(p.sS)))))
    work = rl.stage_work(p, ctx.NSP, s_tri, tuning.get("brief_frame", 0))
    # the per-frame stages' work is per frame; one launch may cover both frames
    # of the pair (the gray conversion, and BRIEF in its fused mode)
    for st in ("gray", "brief", "upsample"):
        if st in work and prof.get(st, (0, 0))[1]:
            frames = 2.0 * a.steps / prof[st][1]
            work[st] = {k: v * frames for k, v in work[st].items()}
    stages = {}
    for st, (ms, n) in prof.items():
        if n == 0:
            continue
        d = {"ms_per_step": ms / a.steps, "launches_per_step": n / a.steps, "avg_launch_ms": ms / n}
        if st in work:
            d["roofline"] = rl.roofline_of(work[st], ms / n / 1e3)
        stages[st] = d
    dom = max((k for k in stages if k in work), key=lambda k: stages[k]["ms_per_step"])
    roof = rl.roofline_of(work[dom], stages[dom]["avg_launch_ms"] / 1e3)
    roof["kernel"] = dom
    roof["work_per_launch"] = {k: v for k, v in work[dom].items()}
    roof["traffic"] = load_traffic(name, dom)
    t_roof = sum(rl.t_peak(work[k]) * st["launches_per_step"] for k, st in stages.items() if k in work)
    roof["pipeline"] = {"t_roof_ms": round(t_roof * 1e3, 4), "t_measured_ms": round(ms_per_step, 4),
                        "frac": round(t_roof * 1e3 / ms_per_step, 4)}
    ctx.flow_pair(rgb1, rgb2, U, V)
    ctx.sync()
    gU, gV = U.download((H, W), np.float32), V.download((H, W), np.float32)
    e_all, e_noc = bm.epe(gU, gV, pr.gt, pr.occ)
    replay = graph_replay(a, bm, p, tuning, rgb1, rgb2, gU, gV, W, H)
    out = {"value": round(W * H / (ms_per_step * 1e3), 3), "unit": "Mpix/s", "ms_per_step": round(ms_per_step, 4),
           "ms_per_step_with_stage_events": round(prof_ms, 4), "graph_replay": replay,
           "config": {"workload": name, "input": "%dx%d RGB uint8" % (W, H), "matched": "%dx%d" % (p.sx, p.sy),
                      "feature_bits": p.N, "iterations": p.iterations, "jfMax": p.jfMax, "median": p.fsize,
                      "trilateral_window": 2 * s_tri, "mode": "isolated pair, inputs resident in HBM",
                      "tuning": tuning},
           "roofline": roof, "stages": stages, "epe": {"all": round(e_all, 4), "noc": round(e_noc, 4)}}
    return out, dict(ctx=ctx, p=p, pr=pr, rgb1=rgb1, rgb2=rgb2, U=U, V=V, gU=gU, gV=gV, tuning=tuning, W=W, H=H)


def graph_replay(a, bm, p, tuning, rgb1, rgb2, gU, gV, W, H):
    """The same isolated pair captured once as a hipGraph (a one-pair batch in
    graph mode, same seed and launch shapes) and replayed K times: what the
    launch sequence costs beyond its kernels.  Checked bit for bit against
    bm_flow_pair's output."""
    b = bm.Batch([p], seed=1, streams=1)
    b.ctx[0].set_tuning(**tuning)
    U, V = b.ctx[0].image_buffer(), b.ctx[0].image_buffer()
    for _ in range(max(1, a.warmup)):
        b.flow([rgb1], [rgb2], [U], [V], mode=bm.BATCH_GRAPH)
    b.sync()
    b.timer_start()
    for _ in range(a.steps):
        b.flow([rgb1], [rgb2], [U], [V], mode=bm.BATCH_GRAPH)
    ms = b.timer_stop() / a.steps
    same = bool(np.array_equal(U.download((H, W), np.float32), gU) and
                np.array_equal(V.download((H, W), np.float32), gV))
    b.close()
    return {"ms_per_step": round(ms, 4), "value": round(W * H / (ms * 1e3), 3), "bit_exact_vs_flow_pair": same}


def bench_streaming(a, s):
    ctx, rgb1, rgb2, U, V, W, H = s["ctx"], s["rgb1"], s["rgb2"], s["U"], s["V"], s["W"], s["H"]
    ctx.stream_frame(rgb1, U, V)
    for _ in range(2):
        ctx.stream_frame(rgb2, U, V)
        ctx.stream_frame(rgb1, U, V)
    ctx.sync()
    ns = max(a.steps, 4)
    ctx.timer_start()
    for i in range(ns):
        ctx.stream_frame(rgb2 if i % 2 == 0 else rgb1, U, V)
    st_ms = ctx.timer_stop() / ns
    # pipelined: each frame's medians and trilateral filter overlap the next
    # frame's descriptors and matching (identical results)
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
    ctx.stream_reset()
    return {"ms_per_frame": round(st_ms, 4), "value": round(W * H / (st_ms * 1e3), 3), "unit": "Mpix/s",
            "pipelined": {"ms_per_frame": round(sp_ms, 4), "value": round(W * H / (sp_ms * 1e3), 3)}}


def _level_launches(prof, steps):
    return (prof.get("level", (0, 0))[1] + prof.get("level_rs", (0, 0))[1]) / steps


# The learned path's configurations: (iterations, jump range, local search
# radius in input pixels, samples).  "learned" is the path bench.py reports as
# the learned leg: one propagation iteration over the jump range 8 (4 level
# launches) after the seeded initialisation and a local search of 6 samples
# around the seeds; the others show the full jump range (9 levels, no local
# search) and no propagation at all.
LEARNED_CONFIGS = (("learned", 1, 8, 16.0, 6), ("full_jump_range", 1, 256, 0.0, 0),
                   ("without_propagation", 0, 256, 0.0, 0))


def bench_learned(a, bm, s, plain_out):
    """The pair through bm_flow_pair_learned (the committed network) in each
    of LEARNED_CONFIGS, timed like the plain path; EPE next to the plain
    method's and next to the same path with an all-zero network (whose
    proposal is the cost-volume argmin alone), so the ratio is what the
    trained network itself gains; executed level launches per pair against
    the plain method's."""
    from briefmatch_amd import learned
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        return None
    p, pr, W, H = s["p"], s["pr"], s["W"], s["H"]
    nets = {"trained": bm.Net(path=learned.DEFAULT_WEIGHTS),
            "zero_network": bm.Net(weights=np.zeros(bm.NET_PARAMS, np.float32))}
    plain_levels = _level_launches(plain_out["_prof"], a.steps)
    out = {}
    for name, iters, jf, ls_px, ls_n in LEARNED_CONFIGS:
        pf = p.copy()
        pf.iterations = iters
        pf.jfMax = jf
        ls_r = learned.local_search_radius(pf, ls_px) if ls_n else 0
        res = {}
        for label, net in nets.items():
            ctx = bm.Context(pf, seed=1)
            ctx.set_tuning(**s["tuning"])
            ctx.set_local_search(ls_r, ls_n)
            U, V = ctx.image_buffer(), ctx.image_buffer()
            for _ in range(a.warmup):
                ctx.flow_pair_learned(net, s["rgb1"], s["rgb2"], U, V)
            ctx.sync()
            ctx.timer_start()
            for _ in range(a.steps):
                ctx.flow_pair_learned(net, s["rgb1"], s["rgb2"], U, V)
            total_ms = ctx.timer_stop()
            ctx.profile_enable(True)
            ctx.profile_read()
            for _ in range(a.steps):
                ctx.flow_pair_learned(net, s["rgb1"], s["rgb2"], U, V)
            prof = ctx.profile_read()
            ctx.profile_enable(False)
            e_all, e_noc = bm.epe(U.download((H, W), np.float32), V.download((H, W), np.float32), pr.gt, pr.occ)
            res[label] = dict(ms=total_ms / a.steps, prof=prof, e_all=e_all, e_noc=e_noc)
            ctx.close()
        t, z = res["trained"], res["zero_network"]
        out[name] = {
            "iterations": iters, "jfMax": jf,
            "local_search": {"radius_input_px": ls_px, "radius_matched_px": ls_r, "samples": ls_n},
            "ms_per_step": round(t["ms"], 4), "value": round(W * H / (t["ms"] * 1e3), 3), "unit": "Mpix/s",
            "propagation_level_launches_per_pair": _level_launches(t["prof"], a.steps),
            "epe": {"all": round(t["e_all"], 4), "noc": round(t["e_noc"], 4)},
            "epe_ratio_vs_plain": round(t["e_all"] / plain_out["epe"]["all"], 4),
            "zero_network_epe": {"all": round(z["e_all"], 4), "noc": round(z["e_noc"], 4)},
            "epe_ratio_vs_zero_network": round(t["e_all"] / z["e_all"], 4),
            "stages_ms_per_step": {k: round(v[0] / a.steps, 4) for k, v in t["prof"].items() if v[1]}}
    for net in nets.values():
        net.close()
    head = out.pop("learned")
    head.update({"plain_iterations": p.iterations, "plain_epe": plain_out["epe"],
                 "plain_propagation_level_launches_per_pair": plain_levels,
                 "note": "the plain method executes one propagation pass of 9 levels (its first iteration is the "
                         "identity after its zero initialisation and runs as its random search inside the "
                         "initialisation kernel); the learned path seeds the field, searches around the seeds and "
                         "propagates over the jump range 8 (4 levels)",
                 "weights": os.path.relpath(learned.DEFAULT_WEIGHTS, ROOT)})
    head.update(out)
    return head


def bench_batch(a, bm, n_pairs, specs=None, frames=None):
    """Uneven pairs (sizes uniform in [320x240, 1920x1080], random aspect,
    32/128-bit, up-sampling 1.5-2.9): every batch mode timed, results compared
    bit for bit between the modes."""
    from briefmatch_amd import synth
    t0 = time.perf_counter()
    if specs is None:
        specs = synth.batch_specs(n_pairs, a.seed)
        frames = synth.batch_pairs(specs)
    gen_s = time.perf_counter() - t0
    params = [bm.derive_params(s["W"], s["H"], s["up"], s["up"], N=s["N"]) for s in specs]
    batch = bm.Batch(params, seed=1, streams=8)
    r1 = [bm.DeviceBuffer(array=f.frame1) for f in frames]
    r2 = [bm.DeviceBuffer(array=f.frame2) for f in frames]
    Us = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p in params]
    Vs = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p in params]
    pix = sum(p.sx_in * p.sy_in for p in params)
    modes = [("per_pair", bm.BATCH_PER_PAIR), ("graph", bm.BATCH_GRAPH), ("multi", bm.BATCH_MULTI)]
    out, fields = {}, {}
    t0 = time.perf_counter()
    sched = batch.autotune(r1, r2, Us, Vs)
    tune_s = time.perf_counter() - t0
    modes.append(("autotuned", bm.BATCH_AUTO))
    for name, mode in modes:
        if name == "graph":                         # the modes with their default schedules
            batch.set_schedule(mode=mode, branches=8, order=0, persist=0)
        elif name == "multi":
            batch.set_schedule(mode=mode, branches=1, order=0, persist=0)
        elif name == "autotuned":
            batch.set_schedule(**sched.as_dict())
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
    ref = fields["per_pair"]
    exact = {name: all(np.array_equal(x[0][0], y[0][0]) and np.array_equal(x[0][1], y[0][1]) and
                       np.array_equal(x[1], y[1]) for x, y in zip(ref, fields[name]))
             for name in fields if name != "per_pair"}
    batch.close()
    best = "autotuned"
    return {"value": out[best]["value"], "unit": "Mpix/s", "ms_per_step": out[best]["ms_per_batch"],
            "mode": best, "schedule": sched.as_dict(), "autotune_seconds": round(tune_s, 2), "modes": out,
            "bit_exact_vs_per_pair": exact,
            "config": {"workload": "batch%d" % n_pairs, "pairs": n_pairs, "input_mpix": round(pix / 1e6, 3),
                       "matched_mpix": round(sum(p.sx * p.sy for p in params) / 1e6, 3),
                       "feature_bits": sorted(set(int(p.N) for p in params)), "streams": 8,
                       "generator_seconds": round(gen_s, 1)}}


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default=MAIN_WORKLOAD)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU oracle baseline and parity leg")
    ap.add_argument("--tri-rows", type=int, default=68,
                    help="trilateral rows of each timed CPU oracle run (1/16 of the frame at 1080p)")
    ap.add_argument("--copy-calibrate", type=int, default=0, metavar="BYTES")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes")
    ap.add_argument("--tune-cache", default=None, help="autotuner cache file (reused per device and geometry)")
    ap.add_argument("--pairs", type=int, default=64, help="pairs of the batch workload")
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-proposal leg")
    ap.add_argument("--no-streaming", action="store_true", help="skip the sequence-mode leg")
    ap.add_argument("--no-legs", action="store_true", help="skip the other workloads")
    a = ap.parse_args()
    if a.gpus != 1:
        print("bench.py: this build measures one GPU only (--gpus %d requested); the flow path shards by "
              "frame pair with no communication, so N GPUs run N independent copies of this benchmark."
              % a.gpus, file=sys.stderr)
        sys.exit(2)

    # the batch leg's pairs are generated first, in spawned worker processes
    batch_specs = batch_frames = None
    if a.workload.startswith("batch") or not a.no_legs:
        from briefmatch_amd import synth
        n = int(a.workload[5:]) if a.workload.startswith("batch") and a.workload[5:].isdigit() else a.pairs
        t0 = time.perf_counter()
        batch_specs = synth.batch_specs(n, a.seed)
        batch_frames = synth.batch_pairs(batch_specs)
        gen_s = time.perf_counter() - t0

    import briefmatch_amd as bm

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

    def base_line(workload, value, ms):
        return {"metric": METRIC, "value": value, "unit": "Mpix/s", "n_gpus": 1, "steps": a.steps,
                "warmup": a.warmup, "ms_per_step": ms, "higher_is_better": True, "scaling": "weak",
                "vs_baseline": None, "dtype": "u32+f32", "data": "synthetic"}

    if a.workload.startswith("batch"):
        b = bench_batch(a, bm, len(batch_specs), batch_specs, batch_frames)
        b["config"]["generator_seconds"] = round(gen_s, 1)
        out = base_line(a.workload, b["value"], b["ms_per_step"])
        out["metric"] = "batched flow Mpix/s (%d uneven pairs)" % len(batch_specs)
        out.update({k: v for k, v in b.items() if k not in ("value", "unit", "ms_per_step")})
        print(json.dumps(out))
        return

    main_out, s = measure_pair(a, bm, a.workload, a.tune_cache)
    prof_main = {k: (v["ms_per_step"] * a.steps, v["launches_per_step"] * a.steps) for k, v in main_out["stages"].items()}
    out = base_line(a.workload, main_out["value"], main_out["ms_per_step"])
    out["config"] = main_out["config"]
    out["roofline"] = main_out["roofline"]
    out["stages"] = main_out["stages"]
    out["ms_per_step_with_stage_events"] = main_out["ms_per_step_with_stage_events"]
    out["graph_replay"] = main_out["graph_replay"]
    out["epe"] = main_out["epe"]
    if not a.no_streaming:
        out["streaming"] = bench_streaming(a, s)
    if not a.no_learned:
        main_out["_prof"] = prof_main
        out["learned"] = bench_learned(a, bm, s, main_out)
    if not a.no_cpu:
        s["ctx"].flow_pair(s["rgb1"], s["rgb2"], s["U"], s["V"])
        s["ctx"].sync()
        off, cost = s["ctx"].field()
        cb, parity = cpu_baseline(s["pr"], s["p"], 1, off, cost, s["gU"], s["gV"], a.tri_rows)
        out["cpu_baseline"] = cb
        out["parity"] = parity
        out["gpu_over_cpu"] = round(out["value"] / cb["value"], 1)
    else:
        out["cpu_baseline"] = None
    s["ctx"].close()
    if not a.no_legs:
        legs = {}
        for name in LEG_WORKLOADS:
            if name == a.workload:
                continue
            lo, ls = measure_pair(a, bm, name, a.tune_cache)
            ls["ctx"].close()
            legs[name] = lo
        b = bench_batch(a, bm, len(batch_specs), batch_specs, batch_frames)
        b["config"]["generator_seconds"] = round(gen_s, 1)
        legs[b["config"]["workload"]] = b
        out["legs"] = legs
    print(json.dumps(out))


if __name__ == "__main__":
    main()
