"""Turns rocprofv3 CSV output into the short text summaries kept under profiles/.

    python -m briefmatch_amd.rocprof_summary stats  <kernel_stats.csv> <out.txt> --cmd "<command>"
    python -m briefmatch_amd.rocprof_summary counters <counter_collection.csv> <kernel_stats.csv> <out.txt> --cmd ...
    python -m briefmatch_amd.rocprof_summary memory <mem1.csv> <mem2.csv> <kernel_stats.csv> <out.txt> --cmd ...
    python -m briefmatch_amd.rocprof_summary traffic <fetch_counter_collection.csv> <write_counter_collection.csv>
           <copy_fetch.csv> <copy_write.csv> <copy_bytes> <out.txt> <traffic.json> --cmd "<commands>"

`counters` averages every counter per kernel over its dispatches and adds, from
the kernel statistics of a trace of the same command, the instruction rates
(VALU and INT32 VALU lane-operations per second against the 256 CU x 64 lanes
x 2.4 GHz issue peak of one MI355X).

`traffic` scales each kernel's mean FETCH_SIZE and WRITE_SIZE (KiB per dispatch)
by what the same counter reads for the plain copy kernel (k_copy) of known
size, and merges the HBM bytes per launch of each stage of that workload (all
kernels of the stage: the split trilateral filter's finishing kernel counts
with it) into
traffic.json under "workloads" -> workload -> stage (other workloads kept),
which bench.py reports as roofline.traffic.  Profile a run whose launch shapes
come from an autotuner cache (bench.py --tune-cache), so each stage runs one
kernel variant.
"""
import argparse
import csv
import json
import re
from collections import OrderedDict, defaultdict


def short_name(name):
    n = name.replace("bm::(anonymous namespace)::", "").replace("bm::net::", "").replace("void ", "")
    m = re.match(r"([A-Za-z_0-9]+(?:<[^>]*>)?)", n)
    return m.group(1) if m else n[:40]


STAGE_OF = OrderedDict([("k_gray", "gray"), ("k_upsample", "upsample"), ("k_brief", "brief"),
                        ("k_init_proposal", "proposal"), ("k_local_search", "local_search"), ("k_init", "init"), ("k_level<", "level"), ("k_level_lanes<", "level"), ("k_random_search", "random_search"),
                        ("k_downsample", "downsample"), ("k_median", "median"), ("k_trilateral", "trilateral"),
                        ("k_copy", "copy"), ("k_features", "proposal"),
                        ("k_block_sums", "proposal"), ("k_far_search", "proposal"), ("k_conv", "proposal"),
                        ("k_proposal_offsets", "proposal")])


def stage_of(short):
    for k, v in STAGE_OF.items():
        if short.startswith(k):
            if v == "level" and ", true" in short:
                return "level_rs"
            return v
    return short


def stats(path, out, cmd):
    rows = list(csv.DictReader(open(path)))
    lines = ["# rocprofv3 kernel statistics", "# command: " + cmd, "",
             "%-34s %7s %12s %12s %12s %7s" % ("kernel", "calls", "total_ms", "avg_us", "min_us", "pct")]
    for r in rows:
        lines.append("%-34s %7d %12.3f %12.2f %12.2f %7.2f" % (
            short_name(r["Name"]), int(r["Calls"]), float(r["TotalDurationNs"]) / 1e6,
            float(r["AverageNs"]) / 1e3, float(r["MinNs"]) / 1e3, float(r["Percentage"])))
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


PEAK_LANE_OPS = 256 * 64 * 2.4e9   # one VALU lane-operation per lane per cycle


def counters(path, stats_path, out, cmd):
    acc = defaultdict(lambda: defaultdict(list))
    for r in csv.DictReader(open(path)):
        acc[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    dur = {}
    for r in csv.DictReader(open(stats_path)):
        dur[short_name(r["Name"])] = float(r["AverageNs"]) * 1e-9
    names = sorted({c for k in acc.values() for c in k})
    lines = ["# rocprofv3 counters, mean per dispatch", "# command: " + cmd,
             "# rates use the average duration of the same kernel in the kernel-trace run;",
             "# VALU lane-op peak = 256 CU x 64 lanes x 2.4 GHz = %.1f T/s" % (PEAK_LANE_OPS / 1e12), ""]
    head = "%-30s" % "kernel" + "".join(" %14s" % n[:14] for n in names)
    extra = "int32_Tops/s valu_frac_peak" if any("SQ_INSTS_VALU" in n for n in names) else ""
    lines.append(head + "  " + extra)
    for k in sorted(acc, key=lambda k: -dur.get(k, 0.0)):
        vals = {n: sum(v) / len(v) for n, v in acc[k].items()}
        row = "%-30s" % k + "".join(" %14.4g" % vals[n] if n in vals else " %14s" % "-" for n in names)
        d = dur.get(k)
        if extra and d:
            i32 = vals.get("SQ_INSTS_VALU_INT32", 0.0) * 64 / d
            va = vals.get("SQ_INSTS_VALU", 0.0) * 64 / d
            row += "  %12.2f %14.3f" % (i32 / 1e12, va / PEAK_LANE_OPS)
        lines.append(row)
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def memory(mem1, mem2, stats_path, out, cmd):
    """Cache and address-unit figures of each kernel from two counter passes:
    l2_hit = TCC_HIT / (TCC_HIT + TCC_MISS); ta_busy = TA_TA_BUSY_sum / (256
    TAs x GRBM_GUI_ACTIVE / 8 XCDs), the fraction of cycles the CUs'
    texture-address units are busy; ta_cyc/vmem = TA busy cycles per vector
    memory read instruction; wait = SQ_WAIT_INST_ANY / SQ_WAVE_CYCLES;
    vmem_rd/wave = SQ_INSTS_VMEM_RD / SQ_WAVES."""
    acc = defaultdict(lambda: defaultdict(list))
    for path in (mem1, mem2):
        for r in csv.DictReader(open(path)):
            acc[short_name(r["Kernel_Name"])][r["Counter_Name"]].append(float(r["Counter_Value"]))
    dur = {short_name(r["Name"]): float(r["AverageNs"]) / 1e3 for r in csv.DictReader(open(stats_path))}
    lines = ["# rocprofv3 cache and address-unit counters, mean per dispatch", "# commands: " + cmd,
             "# " + memory.__doc__.split("\n", 1)[1].replace("\n    ", "\n# ").strip(), "",
             "%-28s %9s %8s %12s %8s %10s %6s %12s" % ("kernel", "dur_us", "l2_hit", "l2_requests", "ta_busy",
                                                    "ta_cyc/vmem", "wait", "vmem_rd/wave")]
    for k in sorted(acc, key=lambda k: -dur.get(k, 0.0)):
        v = {n: sum(x) / len(x) for n, x in acc[k].items()}
        if k.startswith("__amd") or "TCC_HIT_sum" not in v or "TA_TA_BUSY_sum" not in v:
            continue
        req = v["TCC_HIT_sum"] + v["TCC_MISS_sum"]
        cyc = v["GRBM_GUI_ACTIVE"] / 8.0
        lines.append("%-28s %9.1f %8.3f %12.4g %8.2f %10.1f %6.2f %12.1f" % (
            k, dur.get(k, float("nan")), v["TCC_HIT_sum"] / req if req else 0.0, req,
            v["TA_TA_BUSY_sum"] / (256 * cyc) if cyc else 0.0,
            v["TA_TA_BUSY_sum"] / v["SQ_INSTS_VMEM_RD"] if v.get("SQ_INSTS_VMEM_RD") else 0.0,
            v["SQ_WAIT_INST_ANY"] / v["SQ_WAVE_CYCLES"] if v.get("SQ_WAVE_CYCLES") else 0.0,
            v["SQ_INSTS_VMEM_RD"] / v["SQ_WAVES"] if v.get("SQ_WAVES") else 0.0))
    with open(out, "w") as f:
        f.write("\n".join(lines) + "\n")


def _mean_by_kernel(path, counts=None):
    acc = defaultdict(list)
    for r in csv.DictReader(open(path)):
        acc[short_name(r["Kernel_Name"])].append(float(r["Counter_Value"]))
    if counts is not None:
        counts.update({k: len(v) for k, v in acc.items()})
    return {k: sum(v) / len(v) for k, v in acc.items()}


def traffic(fetch, write, cfetch, cwrite, copy_bytes, out, out_json, cmd, workload):
    calls = {}
    f = _mean_by_kernel(fetch, calls)
    w = _mean_by_kernel(write)
    cf = _mean_by_kernel(cfetch)
    cw = _mean_by_kernel(cwrite)
    kib = copy_bytes / 1024.0
    sf = kib / cf["k_copy"]
    sw = kib / cw["k_copy"]
    lines = ["# HBM traffic per launch from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE (KiB per dispatch)",
             "# command(s): " + cmd,
             "# calibration: k_copy of %d bytes reads FETCH_SIZE %.1f KiB, WRITE_SIZE %.1f KiB;" % (
                 copy_bytes, cf["k_copy"], cw["k_copy"]),
             "#   scale factors fetch x%.4f, write x%.4f" % (sf, sw), "",
             "%-34s %14s %14s %14s" % ("kernel", "fetch_MB", "write_MB", "total_MB")]
    try:
        with open(out_json) as fo:
            tj = json.load(fo)
    except (OSError, ValueError):
        tj = {}
    tj = {"_calibration": tj.get("_calibration", {}), "workloads": tj.get("workloads", {})}
    tj["_calibration"][workload] = {"copy_bytes": copy_bytes, "fetch_scale": sf, "write_scale": sw}
    per_stage = defaultdict(list)
    for k in f:
        if k == "k_copy" or k.startswith("__amd"):
            continue
        fb, wb = f[k] * sf * 1024, w.get(k, 0.0) * sw * 1024
        lines.append("%-34s %14.1f %14.1f %14.1f" % (k, fb / 1e6, wb / 1e6, (fb + wb) / 1e6))
        per_stage[stage_of(k)].append((fb + wb, calls.get(k, 1)))
    # a stage launch may be several kernels (the split trilateral filter and its
    # finishing kernel): bytes of all its dispatches over the launches, counted
    # as the dispatches of its most frequent kernel
    tj["workloads"][workload] = {st: sum(b * n for b, n in v) / max(n for _, n in v)
                                 for st, v in per_stage.items()}
    with open(out, "w") as fo:
        fo.write("\n".join(lines) + "\n")
    with open(out_json, "w") as fo:
        json.dump(tj, fo, indent=1)


def main():
    ap = argparse.ArgumentParser()
    sub = ap.add_subparsers(dest="mode", required=True)
    a = sub.add_parser("stats")
    a.add_argument("csv")
    a.add_argument("out")
    a.add_argument("--cmd", required=True)
    b = sub.add_parser("traffic")
    for n in ("fetch", "write", "copy_fetch", "copy_write"):
        b.add_argument(n)
    b.add_argument("copy_bytes", type=int)
    b.add_argument("out")
    b.add_argument("out_json")
    b.add_argument("--cmd", required=True)
    b.add_argument("--workload", required=True)
    c = sub.add_parser("counters")
    c.add_argument("csv")
    c.add_argument("stats_csv")
    c.add_argument("out")
    c.add_argument("--cmd", required=True)
    d = sub.add_parser("memory")
    d.add_argument("mem1_csv")
    d.add_argument("mem2_csv")
    d.add_argument("stats_csv")
    d.add_argument("out")
    d.add_argument("--cmd", required=True)
    args = ap.parse_args()
    if args.mode == "memory":
        memory(args.mem1_csv, args.mem2_csv, args.stats_csv, args.out, args.cmd)
    elif args.mode == "stats":
        stats(args.csv, args.out, args.cmd)
    elif args.mode == "counters":
        counters(args.csv, args.stats_csv, args.out, args.cmd)
    else:
        traffic(args.fetch, args.write, args.copy_fetch, args.copy_write, args.copy_bytes, args.out, args.out_json,
                args.cmd, args.workload)


if __name__ == "__main__":
    main()
