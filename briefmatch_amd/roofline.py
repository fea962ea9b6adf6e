"""Work per launch of each stage of the flow path and the MI355X ceilings it is
measured against (bench.py's `roofline` and `stages` objects).

Every stage has up to three algorithmic work figures per launch:
  hbm   compulsory bytes: each array read once and written once;
  lds   LDS bytes the algorithm itself needs (BRIEF: two 4-byte samples per
        test; the tile fill is overhead, not counted);
  valu  lane-operations on the vector ALUs (BRIEF: subtract + shift-in per
        test; median: the min/max instructions of its selection network;
        trilateral: 15 flop per tap, counted as FLOP against the FP32 peak).
The bound of a stage is the ceiling with the largest time at peak.

Peaks (published or derived from published figures, one MI355X):
  HBM3E 8 TB/s; FP32 157.3 TFLOP/s; LDS 256 CU x 128 B/clk x 2.4 GHz =
  78.6 TB/s; VALU issue 256 CU x 64 lanes x 2.4 GHz = 39.3 T lane-op/s.
"""

HBM_PEAK_GBS = 8000.0
FP32_PEAK_TFLOPS = 157.3
LDS_PEAK_GBS = 256 * 128 * 2.4
VALU_PEAK_TOPS = 256 * 64 * 2.4e9 / 1e12
TRI_FLOPS_PER_TAP = 15         # 3 sub + 3 fma + exp2 + add + 2 fma (fma = 2)
BRIEF_VALU_PER_TEST = 2        # sign of the sample difference shifted into the word


def median_minmax(fs):
    """min/max instructions of the pruned median-selection network for a
    window of fs values: Batcher's odd-even merge network on the next power of
    two, +inf padding folded away, pruned backwards from the middle wire (the
    construction of the median kernel)."""
    n = 2
    while n < fs:
        n *= 2
    comps = []
    p = 1
    while p < n:
        k = p
        while k >= 1:
            j = k % p
            while j + k < n:
                for i in range(min(k, n - j - k)):
                    if (i + j) // (2 * p) == (i + j + k) // (2 * p):
                        comps.append((i + j, i + j + k))
                j += 2 * k
            k //= 2
        p *= 2
    inf = [w >= fs for w in range(n)]
    live = []
    for a, b in comps:
        if inf[b]:
            continue
        if inf[a]:
            inf[a], inf[b] = False, True
            live.append((a, b, 0))
            continue
        live.append((a, b, 3))
    need = [False] * n
    need[fs // 2] = True
    minmax = 0
    for a, b, kind in reversed(live):
        if kind == 0:
            if need[a]:
                need[a], need[b] = False, True
            continue
        na, nb = need[a], need[b]
        if not (na or nb):
            continue
        need[a] = need[b] = True
        minmax += int(na) + int(nb)
    return minmax


def stage_work(p, NSP, s_tri, brief_frame=0):
    """Algorithmic work per launch of each stage for parameters p (a bm_params
    with sx, sy, sx_in, sy_in, N, fsize), descriptor stride NSP and trilateral
    half window s_tri; brief_frame 1: the up-sampled frame is written by its
    own launch and read by BRIEF.  Gray, BRIEF and up-sampling are per frame
    (a caller whose launch covers both frames scales them by two)."""
    Win = p.sx_in * p.sy_in
    M = p.sx * p.sy
    b = p.N / 8.0
    taps = Win * (2 * s_tri) ** 2
    return {
        # frame 1 writes gray8 + FP32, frame 2 gray8 only: the mean of the two launches
        "gray": {"hbm": Win * (3 + 1 + 2.0)},
        # BRIEF of one frame: gray8 in (up-sampling fused) or the FP32 frame in,
        # descriptors out
        "brief": {"hbm": (M * 4.0 if brief_frame else Win * 1.0) + M * NSP * 4.0, "lds": M * p.N * 8.0,
                  "valu": M * p.N * BRIEF_VALU_PER_TEST},
        # bilinear up-sampling of one frame into memory: gray8 in, FP32 frame out
        "upsample": {"hbm": Win * 1.0 + M * 4.0},
        "init": {"hbm": M * (2 * b + 6)},
        "level": {"hbm": M * (2 * b + 12)},          # F1, F2 once, field in + out
        "level_rs": {"hbm": M * (2 * b + 12)},
        "downsample": {"hbm": M * 4.0 + Win * 8.0},
        # one direction, both flow components
        "median": {"hbm": Win * 16.0, "valu": Win * 2.0 * median_minmax(p.fsize)},
        "trilateral": {"flop": taps * TRI_FLOPS_PER_TAP, "hbm": Win * 28.0},
    }


def _ceilings(work, avg_s):
    out = {}
    if "hbm" in work:
        a = work["hbm"] / avg_s / 1e9
        out["hbm"] = {"achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": a / HBM_PEAK_GBS,
                      "t_peak_s": work["hbm"] / (HBM_PEAK_GBS * 1e9)}
    if "lds" in work:
        a = work["lds"] / avg_s / 1e9
        out["lds"] = {"achieved": a, "peak": LDS_PEAK_GBS, "unit": "GB/s", "frac": a / LDS_PEAK_GBS,
                      "t_peak_s": work["lds"] / (LDS_PEAK_GBS * 1e9)}
    if "valu" in work:
        a = work["valu"] / avg_s / 1e12
        out["valu"] = {"achieved": a, "peak": VALU_PEAK_TOPS, "unit": "Tlane-op/s", "frac": a / VALU_PEAK_TOPS,
                       "t_peak_s": work["valu"] / (VALU_PEAK_TOPS * 1e12)}
    if "flop" in work:
        a = work["flop"] / avg_s / 1e12
        out["fp32"] = {"achieved": a, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s", "frac": a / FP32_PEAK_TFLOPS,
                       "t_peak_s": work["flop"] / (FP32_PEAK_TFLOPS * 1e12)}
    return out


def roofline_of(work, avg_s):
    """The binding ceiling of a stage (largest time at peak) as the bench's
    roofline object, plus every ceiling's fraction."""
    c = _ceilings(work, avg_s)
    name = max(c, key=lambda k: c[k]["t_peak_s"])
    bound = {"hbm": "hbm", "lds": "lds", "valu": "valu", "fp32": "valu"}[name]
    r = {"bound": bound, "achieved": c[name]["achieved"], "peak": c[name]["peak"], "unit": c[name]["unit"],
         "frac": c[name]["frac"]}
    if len(c) > 1:
        r["ceilings"] = {k: round(v["frac"], 4) for k, v in c.items()}
    return r


def t_peak(work):
    """Time of the stage at its binding ceiling (seconds)."""
    return max(v["t_peak_s"] for v in _ceilings(work, 1.0).values())
