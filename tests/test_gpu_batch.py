"""Tuning invariance, the autotuner and batches of uneven pairs on the GPU.

Every valid tuning must give bit-identical fields and flows, except for the
trilateral filter's kernel and split, which add the same terms in another
order; a batch (per-pair launches or one hipGraph) must give, for pair i,
exactly bm_flow_pair(ctx_i, pair = i), whose field in turn equals the CPU
oracle."""
import itertools

import numpy as np
import pytest

import briefmatch_amd as bm
import oracle
from briefmatch_amd import synth
from conftest import say

pytestmark = pytest.mark.gpu


def _flow(ctx, pr, pair=0):
    U, V = ctx.flow_numpy(pr.frame1, pr.frame2, pair=pair)
    off, cost = ctx.field()
    return off, cost, U, V


@pytest.mark.parametrize("N", [32, 128, 256, 512])
def test_tuning_does_not_change_results(N):
    say("tuning invariance", N)
    pr = synth.make_pair(150, 100, motion="rects", max_motion=8.0, seed=N)
    p = bm.derive_params(150, 100, 2.9, 2.9, N=N)
    ctx = bm.Context(p, seed=3)
    ref = _flow(ctx, pr)
    t0 = ctx.tuning()
    px = (1, 2, 4) if N <= 128 else ((1, 2) if N <= 256 else (1,))
    for lp, strip, near, band, frame in itertools.product(px, (16, 256), (0, 8, 32, 48), (1, t0.tri_band), (0, 1)):
        if frame and not near:
            continue
        ctx.set_tuning(level_px=lp, strip_rows=strip, brief_near=near, tri_band=band, brief_frame=frame)
        got = _flow(ctx, pr)
        for a, b in zip(ref, got):
            assert np.array_equal(a, b), (lp, strip, near, band, frame)
    ctx.set_tuning(brief_frame=0)
    # the trilateral kernels and splits sum in different orders: equal within the flow tolerance
    for groups in (0, 2, 4):
        for split in (1, 2, 8):
            ctx.set_tuning(tri_groups=groups, tri_split=split)
            got = _flow(ctx, pr)
            for a, b in zip(ref[:2], got[:2]):
                assert np.array_equal(a, b), (groups, split)
            for a, b in zip(ref[2:], got[2:]):
                assert np.abs(a - b).max() <= 1e-3, (groups, split)


@pytest.mark.parametrize("wl,frame,near", [("c1_320x240_u1.5_n32", 0, None), ("c2_1920x1080_u2.9_n128", 1, 32),
                                           ("c3_1280x720_u4.9_n512", 1, 48)])
def test_default_tuning_by_size(wl, frame, near):
    """A new context's BRIEF mode follows its matched frame: fused below 2^21
    pixels, the up-sampled frame in memory above, with the 48-px near halo at
    512 bits (where the pattern reaches past 32 px); the trilateral filter is
    not split until a tuning asks for it."""
    w = synth.WORKLOADS[wl]
    p = bm.derive_params(w["W"], w["H"], w["up"], w["up"], N=w["N"])
    ctx = bm.Context(p, seed=1)
    t = ctx.tuning()
    assert t.brief_frame == frame, t.as_dict()
    if near is not None:
        assert t.brief_near == near, t.as_dict()
    assert t.tri_split == 1
    ctx.close()


def test_tuning_rejects_invalid():
    p = bm.derive_params(64, 48, 2.0, 2.0, N=512)
    ctx = bm.Context(p)
    for kw in ({"level_px": 3}, {"level_px": 2}, {"level_px": 4}, {"strip_rows": 8}, {"brief_near": 12},
               {"tri_band": 0}, {"tri_groups": 1}, {"tri_groups": 8}, {"brief_frame": 2},
               {"tri_split": 0}, {"tri_split": 3}, {"tri_split": 16},
               {"brief_near": 0, "brief_frame": 1}):
        with pytest.raises(bm.BriefMatchError):
            ctx.set_tuning(**kw)


def test_autotune_and_cache(tmp_path):
    say("autotune")
    pr = synth.make_pair(320, 240, motion="rects", seed=5)
    p = bm.derive_params(320, 240, 2.9, 2.9, N=128)
    ctx = bm.Context(p, seed=1)
    ref = _flow(ctx, pr)
    r1, r2 = bm.DeviceBuffer(array=pr.frame1), bm.DeviceBuffer(array=pr.frame2)
    cache = str(tmp_path / "tune.txt")
    t = ctx.autotune(r1, r2, cache)
    assert t.level_px in (1, 2, 4) and 16 <= t.strip_rows <= 4096
    assert len(open(cache).read().splitlines()) == 1
    ctx2 = bm.Context(p, seed=1)
    t2 = ctx2.autotune(r1, r2, cache)                  # served from the cache
    assert t2.as_dict() == t.as_dict()
    assert len(open(cache).read().splitlines()) == 1
    got = _flow(ctx, pr)
    for a, b in zip(ref[:2], got[:2]):
        assert np.array_equal(a, b)
    # the flow up to the trilateral filter's summation order (kernel and split)
    for a, b in zip(ref[2:], got[2:]):
        assert np.abs(a - b).max() <= 1e-3


def _batch_workload(n, seed):
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        W, H = int(rng.integers(64, 200)), int(rng.integers(48, 150))
        N = int(rng.choice([32, 128]))
        u = float(rng.uniform(1.5, 2.9))
        pr = synth.make_pair(W, H, motion="rects", max_motion=6.0, seed=seed * 100 + i)
        out.append((bm.derive_params(W, H, u, u, N=N), pr))
    return out


def test_batch_equals_per_pair_and_oracle():
    say("batch")
    wl = _batch_workload(6, 7)
    # one pair with a different jump range and iteration count (its own level
    # group in the multi-pair launches), one without refinement
    wl[2][0].jfMax, wl[2][0].iterations = 64, 3
    wl[4][0].filter = 0
    batch = bm.Batch([p for p, _ in wl], seed=11, streams=3)
    r1 = [bm.DeviceBuffer(array=pr.frame1) for _, pr in wl]
    r2 = [bm.DeviceBuffer(array=pr.frame2) for _, pr in wl]
    Us = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p, _ in wl]
    Vs = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p, _ in wl]
    results = {}
    # graphs captured, then replayed; before each run the swapped pairs go
    # through the per-pair path, so no mode can pass on a buffer a previous
    # mode left behind
    for mode in (bm.BATCH_PER_PAIR, bm.BATCH_GRAPH, bm.BATCH_GRAPH, bm.BATCH_MULTI, bm.BATCH_MULTI):
        batch.flow(r2, r1, Us, Vs, mode=bm.BATCH_PER_PAIR)
        batch.flow(r1, r2, Us, Vs, mode=mode)
        batch.sync()
        res = []
        for i, (p, _) in enumerate(wl):
            off, cost = batch.ctx[i].field()
            res.append((off, cost, Us[i].download((p.sy_in, p.sx_in), np.float32),
                        Vs[i].download((p.sy_in, p.sx_in), np.float32)))
        results.setdefault(mode, []).append(res)
    base = results[bm.BATCH_PER_PAIR][0]
    for mode in (bm.BATCH_GRAPH, bm.BATCH_MULTI):
        for rep, res in enumerate(results[mode]):
            for i, (a, b) in enumerate(zip(base, res)):
                for k, (x, y) in enumerate(zip(a, b)):
                    assert np.array_equal(x, y), ("mode", mode, "run", rep, "pair", i, "array", k,
                                                  "differing", int((x != y).sum()), "of", x.size)
    for i, (p, pr) in enumerate(wl):
        ctx = bm.Context(p, seed=11)
        single = _flow(ctx, pr, pair=i)
        for x, y in zip(base[i], single):
            assert np.array_equal(x, y), i
        op = oracle.Params()
        for n_, _ in op._fields_:
            setattr(op, n_, getattr(ctx.params, n_))
        eo, ec, oU, oV = oracle.flow_pair(op, pr.frame1, pr.frame2, seed=11, pair=i)
        assert np.array_equal(base[i][0], eo) and np.array_equal(base[i][1], ec), i
        assert np.abs(base[i][2] - oU).max() <= 1e-3 and np.abs(base[i][3] - oV).max() <= 1e-3


def test_batch_graph_follows_configuration_changes():
    """A captured batch graph bakes in launch arguments (near halo, level
    kernel shape, up-sampler, trilateral split) and the BRIEF plan's layout: changing a member
    context's tuning, pattern or up-sampler between two graph launches must
    re-capture, so each graph result equals the per-pair launches under the
    same configuration."""
    say("batch graph after configuration changes")
    wl = _batch_workload(3, 9)
    batch = bm.Batch([p for p, _ in wl], seed=5, streams=2)
    r1 = [bm.DeviceBuffer(array=pr.frame1) for _, pr in wl]
    r2 = [bm.DeviceBuffer(array=pr.frame2) for _, pr in wl]
    Us = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p, _ in wl]
    Vs = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p, _ in wl]

    def run(mode):
        batch.flow(r2, r1, Us, Vs, mode=bm.BATCH_PER_PAIR)   # other contents in every buffer first
        batch.flow(r1, r2, Us, Vs, mode=mode)
        batch.sync()
        return [(batch.ctx[i].field(), Us[i].download((p.sy_in, p.sx_in), np.float32))
                for i, (p, _) in enumerate(wl)]

    def same(a, b):
        return all(np.array_equal(x[0][0], y[0][0]) and np.array_equal(x[0][1], y[0][1]) and
                   np.array_equal(x[1], y[1]) for x, y in zip(a, b))

    for mode in (bm.BATCH_GRAPH, bm.BATCH_MULTI):
        run(mode)                                        # captured with the initial configuration
        c = batch.ctx[1]
        t = c.tuning()
        changes = [lambda: c.set_tuning(brief_near=16 if t.brief_near != 16 else 24, strip_rows=64),
                   lambda: c.set_pattern(np.random.default_rng(3 + mode).integers(-12, 13, 4 * c.params.N)
                                         .astype(np.int16)),
                   lambda: c.set_upsampler(bm.UPSAMPLE_LANCZOS3),   # the multi mode's per-pair path for this pair
                   lambda: c.set_upsampler(bm.UPSAMPLE_BILINEAR),
                   lambda: c.set_tuning(tri_split=4),   # the split filter: the multi mode's per-pair path
                   lambda: c.set_tuning(tri_split=1)]
        for k, change in enumerate(changes):
            before = run(mode)
            change()
            got = run(mode)
            want = run(bm.BATCH_PER_PAIR)
            assert same(got, want), (mode, k)
            # pattern / up-sampler change the result
            assert not same(before[1:2], got[1:2]) or k == 0 or k >= 4, (mode, k)


def test_batch_schedules_and_autotune():
    """Every schedule of the multi-pair and graph modes (job order, resident
    work-groups, branches) gives the per-pair launches' results bit for bit;
    the batch autotuner returns a valid schedule that BATCH_AUTO then runs."""
    say("batch schedules")
    wl = _batch_workload(5, 13)
    batch = bm.Batch([p for p, _ in wl], seed=3, streams=4)
    r1 = [bm.DeviceBuffer(array=pr.frame1) for _, pr in wl]
    r2 = [bm.DeviceBuffer(array=pr.frame2) for _, pr in wl]
    Us = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p, _ in wl]
    Vs = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p, _ in wl]

    def run(mode):
        batch.flow(r2, r1, Us, Vs, mode=bm.BATCH_PER_PAIR)   # other contents in every buffer first
        batch.flow(r1, r2, Us, Vs, mode=mode)
        batch.sync()
        return [(batch.ctx[i].field(), Us[i].download((p.sy_in, p.sx_in), np.float32),
                 Vs[i].download((p.sy_in, p.sx_in), np.float32)) for i, (p, _) in enumerate(wl)]

    def same(a, b):
        return all(np.array_equal(x[0][0], y[0][0]) and np.array_equal(x[0][1], y[0][1]) and
                   np.array_equal(x[1], y[1]) and np.array_equal(x[2], y[2]) for x, y in zip(a, b))

    want = run(bm.BATCH_PER_PAIR)
    for order in (0, 1):
        for lanes in (1, 2, 4):
            for persist in (0, 1, 4):
                batch.set_schedule(order=order, branches=lanes, persist=persist)
                assert same(run(bm.BATCH_MULTI), want), (order, lanes, persist)
        for branches in (1, 4):
            batch.set_schedule(order=order, branches=branches, persist=0)
            assert same(run(bm.BATCH_GRAPH), want), (order, branches)
    s = batch.autotune(r1, r2, Us, Vs)
    assert s.mode in (bm.BATCH_GRAPH, bm.BATCH_MULTI) and 1 <= s.branches <= 4
    assert batch.schedule().as_dict() == s.as_dict()
    assert same(run(bm.BATCH_AUTO), want)
    with pytest.raises(bm.BriefMatchError):
        batch.set_schedule(branches=5)
    with pytest.raises(bm.BriefMatchError):
        batch.set_schedule(mode=7)
