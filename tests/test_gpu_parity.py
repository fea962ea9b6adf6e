"""GPU parity: every stage of the HIP path against the CPU oracle, through the C ABI.

Integer stages (gray, up-sampling, BRIEF, init, propagation, random search, the
whole match field) must agree bit for bit; FP32 stages (down-sampling, median,
trilateral, final flow) within 1e-3 px max-abs.
"""
import numpy as np
import pytest

import oracle
import briefmatch_amd as bm
from briefmatch_amd import synth
from conftest import say

pytestmark = pytest.mark.gpu

FLOW_TOL = 1e-3  # px, max-abs of the FP32 flow against the oracle


def _pair(W, H, seed, **gen):
    gen.setdefault("motion", "rects")
    gen.setdefault("max_motion", 6.0)
    return synth.make_pair(W, H, seed=seed, **gen)


def _oracle_params(p):
    op = oracle.Params()
    for n, _ in op._fields_:
        setattr(op, n, getattr(p, n))
    return op


def _descs(ctx, frame, fused=True):
    """Device descriptors of one RGB frame (returns the buffer, gray8 and L)."""
    rgb = bm.DeviceBuffer(array=frame)
    g8 = bm.DeviceBuffer(ctx.W * ctx.H)
    ctx.gray(rgb, g8=g8)
    F = ctx.desc_buffer()
    L = None
    if fused:
        ctx.brief_gray(g8, F)
    else:
        L = bm.DeviceBuffer(ctx.sx * ctx.sy * 4)
        ctx.upsample(g8, L)
        ctx.brief(L, F)
    ctx.sync()
    return F, g8, L


def _oracle_tables(ctx, p, seed):
    """The oracle's own BRIEF pattern and Gaussian kernel for this
    configuration, asserted equal to the tables the product built: the oracle
    never takes its inputs from the product."""
    pat = oracle.make_pattern(seed, p.N, p.patchArea)
    kern = oracle.gauss_kernel(p.sigma)
    assert np.array_equal(ctx.pattern(), pat), "BRIEF pattern differs from the oracle's"
    assert np.array_equal(ctx.gauss_kernel(), kern), "Gaussian kernel differs from the oracle's"
    return pat, kern


def _oracle_descs(frame, p, pattern):
    g = oracle.gray(frame)
    L = oracle.upsample(g, p.sx, p.sy)
    return oracle.brief(L, pattern, p.N), g, L


def _pad(F, NSP):
    sy, sx, NS = F.shape
    out = np.zeros((sy, sx, NSP), np.uint32)
    out[..., :NS] = F
    return out


@pytest.mark.parametrize("W,H,up,N", [(320, 240, 1.5, 32), (48, 32, 2.0, 512), (61, 37, 2.3, 72),
                                      (96, 64, 1.0, 128), (80, 50, 3.1, 100), (70, 44, 2.4, 200)])
def test_gray_upsample_brief(W, H, up, N):
    say("gray/upsample/brief", W, H, up, N)
    pr = _pair(W, H, seed=W + N)
    p = bm.derive_params(W, H, up, up, N=N)
    ctx = bm.Context(p, seed=11)
    pat, _ = _oracle_tables(ctx, p, 11)
    rgb = bm.DeviceBuffer(array=pr.frame1)
    g8, g32 = bm.DeviceBuffer(W * H), bm.DeviceBuffer(W * H * 4)
    ctx.gray(rgb, g8=g8, g32=g32)
    L = bm.DeviceBuffer(p.sx * p.sy * 4)
    ctx.upsample(g8, L)
    F, Ff = ctx.desc_buffer(), ctx.desc_buffer()
    ctx.brief(L, F)
    ctx.brief_gray(g8, Ff)
    ctx.sync()
    og = oracle.gray(pr.frame1)
    assert np.array_equal(g8.download((H, W), np.uint8), og)
    assert np.array_equal(g32.download((H, W), np.float32), og.astype(np.float32))
    oL = oracle.upsample(og, p.sx, p.sy)
    assert np.array_equal(L.download((p.sy, p.sx), np.float32), oL)
    oF = _pad(oracle.brief(oL, pat, N), ctx.NSP)
    assert np.array_equal(ctx.download_desc(F), oF)
    assert np.array_equal(ctx.download_desc(Ff), oF)


@pytest.mark.parametrize("W,H,up,N", [(23, 9, 1.0, 512), (17, 5, 1.0, 100), (30, 7, 1.3, 256), (70, 3, 1.0, 32)])
def test_brief_frames_smaller_than_a_tile(W, H, up, N):
    """Frames narrower or shorter than one work-group tile (64 x 16 / 32 rows,
    up to 16 waves), with a user pattern whose offsets (up to 20 px) reach past
    every border: both BRIEF kernels clamp each sample to the frame as the
    oracle does."""
    say("brief tiny frames", W, H, up, N)
    pr = _pair(W, H, seed=W * H + N)
    p = bm.derive_params(W, H, up, up, N=N)
    ctx = bm.Context(p, seed=5)
    pat = np.random.default_rng(N + W).integers(-20, 21, 4 * N).astype(np.int16)
    ctx.set_pattern(pat)
    rgb = bm.DeviceBuffer(array=pr.frame1)
    g8 = bm.DeviceBuffer(W * H)
    ctx.gray(rgb, g8=g8)
    L = bm.DeviceBuffer(p.sx * p.sy * 4)
    ctx.upsample(g8, L)
    oL = oracle.upsample(oracle.gray(pr.frame1), p.sx, p.sy)
    oF = _pad(oracle.brief(oL, pat, N), ctx.NSP)
    for near in (8, 16, 24):
        ctx.set_tuning(brief_near=near)
        F, Ff = ctx.desc_buffer(), ctx.desc_buffer()
        ctx.brief(L, F)
        ctx.brief_gray(g8, Ff)
        ctx.sync()
        assert np.array_equal(ctx.download_desc(F), oF), ("frame", near)
        assert np.array_equal(ctx.download_desc(Ff), oF), ("fused", near)


@pytest.mark.parametrize("N", [72, 128])
def test_brief_pattern_halo_boundaries(N):
    """User patterns whose tests sit exactly at the near halo, one pixel beyond
    it, at zero and far out, in words that mix them, words of near tests only
    and words of far tests only (and a partial last word for N = 72): the
    fused BRIEF (one bit record per bit in bit order, far bits through their
    own records) equals the oracle for every near halo the tuner can pick,
    and so does the BRIEF kernel reading an up-sampled frame (its tile and far
    tests from the frame, the same records)."""
    say("brief pattern boundaries", N)
    W, H, up = 96, 64, 2.0
    pr = _pair(W, H, seed=N)
    p = bm.derive_params(W, H, up, up, N=N)
    ctx = bm.Context(p, seed=3)
    oL = oracle.upsample(oracle.gray(pr.frame1), p.sx, p.sy)
    rgb = bm.DeviceBuffer(array=pr.frame1)
    g8 = bm.DeviceBuffer(W * H)
    ctx.gray(rgb, g8=g8)
    L = bm.DeviceBuffer(p.sx * p.sy * 4)
    ctx.upsample(g8, L)
    rng = np.random.default_rng(N)
    for rn in (8, 16, 24, 32, 40, 48):
        pat = np.zeros((N, 4), np.int16)
        for b in range(N):
            kind = (b // 32) % 3 if b // 32 < 3 else rng.integers(0, 3)
            near = rng.integers(-rn, rn + 1, 4)
            near[rng.integers(0, 4)] = rng.choice([-rn, rn, 0])
            if kind == 1 or (kind == 0 and b % 2 == 0):
                pat[b] = near
            else:
                far = near.copy()
                far[rng.integers(0, 4)] = rng.choice([-(rn + 1), rn + 1, -60, 60])
                pat[b] = far
        ctx.set_tuning(brief_near=rn)
        ctx.set_pattern(pat.ravel())
        F = ctx.desc_buffer()
        ctx.brief_gray(g8, F)
        ctx.sync()
        oF = _pad(oracle.brief(oL, pat.ravel(), N), ctx.NSP)
        assert np.array_equal(ctx.download_desc(F), oF), (rn, ctx.tuning().brief_near)
        Ff = ctx.desc_buffer()
        ctx.brief(L, Ff)
        ctx.sync()
        assert np.array_equal(ctx.download_desc(Ff), oF), ("frame", rn)
        print("[brief halo]", N, rn, ctx.tuning().brief_near, flush=True)


def test_set_pattern_after_unfit_plan():
    """A context whose fused-BRIEF tiles do not fit (no up-sampling and a patch
    radius of 160 px, offsets up to about 70 px: the gray tile of one
    work-group exceeds its bound) starts
    on the stand-alone kernel; a pattern of small offsets set afterwards brings
    the fused kernel back with its full plan in the buffers sized at creation,
    and both give the oracle's descriptors."""
    say("set_pattern after unfit plan")
    W, H, N = 640, 480, 128
    pr = _pair(W, H, seed=5)
    p = bm.derive_params(W, H, 1.0, 1.0, patch_radius=20.0, N=N)
    ctx = bm.Context(p, seed=4)
    assert ctx.tuning().brief_near == 0
    pat, _ = _oracle_tables(ctx, p, 4)
    assert np.abs(pat).max() > 48
    oL = oracle.upsample(oracle.gray(pr.frame1), p.sx, p.sy)
    F, _, _ = _descs(ctx, pr.frame1)
    assert np.array_equal(ctx.download_desc(F), _pad(oracle.brief(oL, pat, N), ctx.NSP))
    small = np.random.default_rng(9).integers(-8, 9, 4 * N).astype(np.int16)
    ctx.set_pattern(small)
    assert ctx.tuning().brief_near > 0
    F, _, _ = _descs(ctx, pr.frame1)
    assert np.array_equal(ctx.download_desc(F), _pad(oracle.brief(oL, small, N), ctx.NSP))
    # the whole-path call in the frame mode, then the wide pattern again:
    # without a near plan the context falls back to the fused mode, and its
    # tuning stays settable
    def pair_descs():
        r1, r2 = bm.DeviceBuffer(array=pr.frame1), bm.DeviceBuffer(array=pr.frame2)
        U, V = ctx.image_buffer(), ctx.image_buffer()
        ctx.flow_pair(r1, r2, U, V)
        ctx.sync()
        return bm._download_raw(ctx.frames()[1][0], (ctx.sy, ctx.sx, ctx.NSP), np.uint32)
    ctx.set_tuning(brief_frame=1)
    assert np.array_equal(pair_descs(), _pad(oracle.brief(oL, small, N), ctx.NSP))
    ctx.set_pattern(pat)
    t = ctx.tuning()
    assert t.brief_near == 0 and t.brief_frame == 0
    ctx.set_tuning(**t.as_dict())
    assert np.array_equal(pair_descs(), _pad(oracle.brief(oL, pat, N), ctx.NSP))


@pytest.mark.parametrize("N,double_prop", [(32, True), (128, True), (128, False), (256, True), (512, True)])
def test_levels_random_search_init(N, double_prop):
    """Level / random-search / init parity from a random (non-trivial) snapshot."""
    say("levels", N, double_prop)
    W, H, up = 72, 40, 2.0
    pr = _pair(W, H, seed=3 + N)
    p = bm.derive_params(W, H, up, up, N=N, doubleProp=int(double_prop))
    ctx = bm.Context(p, seed=5)
    pat, _ = _oracle_tables(ctx, p, 5)
    F1, _, _ = _descs(ctx, pr.frame1)
    F2, _, _ = _descs(ctx, pr.frame2)
    oF1, _, _ = _oracle_descs(pr.frame1, p, pat)
    oF2, _, _ = _oracle_descs(pr.frame2, p, pat)
    rng = np.random.default_rng(N)
    sx, sy = p.sx, p.sy
    xx, yy = np.meshgrid(np.arange(sx), np.arange(sy))
    dx = np.clip(rng.integers(-25, 26, (sy, sx)), -xx, sx - 1 - xx).astype(np.int16)
    dy = np.clip(rng.integers(-25, 26, (sy, sx)), -yy, sy - 1 - yy).astype(np.int16)
    off0 = (dx.view(np.uint16).astype(np.uint32) | (dy.view(np.uint16).astype(np.uint32) << 16))
    # the stored cost of every entry is the Hamming cost of its offset, as in any
    # field the library produces (the precondition of bm_propagate_level)
    tx, ty = xx + dx, yy + dy
    x = np.bitwise_xor(oF1, oF2[ty, tx])
    cost0 = np.unpackbits(np.ascontiguousarray(x).view(np.uint8), axis=-1).sum(-1).astype(np.uint16)
    oi, ci = bm.DeviceBuffer(array=off0), bm.DeviceBuffer(array=cost0)
    oo, co = ctx.field_buffers()
    for p_lvl in (256, 31, 8, 1):
        for pd in (-1, 1):
            ctx.propagate_level(F1, F2, oi, ci, oo, co, p_lvl, pd)
            ctx.sync()
            eo, ec = oracle.propagate_level(oF1, oF2, p_lvl, pd, double_prop, off0, cost0)
            assert np.array_equal(oo.download((sy, sx), np.uint32), eo), (p_lvl, pd)
            assert np.array_equal(co.download((sy, sx), np.uint16), ec), (p_lvl, pd)
    # random search, in place
    ctx.random_search(F1, F2, oi, ci, iteration=3, pair=7)
    ctx.sync()
    eo, ec = oracle.random_search(oF1, oF2, 3, 7, 5, off0, cost0)
    assert np.array_equal(oi.download((sy, sx), np.uint32), eo)
    assert np.array_equal(ci.download((sy, sx), np.uint16), ec)
    # temporal init from the random field
    prev = bm.DeviceBuffer(array=off0)
    ctx.init(F1, F2, oo, co, first=False, prev_off=prev)
    ctx.sync()
    eo, ec = oracle.init(oF1, oF2, first=False, prev_off=off0)
    assert np.array_equal(oo.download((sy, sx), np.uint32), eo)
    assert np.array_equal(co.download((sy, sx), np.uint16), ec)


@pytest.mark.parametrize("W,H,up,N,iters,jf", [(320, 240, 1.5, 32, 2, 256), (100, 70, 2.9, 128, 3, 64),
                                               (64, 48, 4.9, 512, 2, 256), (90, 60, 2.0, 64, 1, 100)])
def test_match_field(W, H, up, N, iters, jf):
    say("match", W, H, up, N, iters, jf)
    pr = _pair(W, H, seed=N + iters)
    p = bm.derive_params(W, H, up, up, N=N, iterations=iters, jfMax=jf)
    ctx = bm.Context(p, seed=9)
    F1, _, _ = _descs(ctx, pr.frame1)
    F2, _, _ = _descs(ctx, pr.frame2)
    off, cost = ctx.field_buffers()
    ctx.match(F1, F2, off, cost, first=True, pair=4)
    ctx.sync()
    pat, _ = _oracle_tables(ctx, p, 9)
    oF1, _, _ = _oracle_descs(pr.frame1, p, pat)
    oF2, _, _ = _oracle_descs(pr.frame2, p, pat)
    eo, ec = oracle.match(oF1, oF2, iters, jf, True, True, None, 4, 9)
    assert np.array_equal(off.download((p.sy, p.sx), np.uint32), eo)
    assert np.array_equal(cost.download((p.sy, p.sx), np.uint16), ec)
    # propagate + separate random search equals the fused match
    off2, cost2 = ctx.field_buffers()
    ctx.init(F1, F2, off2, cost2, first=True)
    pd = -1
    for i in range(iters):
        ctx.propagate(F1, F2, off2, cost2, pd)
        ctx.random_search(F1, F2, off2, cost2, iteration=i, pair=4)
        pd = -pd
    ctx.sync()
    assert np.array_equal(off2.download((p.sy, p.sx), np.uint32), eo)
    assert np.array_equal(cost2.download((p.sy, p.sx), np.uint16), ec)


def test_refinement_stages():
    """Down-sampling, 1-D medians and the trilateral filter on the oracle's field."""
    say("refine stages")
    W, H, up = 160, 120, 2.9
    pr = _pair(W, H, seed=21, max_motion=10.0)
    p = bm.derive_params(W, H, up, up, N=64)
    ctx = bm.Context(p, seed=2)
    pat, kern = _oracle_tables(ctx, p, 2)
    oF1, og, _ = _oracle_descs(pr.frame1, p, pat)
    oF2, _, _ = _oracle_descs(pr.frame2, p, pat)
    eo, _ = oracle.match(oF1, oF2, 2, 256, True, True, None, 0, 2)
    U, V = ctx.image_buffer(), ctx.image_buffer()
    d_eo = bm.DeviceBuffer(array=eo)
    ctx.downsample(d_eo, U, V)
    ctx.sync()
    oU, oV = oracle.downsample(eo, W, H, kern)
    gU, gV = U.download((H, W), np.float32), V.download((H, W), np.float32)
    assert np.array_equal(gU, oU) and np.array_equal(gV, oV)
    # medians on identical inputs are exact order statistics
    for d in (0, 1):
        out, src = ctx.image_buffer(), bm.DeviceBuffer(array=oU)
        ctx.median1d(src, out, d)
        ctx.sync()
        assert np.array_equal(out.download((H, W), np.float32), oracle.median1d(oU, d, p.fsize))
    oUm = oracle.median1d(oracle.median1d(oU, 1, p.fsize), 0, p.fsize)
    oVm = oracle.median1d(oracle.median1d(oV, 1, p.fsize), 0, p.fsize)
    guide = og.astype(np.float32)
    Uo, Vo = ctx.image_buffer(), ctx.image_buffer()
    ins = [bm.DeviceBuffer(array=a) for a in (oU, oV, oUm, oVm, guide)]
    ctx.trilateral(*ins, Uo, Vo)
    ctx.sync()
    tU, tV = oracle.trilateral(oU, oV, oUm, oVm, guide, p.sEPE ** 2, p.sI ** 2, p.sS ** 2)
    assert np.abs(Uo.download((H, W), np.float32) - tU).max() <= FLOW_TOL
    assert np.abs(Vo.download((H, W), np.float32) - tV).max() <= FLOW_TOL
    # the composed refinement entry point
    ins = [bm.DeviceBuffer(array=a) for a in (oU, oV, guide)]
    ctx.refine(*ins, Uo, Vo)
    ctx.sync()
    assert np.abs(Uo.download((H, W), np.float32) - tU).max() <= FLOW_TOL


@pytest.mark.parametrize("W,H,up,sigma_rad", [(97, 61, 4.9, 0.025), (130, 70, 2.9, 0.4), (64, 48, 12.0, 0.025),
                                               (300, 9, 1.7, 0.025)])
def test_downsample_paths(W, H, up, sigma_rad):
    """Down-sampling on random fields: the LDS-tiled kernel (windows up to
    64 KB) and the direct kernel (larger windows, here x12 up-sampling) equal
    the oracle bit for bit (the same products and sums, each rounded), on
    ragged tiles and wide Gaussians."""
    say("downsample", W, H, up, sigma_rad)
    p = bm.derive_params(W, H, up, up, N=32, sigma_ds=sigma_rad)
    ctx = bm.Context(p, seed=1)
    rng = np.random.default_rng(W + H)
    dx = rng.integers(-300, 301, (p.sy, p.sx)).astype(np.int32)
    dy = rng.integers(-300, 301, (p.sy, p.sx)).astype(np.int32)
    off = ((dx & 0xffff) | ((dy & 0xffff) << 16)).astype(np.uint32)
    U, V = ctx.image_buffer(), ctx.image_buffer()
    ctx.downsample(bm.DeviceBuffer(array=off), U, V)
    ctx.sync()
    _, kern = _oracle_tables(ctx, p, 1)
    oU, oV = oracle.downsample(off, W, H, kern)
    gU, gV = U.download((H, W), np.float32), V.download((H, W), np.float32)
    print("[ds]", W, H, up, len(kern), np.abs(gU - oU).max(), np.abs(gV - oV).max(), flush=True)
    assert np.array_equal(gU, oU) and np.array_equal(gV, oV)


@pytest.mark.parametrize("W,H,sS", [(150, 100, 6.0), (67, 45, 12.0), (40, 30, 20.0), (129, 17, 3.0)])
def test_trilateral_kernels(W, H, sS):
    """Every trilateral kernel (VALU sums; MFMA sums with 2 or 4 row groups),
    whole or with each tile's source rows split over 2, 4 or 8 work-groups,
    against the oracle on ragged frames, windows up to wider than the frame."""
    say("trilateral kernels", W, H, sS)
    rng = np.random.default_rng(W * 1000 + H)
    p = bm.derive_params(W, H, 1.0, 1.0, N=32)
    p.sS = sS
    ctx = bm.Context(p, seed=1)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    # piecewise-smooth flow with a motion boundary, noise, and a textured guide
    U = (np.where(xx > W / 2, 6.0, -3.0) + 0.05 * yy + rng.normal(0, 0.7, (H, W))).astype(np.float32)
    V = (np.where(yy > H / 3, -4.0, 2.5) + rng.normal(0, 0.7, (H, W))).astype(np.float32)
    Um = (U + rng.normal(0, 0.3, (H, W))).astype(np.float32)
    Vm = (V + rng.normal(0, 0.3, (H, W))).astype(np.float32)
    guide = rng.integers(0, 256, (H, W)).astype(np.float32)
    tU, tV = oracle.trilateral(U, V, Um, Vm, guide, p.sEPE ** 2, p.sI ** 2, p.sS ** 2)
    ins = [bm.DeviceBuffer(array=a) for a in (U, V, Um, Vm, guide)]
    got = {}
    for groups in (0, 2, 4):
        for split in (1, 2, 4, 8):
            ctx.set_tuning(tri_groups=groups, tri_split=split)
            Uo, Vo = ctx.image_buffer(), ctx.image_buffer()
            ctx.trilateral(*ins, Uo, Vo)
            ctx.sync()
            gU, gV = Uo.download((H, W), np.float32), Vo.download((H, W), np.float32)
            eu, ev = np.abs(gU - tU).max(), np.abs(gV - tV).max()
            print("[tri]", W, H, sS, groups, split, eu, ev, flush=True)
            assert eu <= FLOW_TOL and ev <= FLOW_TOL, (groups, split, eu, ev)
            got[groups, split] = (gU, gV)
    for key, (gU, gV) in got.items():
        # every variant sums the same terms in another order
        assert np.abs(gU - got[0, 1][0]).max() <= FLOW_TOL and np.abs(gV - got[0, 1][1]).max() <= FLOW_TOL, key


@pytest.mark.parametrize("W,H,sS", [(160, 120, None), (97, 61, 6.0)])
def test_trilateral_large_flow(W, H, sS):
    """Flows of 35-45 px with medians 2 px off: a fifth of the pixels have a
    weight sum below the 1e-5 floor, where the normalisation no longer cancels
    an error common to a pixel's weights.  Differences taken after scaling by
    the range factor (rounding relative to the flow, not the difference) were
    off by 1.5e-3 px here; every kernel stays within 1e-3 px of the oracle."""
    say("trilateral large flow", W, H, sS)
    rng = np.random.default_rng(77)
    p = bm.derive_params(W, H, 1.0, 1.0, N=32)
    if sS is not None:
        p.sS = sS
    ctx = bm.Context(p, seed=1)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    U = (np.where(xx > W / 3, 38.0, -41.0) + 0.02 * yy + rng.normal(0, 1.5, (H, W))).astype(np.float32)
    V = (np.where(yy > H / 2, -35.0, 44.0) + rng.normal(0, 1.5, (H, W))).astype(np.float32)
    Um = (U + rng.normal(0, 2.0, (H, W))).astype(np.float32)
    Vm = (V + rng.normal(0, 2.0, (H, W))).astype(np.float32)
    guide = rng.integers(0, 256, (H, W)).astype(np.float32)
    tU, tV = oracle.trilateral(U, V, Um, Vm, guide, p.sEPE ** 2, p.sI ** 2, p.sS ** 2)
    ins = [bm.DeviceBuffer(array=a) for a in (U, V, Um, Vm, guide)]
    for groups in (0, 2, 4):
        for split in (1, 4):
            ctx.set_tuning(tri_groups=groups, tri_split=split)
            Uo, Vo = ctx.image_buffer(), ctx.image_buffer()
            ctx.trilateral(*ins, Uo, Vo)
            ctx.sync()
            eu = np.abs(Uo.download((H, W), np.float32) - tU).max()
            ev = np.abs(Vo.download((H, W), np.float32) - tV).max()
            print("[tri large]", W, H, sS, groups, split, eu, ev, flush=True)
            assert eu <= FLOW_TOL and ev <= FLOW_TOL, (groups, split, eu, ev)


@pytest.mark.parametrize("filt", [1, 0])
def test_flow_pair_c1(filt):
    """The 320x240 / u1.5 / 32-bit workload, whole path, against the oracle."""
    say("flow_pair c1 filter", filt)
    pr = synth.workload_pair("c1_320x240_u1.5_n32", seed=1)
    p = bm.derive_params(320, 240, 1.5, 1.5, N=32, filter=filt)
    ctx = bm.Context(p, seed=7)
    U, V = ctx.flow_numpy(pr.frame1, pr.frame2, pair=3)
    off, cost = ctx.field()
    eo, ec, oU, oV = oracle.flow_pair(_oracle_params(p), pr.frame1, pr.frame2, seed=7, pair=3)
    assert np.array_equal(off, eo)
    assert np.array_equal(cost, ec)
    assert np.abs(U - oU).max() <= FLOW_TOL and np.abs(V - oV).max() <= FLOW_TOL
    e_all, _ = bm.epe(U, V, pr.gt, pr.occ)
    z_all, _ = bm.epe(np.zeros_like(U), np.zeros_like(V), pr.gt, pr.occ)
    # the refined flow is far better than zero flow; the raw decimated field is noisier
    assert e_all < (0.5 if filt else 0.95) * z_all


@pytest.mark.parametrize("W,H,up,N", [(120, 90, 2.0, 64), (80, 60, 2.5, 512)])
def test_sequence_temporal_init(W, H, up, N):
    """Sequence mode: three frames, temporal init from the previous field (at
    512 bits the initialisation and level kernels run two lanes per pixel)."""
    say("stream 3 frames", W, H, up, N)
    base = _pair(W, H, seed=31, max_motion=5.0)
    f3 = _pair(W, H, seed=32, max_motion=5.0).frame2
    frames = [base.frame1, base.frame2, f3]
    p = bm.derive_params(W, H, up, up, N=N)
    ctx = bm.Context(p, seed=13)
    U, V = ctx.image_buffer(), ctx.image_buffer()
    outs = []
    keep = []
    for k, fr in enumerate(frames):
        keep.append(bm.DeviceBuffer(array=fr))
        has = ctx.stream_frame(keep[-1], U, V)
        assert has == (k > 0)
        if has:
            ctx.sync()
            outs.append((ctx.field(), U.download((H, W), np.float32), V.download((H, W), np.float32)))
    pat, kern = _oracle_tables(ctx, p, 13)
    oF = [_oracle_descs(f, p, pat) for f in frames]
    prev = None
    for k in range(2):
        first = prev is None
        eo, ec = oracle.match(oF[k][0], oF[k + 1][0], 2, 256, True, first, prev, k, 13)
        (go, gc), gU, gV = outs[k]
        assert np.array_equal(go, eo), k
        assert np.array_equal(gc, ec), k
        oU, oV = oracle.downsample(eo, W, H, kern)
        oUm = oracle.median1d(oracle.median1d(oU, 1, p.fsize), 0, p.fsize)
        oVm = oracle.median1d(oracle.median1d(oV, 1, p.fsize), 0, p.fsize)
        tU, tV = oracle.trilateral(oU, oV, oUm, oVm, oF[k][1].astype(np.float32), p.sEPE ** 2, p.sI ** 2,
                                   p.sS ** 2)
        assert np.abs(gU - tU).max() <= FLOW_TOL and np.abs(gV - tV).max() <= FLOW_TOL
        prev = eo


@pytest.mark.parametrize("W,H,up,N", [(160, 120, 2.0, 64), (97, 61, 2.9, 128)])
def test_sequence_pipelined(W, H, up, N):
    """Pipelined sequence mode (refinement tail on the second stream) gives the
    serial mode's flow bit for bit, frame by frame, with one output pair per
    frame and a single synchronisation at the end; a whole-pair call after it
    sees a settled context."""
    say("stream pipelined", W, H, up, N)
    frames = [_pair(W, H, seed=40 + k, max_motion=5.0).frame1 for k in range(5)]
    p = bm.derive_params(W, H, up, up, N=N)
    keep = [bm.DeviceBuffer(array=f) for f in frames]
    res = {}
    for piped in (False, True):
        ctx = bm.Context(p, seed=5)
        if piped:
            ctx.set_pipelining(True)
        outs = [(ctx.image_buffer(), ctx.image_buffer()) for _ in frames]
        for k in range(len(frames)):
            ctx.stream_frame(keep[k], *outs[k])
        ctx.sync()
        res[piped] = [(U.download((H, W), np.float32), V.download((H, W), np.float32)) for U, V in outs[1:]]
        # a whole-pair call right after pipelined frames (no sync in between)
        for k in range(len(frames)):
            ctx.stream_frame(keep[k], *outs[k])
        U, V = ctx.image_buffer(), ctx.image_buffer()
        ctx.flow_pair(keep[0], keep[1], U, V)
        ctx.sync()
        res[(piped, "pair")] = (U.download((H, W), np.float32), V.download((H, W), np.float32))
        ctx.close()
    for a, b in zip(res[False], res[True]):
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    assert np.array_equal(res[(False, "pair")][0], res[(True, "pair")][0])
    assert np.array_equal(res[(False, "pair")][1], res[(True, "pair")][1])


def test_sequence_pipelined_interleaved():
    """Pipelining with the refinement off (serial path), switched off in the
    middle of a sequence, and a learned-path call right after pipelined frames:
    each result equals the one of a context that never pipelined."""
    say("stream pipelined, interleaved calls")
    W, H, up, N = 128, 96, 2.0, 64
    frames = [_pair(W, H, seed=60 + k, max_motion=4.0).frame1 for k in range(4)]
    keep = [bm.DeviceBuffer(array=f) for f in frames]
    net = bm.Net(weights=np.zeros(bm.NET_PARAMS, np.float32))
    out = {}
    for piped in (False, True):
        for filt in (1, 0):
            p = bm.derive_params(W, H, up, up, N=N, filter=filt)
            ctx = bm.Context(p, seed=9)
            ctx.set_pipelining(piped)
            bufs = [(ctx.image_buffer(), ctx.image_buffer()) for _ in frames]
            for k in range(3):
                ctx.stream_frame(keep[k], *bufs[k])
            if piped:
                ctx.set_pipelining(False)  # settles the pending tail
            ctx.stream_frame(keep[3], *bufs[3])
            if filt:
                ctx.set_pipelining(piped)
                ctx.stream_frame(keep[0], *bufs[0])  # pending tail again
                Ul, Vl = ctx.image_buffer(), ctx.image_buffer()
                ctx.flow_pair_learned(net, keep[1], keep[2], Ul, Vl)
            ctx.sync()
            # the sequence's first frame has no flow: its buffers are written only
            # when frame 0 comes round again (filt = 1)
            got = [(U.download((H, W), np.float32), V.download((H, W), np.float32))
                   for U, V in (bufs if filt else bufs[1:])]
            if filt:
                got.append((Ul.download((H, W), np.float32), Vl.download((H, W), np.float32)))
            out[(piped, filt)] = got
            ctx.close()
    for filt in (1, 0):
        for a, b in zip(out[(False, filt)], out[(True, filt)]):
            assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1]), filt


def _field_invariants(off, cost, F1, F2, NS, row0=0, row_step=1):
    """Size-independent checks of a final field on the rows row0::row_step:
    in-image targets, and the stored cost is the Hamming distance between the
    pixel's descriptor and the descriptor at the stored offset."""
    sy, sx = off.shape
    rows = np.arange(row0, sy, row_step)
    o = off[rows]
    dx = (o & 0xFFFF).astype(np.uint16).view(np.int16).astype(np.int64)
    dy = (o >> 16).astype(np.uint16).view(np.int16).astype(np.int64)
    xx = np.arange(sx)[None, :]
    tx, ty = xx + dx, rows[:, None] + dy
    assert (tx >= 0).all() and (tx < sx).all() and (ty >= 0).all() and (ty < sy).all()
    x = np.bitwise_xor(F1[rows][..., :NS], F2[ty, tx][..., :NS])
    pc = np.unpackbits(np.ascontiguousarray(x).view(np.uint8), axis=-1).sum(-1)
    assert np.array_equal(pc, cost[rows].astype(np.int64))


@pytest.mark.slow
@pytest.mark.parametrize("name,seed", [("c2_1920x1080_u2.9_n128", 0), ("c3_1280x720_u4.9_n512", 0)])
def test_full_size_match_field(name, seed):
    """The 1080p/128-bit and 720p/512-bit workloads: the product's BRIEF pattern
    (patch radius 191.65 / 215.88 matched pixels) and Gaussian (11 / 13 taps)
    equal the oracle's, the match field is bit-exact against the oracle run on
    its own tables, the field invariants hold, and the final flow is within
    1e-3 px of the oracle's on every pixel of the frame."""
    say("full size", name)
    wl = synth.WORKLOADS[name]
    pr = synth.workload_pair(name, seed=seed)
    W, H = wl["W"], wl["H"]
    p = bm.derive_params(W, H, wl["up"], wl["up"], N=wl["N"])
    ctx = bm.Context(p, seed=1)
    p = ctx.params
    U, V = ctx.flow_numpy(pr.frame1, pr.frame2, pair=0)
    off, cost = ctx.field()
    pat, kern = _oracle_tables(ctx, p, 1)
    assert len(kern) == {"c2_1920x1080_u2.9_n128": 11, "c3_1280x720_u4.9_n512": 13}[name]
    oF1, og, _ = _oracle_descs(pr.frame1, p, pat)
    oF2, _, _ = _oracle_descs(pr.frame2, p, pat)
    eo, ec = oracle.match(oF1, oF2, 2, 256, True, True, None, 0, 1)
    assert np.array_equal(off, eo)
    assert np.array_equal(cost, ec)
    _field_invariants(off, cost, oF1, oF2, (p.N + 31) // 32, row0=3, row_step=61)
    del oF1, oF2
    oU, oV = oracle.downsample(eo, W, H, kern)
    oUm = oracle.median1d(oracle.median1d(oU, 1, p.fsize), 0, p.fsize)
    oVm = oracle.median1d(oracle.median1d(oV, 1, p.fsize), 0, p.fsize)
    tU, tV = oracle.trilateral(oU, oV, oUm, oVm, og.astype(np.float32), p.sEPE ** 2, p.sI ** 2, p.sS ** 2)
    du, dv = float(np.abs(U - tU).max()), float(np.abs(V - tV).max())
    print("[full size flow]", name, "rows", H, "max |du| %.2e |dv| %.2e" % (du, dv), flush=True)
    assert du <= FLOW_TOL and dv <= FLOW_TOL
    e_all, e_noc = bm.epe(U, V, pr.gt, pr.occ)
    print("[epe]", name, e_all, e_noc, flush=True)


@pytest.mark.parametrize("W,H,up", [(320, 240, 1.5), (61, 37, 2.3), (96, 64, 1.0), (40, 30, 4.9)])
def test_upsample_lanczos_bit_exact(W, H, up):
    """Lanczos-3 mode (the reference's up-sampler) against the oracle."""
    say("lanczos", W, H, up)
    p = bm.derive_params(W, H, up, up, N=64)
    ctx = bm.Context(p, seed=1)
    g = np.random.default_rng(W).integers(0, 256, (H, W)).astype(np.uint8)
    g8 = bm.DeviceBuffer(array=g)
    L = bm.DeviceBuffer(ctx.sx * ctx.sy * 4)
    ctx.upsample_lanczos(g8, L)
    ctx.sync()
    assert np.array_equal(L.download((ctx.sy, ctx.sx), np.float32), oracle.upsample_lanczos(g, ctx.sx, ctx.sy))


def test_flow_pair_lanczos_c1():
    """The whole pair with the Lanczos-3 up-sampler: match field bit-exact,
    flow within 1e-3 px of the oracle run with the same up-sampler."""
    say("flow_pair c1 lanczos")
    pr = synth.workload_pair("c1_320x240_u1.5_n32", seed=1)
    p = bm.derive_params(320, 240, 1.5, 1.5, N=32)
    ctx = bm.Context(p, seed=7)
    ctx.set_upsampler(bm.UPSAMPLE_LANCZOS3)
    assert ctx.upsampler() == bm.UPSAMPLE_LANCZOS3
    U, V = ctx.flow_numpy(pr.frame1, pr.frame2, pair=3)
    off, cost = ctx.field()
    eo, ec, oU, oV = oracle.flow_pair(_oracle_params(p), pr.frame1, pr.frame2, seed=7, pair=3, lanczos=True)
    assert np.array_equal(off, eo) and np.array_equal(cost, ec)
    assert np.abs(U - oU).max() <= FLOW_TOL and np.abs(V - oV).max() <= FLOW_TOL
    e_l, _ = bm.epe(U, V, pr.gt, pr.occ)
    bctx = bm.Context(p, seed=7)
    bU, bV = bctx.flow_numpy(pr.frame1, pr.frame2, pair=3)
    e_b, _ = bm.epe(bU, bV, pr.gt, pr.occ)
    print("[epe] c1 lanczos %.4f bilinear %.4f" % (e_l, e_b), flush=True)
    assert e_l < 1.5 * e_b and e_b < 1.5 * e_l
