"""Learned proposal on the GPU, through the C ABI.

Integer parts (feature map, far block match, labels, offsets, seeded
initialisation, the match field of the learned path) agree with the oracle bit for bit.  The network's
forward output and gradients are compared with the oracle's float64 network
evaluated with the same BF16 rounding points: the device accumulates in FP32
in another order, so they agree to a relative 2e-2 per layer (gradients) and
2e-2 of the output range (forward).  Training lowers the loss and is
bit-reproducible.  The local search around the seeds matches the oracle bit
for bit.  With the committed weights, the learned path with a local search and
4 level launches beats the plain method's two-iteration EPE at 1080p, and the
trained network beats the all-zero network."""
import os

import numpy as np
import pytest

import briefmatch_amd as bm
import oracle
from briefmatch_amd import learned, synth
from conftest import say
from oracle import net as onet

pytestmark = pytest.mark.gpu

FLOW_TOL = 1e-3


def _flow_ctx(W, H, up, N, seed=3, **kw):
    p = bm.derive_params(W, H, up, up, N=N, **kw)
    ctx = bm.Context(p, seed=seed)
    pr = synth.make_pair(W, H, motion="rects", max_motion=min(W, H) / 10.0, seed=seed)
    r1 = bm.DeviceBuffer(array=pr.frame1)
    r2 = bm.DeviceBuffer(array=pr.frame2)
    U, V = ctx.image_buffer(), ctx.image_buffer()
    ctx.flow_pair(r1, r2, U, V)
    ctx.sync()
    return ctx, pr, (r1, r2, U, V)


def _frames(ctx):
    g8, F = ctx.frames()
    Fs = [bm._download_raw(f, (ctx.sy, ctx.sx, ctx.NSP), np.uint32) for f in F]
    gs = [bm._download_raw(g, (ctx.H, ctx.W), np.uint8) for g in g8]
    return g8, F, Fs, gs


@pytest.mark.parametrize("W,H,up,N", [(320, 240, 1.5, 32), (200, 136, 2.9, 128), (130, 75, 1.0, 512),
                                      (97, 43, 2.2, 64)])
def test_features_and_labels_bit_exact(W, H, up, N):
    say("features", W, H, up, N)
    ctx, pr, (r1, r2, U, V) = _flow_ctx(W, H, up, N)
    cw, ch = ctx.proposal_dims()
    assert (cw, ch) == (W // 8, H // 8)
    g8, F, Fs, gs = _frames(ctx)
    feat = bm.DeviceBuffer(ch * cw * 128 * 2)
    lab = bm.DeviceBuffer(ch * cw * 2 * 4)
    ctx.proposal_features(g8[0], g8[1], feat)
    ctx.coarse_flow(U, V, lab)
    ctx.sync()
    want = onet.features(gs[0], gs[1])
    assert np.array_equal(feat.download((ch, cw, 128), np.uint16), want)
    Uh, Vh = U.download((H, W), np.float32), V.download((H, W), np.float32)
    assert np.array_equal(lab.download((ch, cw, 2), np.float32), onet.coarse_flow(Uh, Vh))


@pytest.mark.parametrize("W,H,up,N", [(320, 240, 1.5, 32), (200, 136, 2.9, 128), (97, 43, 2.2, 512),
                                      (120, 64, 2.0, 256)])
def test_offsets_and_seeded_init_bit_exact(W, H, up, N):
    say("seeded init", W, H, up, N)
    ctx, pr, _ = _flow_ctx(W, H, up, N, seed=5)
    cw, ch = ctx.proposal_dims()
    g8, F, Fs, gs = _frames(ctx)
    rng = np.random.default_rng(W + N)
    out = (rng.standard_normal((ch, cw, 16)) * 1.5).astype(np.float32)
    out[0, 0, 0] = np.nan
    out[-1, -1, 1] = 1e30
    dout = bm.DeviceBuffer(array=out)
    feat = bm.DeviceBuffer(ch * cw * 256)
    ctx.proposal_features(g8[0], g8[1], feat)
    coff = bm.DeviceBuffer(ch * cw * 4 * 4)
    ctx.proposal_offsets(dout, feat, coff)
    ctx.sync()
    got_coff = coff.download((ch, cw, 4), np.int32)
    want_feat = onet.features(gs[0], gs[1])
    want_coff = onet.proposal_offsets(out, want_feat, W, H, ctx.sx, ctx.sy)
    assert np.array_equal(got_coff, want_coff)
    # with a far match: random displacements of every ring, so that both sides
    # of the override are taken
    far = rng.integers(-5, 6, (H // 16, W // 16, 2)).astype(np.int32)
    coff_far = bm.DeviceBuffer(ch * cw * 4 * 4)
    ctx.proposal_offsets(dout, feat, coff_far, far=bm.DeviceBuffer(array=far))
    ctx.sync()
    want_far = onet.proposal_offsets(out, want_feat, W, H, ctx.sx, ctx.sy, far=far)
    assert np.array_equal(coff_far.download((ch, cw, 4), np.int32), want_far)
    assert not np.array_equal(want_far, want_coff)
    M = ctx.sx * ctx.sy
    dx = rng.integers(-9, 10, (ctx.sy, ctx.sx))
    dy = rng.integers(-9, 10, (ctx.sy, ctx.sx))
    prev = ((dx & 0xFFFF) | ((dy & 0xFFFF) << 16)).astype(np.uint32)
    dprev = bm.DeviceBuffer(array=prev)
    off = bm.DeviceBuffer(M * 4)
    cost = bm.DeviceBuffer(M * 2)
    for first in (True, False):
        ctx.init_proposal(F[0], F[1], coff, off, cost, first=first, prev_off=None if first else dprev)
        ctx.sync()
        eo, ec = onet.init_proposal(Fs[0], Fs[1], want_coff, W, H, first=first, prev_off=prev)
        assert np.array_equal(off.download((ctx.sy, ctx.sx), np.uint32), eo)
        assert np.array_equal(cost.download((ctx.sy, ctx.sx), np.uint16), ec)


def _gray_pair(rng, H, W, shift):
    g1 = rng.integers(0, 256, (H, W)).astype(np.uint8)
    g2 = np.roll(g1, shift, axis=(0, 1))
    g2[rng.random((H, W)) < 0.05] = 0          # some noise, so that costs are not all zero
    return g1, g2


@pytest.mark.parametrize("W,H,shift", [(320, 240, (0, 64)), (272, 256, (-48, 16)), (97, 43, (3, -17)),
                                       (1920, 1080, (37, -70)), (16, 16, (0, 0)), (40, 33, (5, 9))])
def test_far_search_bit_exact(W, H, shift):
    """bm_proposal_far = the oracle's far block match, bit for bit, on moved
    random frames (odd grids, the smallest 16x16 frame, a 1080p frame)."""
    say("far search", W, H, shift)
    p = bm.derive_params(W, H, 1.5, 1.5, N=32)
    ctx = bm.Context(p, seed=1)
    rng = np.random.default_rng(W * H)
    g1, g2 = _gray_pair(rng, H, W, shift)
    far = bm.DeviceBuffer(array=np.full((H // 16, W // 16, 2), -99, np.int32))
    ctx.proposal_far(bm.DeviceBuffer(array=g1), bm.DeviceBuffer(array=g2), far)
    ctx.sync()
    want = onet.far_search(g1, g2)
    assert np.array_equal(far.download((H // 16, W // 16, 2), np.int32), want)


def test_far_search_errors():
    """Frames below 16x16 cannot hold a block: a status, no launch."""
    p = bm.derive_params(40, 15, 1.5, 1.5, N=32)
    ctx = bm.Context(p, seed=1)
    g = bm.DeviceBuffer(array=np.zeros((15, 40), np.uint8))
    far = bm.DeviceBuffer(64)
    with pytest.raises(bm.BriefMatchError):
        ctx.proposal_far(g, g, far)
    cw, ch = ctx.proposal_dims()
    feat, nout, coff = bm.DeviceBuffer(ch * cw * 256), bm.DeviceBuffer(ch * cw * 64), bm.DeviceBuffer(ch * cw * 16)
    with pytest.raises(bm.BriefMatchError):
        ctx.proposal_offsets(nout, feat, coff, far=far)


def _random_weights(seed):
    rng = np.random.default_rng(seed)
    w = learned.init_weights(seed)
    w[-16:-14] = rng.standard_normal(2).astype(np.float32)
    # non-zero biases everywhere so that the bias gradients are exercised
    for (W_, b), (ci, co, _, _) in zip(onet.split_params(w), onet.LAYERS):
        b[:] = rng.standard_normal(co).astype(np.float32) * 0.05
    return w


def _random_feat(rng, n, ch, cw):
    f = rng.uniform(-1, 1, (n, ch, cw, 128)).astype(np.float32)
    f[..., 123:] = 0.0
    return onet.bf16_bits(f)


# below 16384 pixels the waves split the output channels; 32400 (the 1080p
# map): one pixel tile per wave; 65536 and 262144: two and four
@pytest.mark.parametrize("n,ch,cw", [(1, 9, 13), (3, 15, 20), (1, 64, 70), (1, 135, 240), (1, 256, 256),
                                     (4, 256, 256)])
def test_forward_matches_oracle(n, ch, cw):
    say("net forward", n, ch, cw)
    rng = np.random.default_rng(ch)
    w = _random_weights(ch)
    feat = _random_feat(rng, n, ch, cw)
    net = bm.Net(weights=w)
    assert np.array_equal(net.weights(), w)
    dfeat = bm.DeviceBuffer(array=feat)
    dout = bm.DeviceBuffer(n * ch * cw * 16 * 4)
    net.forward(dfeat, n, ch, cw, dout)
    got = dout.download((n, ch, cw, 16), np.float32)
    want, _ = onet.forward(w, feat, quant=True)
    scale = np.abs(want[..., :2]).max()
    err = np.abs(got[..., :2] - want[..., :2])
    print("[net] forward max err %.3e of range %.3e, mean %.3e" % (err.max(), scale, err.mean()), flush=True)
    assert err.max() <= 2e-2 * scale
    assert err.mean() <= 2e-3 * scale
    assert not got[..., 2:].any() or np.abs(got[..., 2:] - want[..., 2:]).max() <= 2e-2 * scale


@pytest.mark.parametrize("n,ch,cw", [(1, 9, 13), (2, 30, 40), (1, 128, 160), (2, 160, 220)])
def test_gradients_match_oracle(n, ch, cw):
    say("net gradients", n, ch, cw)
    rng = np.random.default_rng(100 + ch)
    w = _random_weights(7 + ch)
    feat = _random_feat(rng, n, ch, cw)
    label = (rng.standard_normal((n, ch, cw, 2)) * 3.0).astype(np.float32)
    label[0, 0, 0, 0] = np.nan
    net = bm.Net(weights=w)
    loss, g = net.gradients(bm.DeviceBuffer(array=feat), bm.DeviceBuffer(array=label), n, ch, cw)
    wl, wg = onet.loss_and_grad(w, feat, label, quant=True)
    print("[net] loss gpu %.6f oracle %.6f" % (loss, wl), flush=True)
    assert abs(loss - wl) <= 1e-4 * abs(wl)
    for l, ((gW, gb), (eW, eb)) in enumerate(zip(onet.split_params(g), onet.split_params(wg))):
        for name, a, b in (("W", gW, eW), ("b", gb, eb)):
            if l == len(onet.LAYERS) - 1:
                assert not a[2:].any() if name == "b" else not a[:, 2:].any()
            rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
            print("[net] layer %d %s rel err %.2e" % (l, name, rel), flush=True)
            assert rel <= 2e-2, (l, name, rel)


def test_training_lowers_loss_and_is_reproducible():
    """150 Adam steps on one mini-batch of four 160x120 pairs (the network can
    fit it: the float64 oracle network goes from 1.75 to 0.70 in the same
    schedule) lower its loss below 70%; two runs give identical weights."""
    say("training")
    ds = learned.build_dataset(4, seed0=77, W=160, H=120, up=1.0, N=64, workers=1, log=None)
    w0 = learned.init_weights(3)
    nets = [bm.Net(weights=w0), bm.Net(weights=w0)]
    before, _ = nets[0].gradients(ds.feat, ds.label, ds.rows, ds.ch, ds.cw)
    for net in nets:
        for _ in range(150):
            net.train_step(ds.feat, ds.label, ds.rows, [0, 1, 2, 3], ds.ch, ds.cw, 1e-3)
    after, _ = nets[0].gradients(ds.feat, ds.label, ds.rows, ds.ch, ds.cw)
    print("[train] loss %.4f -> %.4f" % (before, after), flush=True)
    assert after < 0.7 * before
    assert np.array_equal(nets[0].weights(), nets[1].weights())


@pytest.mark.parametrize("W,H,up,N,radius,samples", [(320, 240, 1.5, 32, 16, 5), (200, 136, 2.9, 128, 24, 6),
                                                      (130, 75, 1.0, 512, 3, 4), (97, 43, 2.2, 64, 200, 16)])
def test_local_search_bit_exact(W, H, up, N, radius, samples):
    """bm_local_search on a converged field of a real pair (and on its zero
    field, where most samples are taken) against the oracle; the last case's
    radius exceeds the frame, so the clamp acts on most samples."""
    say("local search", W, H, up, N, radius, samples)
    ctx, pr, _ = _flow_ctx(W, H, up, N)
    ctx.set_local_search(radius, samples)
    assert ctx.local_search_config() == (radius, samples)
    g8, F, Fs, gs = _frames(ctx)
    oF1, oF2 = Fs[0][..., :ctx.NS].copy(), Fs[1][..., :ctx.NS].copy()
    off0, cost0 = ctx.field()
    zo, zc = oracle.init(oF1, oF2, first=True)
    for k, (o_in, c_in) in enumerate(((off0, cost0), (zo, zc))):
        off = bm.DeviceBuffer(array=np.ascontiguousarray(o_in))
        cost = bm.DeviceBuffer(array=np.ascontiguousarray(c_in))
        ctx.local_search(F[0], F[1], off, cost, iteration=2 + k, pair=9)
        ctx.sync()
        eo, ec = onet.local_search(oF1, oF2, o_in, c_in, radius, samples, 2 + k, 9, 3)
        go = off.download((ctx.sy, ctx.sx), np.uint32)
        gc = cost.download((ctx.sy, ctx.sx), np.uint16)
        assert np.array_equal(go, eo) and np.array_equal(gc, ec)
        if k == 1:
            assert (go != zo).mean() > 0.02     # the search moved pixels off zero
    ctx.close()


def test_local_search_errors():
    p = bm.derive_params(64, 48, 1.0, 1.0, N=32)
    ctx = bm.Context(p, seed=1)
    assert ctx.local_search_config() == (0, 0)
    for r, n in ((0, 3), (5, -1), (5, 17), (4097, 2)):
        with pytest.raises(bm.BriefMatchError):
            ctx.set_local_search(r, n)
    ctx.set_local_search(0, 0)
    ctx.set_local_search(4096, 16)
    assert ctx.local_search_config() == (4096, 16)
    ctx.close()


@pytest.mark.parametrize("motion,W,H,ux,uy,ls", [("c1", 320, 240, 1.5, 1.5, None), ("far", 320, 240, 1.5, 1.5, None),
                                                ("small", 56, 14, 1.5, 1.5, None), ("c1", 200, 150, 0.0, 0.0, None),
                                                ("c1", 320, 240, 1.5, 1.5, (12, 5)),
                                                ("far", 320, 240, 0.0, 0.0, (30, 6))])
def test_learned_path_equals_composed_oracle(motion, W, H, ux, uy, ls):
    """bm_flow_pair_learned = features + far match -> network -> offsets ->
    seeded init -> random search -> propagation + random search ->
    refinement, each stage on the oracle except the network (its device output
    is the oracle's input); "far": regions moving up to 60 pixels, where the
    far match overrides the network in many cells; "small": 56 x 14 frames,
    one coarse row and no 16-pixel block row, so there is no far match; the
    fourth case runs the reference's default anisotropic up-sampling (2.740 x
    3.052, of_pipeline.cpp:178-181), where the coarse-to-matched offset scales
    differ per axis; the last two add the local search around the seeds
    (radius, samples) after the seeded initialisation's random search."""
    say("learned path", motion, W, H, ux, uy, ls)
    N = 32
    p = bm.derive_params(W, H, ux, uy, N=N)
    if (ux, uy) == (0.0, 0.0):
        assert p.sx * H != p.sy * W  # different scale factors per axis
    p.iterations = 1
    ctx = bm.Context(p, seed=11)
    if ls:
        ctx.set_local_search(*ls)
    path = learned.DEFAULT_WEIGHTS
    net = bm.Net(path=path) if os.path.exists(path) else bm.Net(weights=_random_weights(1))
    if motion == "c1":
        pr = synth.workload_pair("c1_320x240_u1.5_n32", seed=2)
    elif motion == "far":
        pr = synth.make_pair(W, H, motion="regions", n_regions=6, max_motion=60.0, seed=1)
    else:
        pr = synth.make_pair(W, H, motion="rects", n_rects=1, max_motion=3.0, seed=4)
    r1 = bm.DeviceBuffer(array=pr.frame1)
    r2 = bm.DeviceBuffer(array=pr.frame2)
    U, V = ctx.image_buffer(), ctx.image_buffer()
    ctx.flow_pair_learned(net, r1, r2, U, V, pair=6)
    ctx.sync()
    off, cost = ctx.field()
    g8, F, Fs, gs = _frames(ctx)
    cw, ch = ctx.proposal_dims()
    feat = bm.DeviceBuffer(ch * cw * 256)
    ctx.proposal_features(g8[0], g8[1], feat)
    ctx.sync()
    nout = bm.DeviceBuffer(ch * cw * 64)
    net.forward(feat, 1, ch, cw, nout)
    out = nout.download((ch, cw, 16), np.float32)
    far = onet.far_search(gs[0], gs[1])
    if motion == "far":
        assert (np.abs(far).max(axis=-1) >= onet.FAR_MIN).sum() > 50
    coff = onet.proposal_offsets(out, onet.features(gs[0], gs[1]), W, H, ctx.sx, ctx.sy, far=far)
    oF1, oF2 = Fs[0][..., :ctx.NS].copy(), Fs[1][..., :ctx.NS].copy()
    eo, ec = onet.init_proposal(oF1, oF2, coff, W, H)
    eo, ec = oracle.random_search(oF1, oF2, 1, 6, 11, eo, ec)     # after the seeded init: index = iterations
    if ls:
        eo, ec = onet.local_search(oF1, oF2, eo, ec, ls[0], ls[1], 1, 6, 11)
    eo, ec = oracle.propagate(oF1, oF2, -1, p.jfMax, True, eo, ec)
    eo, ec = oracle.random_search(oF1, oF2, 0, 6, 11, eo, ec)
    assert np.array_equal(off, eo) and np.array_equal(cost, ec)
    kern = oracle.gauss_kernel(ctx.params.sigma)
    assert np.array_equal(ctx.gauss_kernel(), kern)
    oU, oV = oracle.downsample(eo, W, H, kern)
    oUm = oracle.median1d(oracle.median1d(oU, 1, p.fsize), 0, p.fsize)
    oVm = oracle.median1d(oracle.median1d(oV, 1, p.fsize), 0, p.fsize)
    tU, tV = oracle.trilateral(oU, oV, oUm, oVm, gs[0].astype(np.float32), p.sEPE ** 2, p.sI ** 2, p.sS ** 2)
    assert np.abs(U.download((H, W), np.float32) - tU).max() <= FLOW_TOL
    assert np.abs(V.download((H, W), np.float32) - tV).max() <= FLOW_TOL


@pytest.mark.skipif(not os.path.exists(learned.DEFAULT_WEIGHTS), reason="no trained proposal weights")
def test_trained_proposal_gain():
    """What the learned path gains, and what the trained network itself gains
    over the same path with an all-zero network (its proposal is then the
    cost-volume argmin and far match alone), in the configuration the
    benchmark reports: the seeded initialisation, a local search of 6 samples
    around the seeds (radius 16 input pixels), one propagation iteration over
    the jump range 8 (4 level launches; the plain method executes 9).
    Measured with the committed weights, mean EPE: 1920x1080 / u2.9 / 128-bit
    pairs (seeds 0-7) 0.970 against 1.224 (plain, two iterations) and 1.173
    (zero network); 1280x720 / u4.9 / 512-bit pairs moving up to 60 px (seeds
    0-3) 4.58 against 30.4 and 6.11; 320x240 / u1.5 / 32-bit pairs (seeds 0-3)
    0.569 against 0.361 and 0.876 (motion below 8 px: the plain method's zero
    start is already close, and the short jump range costs more than the seeds
    give).  Asserted: the target on the 1080p pairs (within 2 % of the plain
    method; here below it), the large-motion gain, and the network's own gain
    over the zero network on all three workloads."""
    net = bm.Net(path=learned.DEFAULT_WEIGHTS)
    log = lambda m: print(m, flush=True)
    cfg = dict(fewer=1, jf_max=8, local_search=(16.0, 6), log=log)
    say("learned gain c2")
    c2 = learned.evaluate(net, seeds=tuple(range(8)), **cfg)
    say("learned gain c3")
    c3 = learned.evaluate(net, workload="c3_1280x720_u4.9_n512", seeds=(0, 1, 2, 3), **cfg)
    say("learned gain c1")
    c1 = learned.evaluate(net, workload="c1_320x240_u1.5_n32", seeds=(0, 1, 2, 3), **cfg)
    print("[gain] c2", c2, "\n[gain] c3", c3, "\n[gain] c1", c1, flush=True)
    assert c2["learned_fewer"] <= 1.02 * c2["plain"]
    assert c2["learned_fewer"] < 0.1 * c2["plain_fewer"]
    assert c2["learned_fewer"] <= 0.92 * c2["argmin_fewer"]
    assert c3["learned_fewer"] <= 0.2 * c3["plain"]
    assert c3["learned_fewer"] <= 0.85 * c3["argmin_fewer"]
    assert c1["learned_fewer"] <= 0.85 * c1["argmin_fewer"]
    net.close()


@pytest.mark.skipif(not os.path.exists(learned.DEFAULT_WEIGHTS), reason="no trained proposal weights")
def test_network_gain_without_propagation():
    """What the trained network itself gains over the all-zero network (the
    method without it: the seeded initialisation with the cost-volume argmin
    and far match alone) when no propagation iteration follows the seeded
    initialisation, so the flow is set by the seeds and one random search.
    Measured with the committed weights: mean EPE 2.74 against 7.33 on the
    1920x1080 / u2.9 / 128-bit pairs (seeds 0-7), 12.0 against 18.0 on the
    1280x720 pairs moving up to 60 px (seeds 0-3) and 1.17 against 2.46 on the
    320x240 pairs (seeds 0-3).  With one iteration the propagation spreads
    either set of seeds over the regions and the two are level
    (test_trained_proposal_gain)."""
    net = bm.Net(path=learned.DEFAULT_WEIGHTS)
    log = lambda m: print(m, flush=True)
    bounds = (("c2_1920x1080_u2.9_n128", tuple(range(8)), 0.6), ("c3_1280x720_u4.9_n512", (0, 1, 2, 3), 0.8),
              ("c1_320x240_u1.5_n32", (0, 1, 2, 3), 0.8))
    for workload, seeds, bound in bounds:
        say("network gain without propagation", workload)
        r = learned.evaluate(net, workload=workload, seeds=seeds, fewer=0, log=log)
        print("[gain0]", workload, r, flush=True)
        assert r["learned_fewer"] <= bound * r["argmin_fewer"], workload
    net.close()


def _staged_learned_epe(ctx, net, p, pr, with_far):
    """The learned path through its public stages (features, far match,
    network, offsets, seeded init, random search, one propagation iteration,
    refinement), with or without the far match; returns the EPE."""
    W, H = ctx.W, ctx.H
    r1, r2 = bm.DeviceBuffer(array=pr.frame1), bm.DeviceBuffer(array=pr.frame2)
    U, V = ctx.image_buffer(), ctx.image_buffer()
    ctx.flow_pair(r1, r2, U, V)                    # gray frames and descriptors of the pair
    g8, F = ctx.frames()
    cw, ch = ctx.proposal_dims()
    feat = bm.DeviceBuffer(ch * cw * 256)
    ctx.proposal_features(g8[0], g8[1], feat)
    far = None
    if with_far:
        far = bm.DeviceBuffer((H // 16) * (W // 16) * 8)
        ctx.proposal_far(g8[0], g8[1], far)
    ctx.sync()
    nout = bm.DeviceBuffer(ch * cw * 64)
    net.forward(feat, 1, ch, cw, nout)
    coff = bm.DeviceBuffer(ch * cw * 16)
    ctx.proposal_offsets(nout, feat, coff, far=far)
    off, cost = ctx.field_buffers()
    ctx.init_proposal(F[0], F[1], coff, off, cost, first=True)
    ctx.random_search(F[0], F[1], off, cost, iteration=p.iterations, pair=0)
    ctx.propagate(F[0], F[1], off, cost, -1)
    ctx.random_search(F[0], F[1], off, cost, iteration=0, pair=0)
    U0, V0 = ctx.image_buffer(), ctx.image_buffer()
    ctx.downsample(off, U0, V0)
    g32 = ctx.image_buffer()
    ctx.gray(r1, None, g32)
    ctx.refine(U0, V0, g32, U, V)
    ctx.sync()
    return bm.epe(U.download((H, W), np.float32), V.download((H, W), np.float32), pr.gt)[0]


@pytest.mark.skipif(not os.path.exists(learned.DEFAULT_WEIGHTS), reason="no trained proposal weights")
def test_far_match_gain():
    """With one iteration, the far match lowers the learned path's EPE by at
    least 10 % on 1280x720 pairs moving up to 60 px (beyond the cost volume's
    40 px; measured 11 % with the committed weights on these two seeds) and
    costs at most 3 % on the 1920x1080 pairs (measured 0.8 %); the composed
    stages with the far match are bm_flow_pair_learned itself."""
    net = bm.Net(path=learned.DEFAULT_WEIGHTS)
    for name, seeds in (("c3_1280x720_u4.9_n512", (0, 1)), ("c2_1920x1080_u2.9_n128", (0, 1))):
        wl = synth.WORKLOADS[name]
        p = bm.derive_params(wl["W"], wl["H"], wl["up"], wl["up"], N=wl["N"])
        p.iterations = 1
        ctx = bm.Context(p, seed=0)
        e = {True: [], False: []}
        for seed in seeds:
            say("far match gain", name, seed)
            pr = synth.workload_pair(name, seed=seed)
            for with_far in (True, False):
                e[with_far].append(_staged_learned_epe(ctx, net, p, pr, with_far))
            U, V = ctx.flow_numpy_learned(net, pr.frame1, pr.frame2)
            assert bm.epe(U, V, pr.gt)[0] == e[True][-1]
        far, nofar = float(np.mean(e[True])), float(np.mean(e[False]))
        print("[far gain] %s with %.3f without %.3f" % (name, far, nofar), flush=True)
        if name.startswith("c3"):
            assert far <= 0.9 * nofar
        else:
            assert far <= 1.03 * nofar


def test_net_api_errors_and_roundtrip(tmp_path):
    """Invalid calls fail with a status before any launch; weights survive a
    save / load round trip bit for bit."""
    w = _random_weights(4)
    net = bm.Net(weights=w)
    path = str(tmp_path / "w.bnet")
    net.save(path)
    assert np.array_equal(bm.Net(path=path).weights(), w)
    raw = open(path, "rb").read()
    (tmp_path / "short.bnet").write_bytes(raw[:-4])
    (tmp_path / "magic.bnet").write_bytes(b"XXXXXXXX" + raw[8:])
    for bad in ("short.bnet", "magic.bnet", "missing.bnet"):
        with pytest.raises(bm.BriefMatchError):
            net.load(str(tmp_path / bad))
    with pytest.raises(bm.BriefMatchError):
        net.set_weights(w[:-1])
    bad_w = w.copy()
    bad_w[7] = np.nan
    with pytest.raises(bm.BriefMatchError):
        net.set_weights(bad_w)
    assert np.array_equal(net.weights(), w)                 # rejected calls leave the weights alone
    feat = bm.DeviceBuffer(2 * 4 * 5 * 256)
    lab = bm.DeviceBuffer(2 * 4 * 5 * 8)
    feat.zero()
    lab.zero()
    for idx in ([0, 2], [-1], list(range(257))):            # row outside the set / batch over 256
        with pytest.raises(bm.BriefMatchError):
            net.train_step(feat, lab, 2, idx, 4, 5, 1e-3)
    with pytest.raises(bm.BriefMatchError):
        net.train_step(feat, lab, 2, [0, 1], 4, 5, float("nan"))
    assert np.array_equal(net.weights(), w)
    tiny = bm.Context(bm.derive_params(7, 7, 2.0, 2.0, N=32))
    rgb = bm.DeviceBuffer(array=np.zeros((7, 7, 3), np.uint8))
    U, V = tiny.image_buffer(), tiny.image_buffer()
    with pytest.raises(bm.BriefMatchError):
        tiny.flow_pair_learned(net, rgb, rgb, U, V)
