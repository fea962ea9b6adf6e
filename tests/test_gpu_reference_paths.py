"""The reference's own configuration and fixture on the GPU, against the oracle.

* The default anisotropic up-sampling 2.740 x 3.052 (of_pipeline.cpp:178-181),
  which every other test replaces by equal factors: a swapped scale factor or
  an sx-for-sy mistake in the up-sampling, descriptor, NN decimation or
  down-sampling tables would only show here.
* data/RubberWhale frames 10 -> 11 (README.md:93-100) through bm_flow_pair and
  through the CLI with no -ux/-uy/-fl, both against the oracle run on the same
  frames.
* Parameter paths of the match loop (briefmatch.cpp:248-293): iterations = 0
  (the range is [0, 100], of_pipeline.cpp:119), doubleProp = false through the
  whole pair, and sequences with tInit = false (briefmatch.h:74,
  briefmatch.cpp:265-268).
Match fields are bit-exact, flows within 1e-3 px of the oracle."""
import os
import subprocess

import numpy as np
import pytest

import briefmatch_amd as bm
import oracle
from briefmatch_amd import io as bio
from briefmatch_amd import synth
from conftest import say

pytestmark = pytest.mark.gpu

FLOW_TOL = 1e-3
GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rubberwhale")


def _oracle_params(p):
    op = oracle.Params()
    for n, _ in op._fields_:
        setattr(op, n, getattr(p, n))
    return op


def _check_pair(ctx, frame1, frame2, seed, pair, what):
    """bm_flow_pair against oracle.flow_pair, which builds its own pattern and
    Gaussian from the parameters and the seed."""
    U, V = ctx.flow_numpy(frame1, frame2, pair=pair)
    off, cost = ctx.field()
    eo, ec, oU, oV = oracle.flow_pair(_oracle_params(ctx.params), frame1, frame2, seed=seed, pair=pair)
    assert np.array_equal(off, eo), what
    assert np.array_equal(cost, ec), what
    du, dv = float(np.abs(U - oU).max()), float(np.abs(V - oV).max())
    print("[ref path]", what, "max |du| %.2e |dv| %.2e" % (du, dv), flush=True)
    assert du <= FLOW_TOL and dv <= FLOW_TOL, what
    return U, V


@pytest.mark.parametrize("W,H,N", [(200, 150, 128), (157, 93, 32), (96, 130, 72), (64, 48, 512)])
def test_default_anisotropic_upsampling(W, H, N):
    """derive_params with no factors: 2.740 x 3.052, so sx / W != sy / H."""
    say("default up-sampling", W, H, N)
    p = bm.derive_params(W, H, N=N)
    assert p.sx == int(np.float32(2.740) * np.float32(W)) and p.sy == int(np.float32(3.052) * np.float32(H))
    assert p.sx * H != p.sy * W
    pr = synth.make_pair(W, H, motion="rects", n_rects=3, max_motion=6.0, seed=W + N)
    ctx = bm.Context(p, seed=17)
    _check_pair(ctx, pr.frame1, pr.frame2, 17, 2, "aniso %dx%d N%d" % (W, H, N))


def test_default_anisotropic_sequence():
    """Sequence mode at the default factors: temporal initialisation follows
    the previous field through offsets whose x and y scales differ."""
    say("default up-sampling, sequence")
    W, H = 120, 90
    p = bm.derive_params(W, H, N=64)
    frames = [synth.make_pair(W, H, motion="rects", max_motion=4.0, seed=70 + k).frame1 for k in range(3)]
    ctx = bm.Context(p, seed=23)
    U, V = ctx.image_buffer(), ctx.image_buffer()
    keep = [bm.DeviceBuffer(array=f) for f in frames]
    fields = []
    for k in range(3):
        if ctx.stream_frame(keep[k], U, V):
            ctx.sync()
            fields.append(ctx.field())
    pat = oracle.make_pattern(23, p.N, p.patchArea)
    assert np.array_equal(ctx.pattern(), pat)
    oF = [oracle.brief(oracle.upsample(oracle.gray(f), p.sx, p.sy), pat, p.N) for f in frames]
    prev = None
    for k in range(2):
        eo, ec = oracle.match(oF[k], oF[k + 1], p.iterations, p.jfMax, True, prev is None, prev, k, 23)
        assert np.array_equal(fields[k][0], eo) and np.array_equal(fields[k][1], ec), k
        prev = eo


def _rubberwhale():
    f10 = bio.read_image(os.path.join(GOLD, "frame10.png"))
    f11 = bio.read_image(os.path.join(GOLD, "frame11.png"))
    gU, gV = bio.read_flo(os.path.join(GOLD, "flow10.flo"))
    return f10, f11, np.stack([gU, gV], -1)


def test_rubberwhale_reference_defaults():
    """The reference's fixture at its default configuration (584 x 388 ->
    1600 x 1184 matched pixels, 128-bit descriptors): bit-exact field and flow
    within 1e-3 px of the oracle; EPE against flow10.flo reported."""
    say("RubberWhale 10->11 defaults")
    f10, f11, gt = _rubberwhale()
    p = bm.derive_params(584, 388)
    assert (p.sx, p.sy, p.N) == (1600, 1184, 128)
    ctx = bm.Context(p, seed=0)
    U, V = _check_pair(ctx, f10, f11, 0, 0, "RubberWhale defaults")
    e, _ = bm.epe(U, V, gt)
    z, _ = bm.epe(np.zeros_like(U), np.zeros_like(V), gt)
    print("[epe] RubberWhale 10->11 at the reference defaults: %.4f (zero flow %.4f)" % (e, z), flush=True)
    assert e < 0.5 * z


@pytest.mark.skipif(not os.path.exists(bio.CLI_PATH), reason="CLI not built")
def test_cli_reference_defaults(tmp_path):
    """The CLI with only -i, -f and -o (every other option at the reference's
    default) equals bm_flow_pair at the default parameters, and so the oracle."""
    say("cli defaults")
    args = [bio.CLI_PATH, "-i", os.path.join(GOLD, "frame%02d.png"), "-f", "10:11",
            "-o", str(tmp_path / "flow%02d.flo")]
    r = subprocess.run(args, capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    U, V = bio.read_flo(str(tmp_path / "flow11.flo"))
    f10, f11, _ = _rubberwhale()
    ctx = bm.Context(bm.derive_params(584, 388), seed=0)
    pU, pV = _check_pair(ctx, f10, f11, 0, 0, "RubberWhale via bm_flow_pair")
    assert np.array_equal(U, pU) and np.array_equal(V, pV)


@pytest.mark.parametrize("iters", [0, 1, 3])
def test_iteration_counts(iters):
    """iterations = 0 leaves the initialisation (no propagation and no random
    search, briefmatch.cpp:274); 1 and 3 run the fused first iteration and an
    odd number of pd flips."""
    say("iterations", iters)
    W, H = 160, 120
    p = bm.derive_params(W, H, 2.0, 2.0, N=64, iterations=iters)
    pr = synth.make_pair(W, H, motion="rects", max_motion=5.0, seed=90 + iters)
    ctx = bm.Context(p, seed=29)
    _check_pair(ctx, pr.frame1, pr.frame2, 29, 1, "iterations %d" % iters)
    if iters == 0:
        off, _ = ctx.field()
        assert not off.any()  # zero initialisation only


@pytest.mark.parametrize("W,H,up,N", [(160, 120, 2.0, 64), (120, 80, 0.0, 128)])
def test_double_prop_off_whole_pair(W, H, up, N):
    """doubleProp = false: two propagation candidates per level through the
    whole pair, also at the default anisotropic factors."""
    say("doubleProp off", W, H, up, N)
    p = bm.derive_params(W, H, up, up, N=N, doubleProp=0)
    pr = synth.make_pair(W, H, motion="rects", max_motion=5.0, seed=W)
    ctx = bm.Context(p, seed=31)
    _check_pair(ctx, pr.frame1, pr.frame2, 31, 0, "doubleProp off %dx%d" % (W, H))


@pytest.mark.parametrize("t_init", [0, 1])
def test_sequence_temporal_init_switch(t_init):
    """Four frames in sequence mode: with tInit = false every pair starts from
    the zero field (pair index = frame counter), with tInit = true from the
    previous field; flows against the oracle's refinement of its own field."""
    say("sequence tInit", t_init)
    W, H = 112, 84
    p = bm.derive_params(W, H, 2.0, 2.0, N=64, tInit=t_init)
    frames = [synth.make_pair(W, H, motion="rects", max_motion=4.0, seed=80 + k).frame1 for k in range(4)]
    ctx = bm.Context(p, seed=37)
    p = ctx.params
    keep = [bm.DeviceBuffer(array=f) for f in frames]
    U, V = ctx.image_buffer(), ctx.image_buffer()
    outs = []
    for k in range(4):
        if ctx.stream_frame(keep[k], U, V):
            ctx.sync()
            outs.append((ctx.field(), U.download((H, W), np.float32), V.download((H, W), np.float32)))
    pat = oracle.make_pattern(37, p.N, p.patchArea)
    kern = oracle.gauss_kernel(p.sigma)
    assert np.array_equal(ctx.pattern(), pat) and np.array_equal(ctx.gauss_kernel(), kern)
    gray = [oracle.gray(f) for f in frames]
    oF = [oracle.brief(oracle.upsample(g, p.sx, p.sy), pat, p.N) for g in gray]
    prev = None
    for k in range(3):
        first = prev is None or not t_init
        eo, ec = oracle.match(oF[k], oF[k + 1], p.iterations, p.jfMax, True, first, None if first else prev, k,
                              37)
        (go, gc), gU, gV = outs[k]
        assert np.array_equal(go, eo) and np.array_equal(gc, ec), k
        oU, oV = oracle.downsample(eo, W, H, kern)
        oUm = oracle.median1d(oracle.median1d(oU, 1, p.fsize), 0, p.fsize)
        oVm = oracle.median1d(oracle.median1d(oV, 1, p.fsize), 0, p.fsize)
        tU, tV = oracle.trilateral(oU, oV, oUm, oVm, gray[k].astype(np.float32), p.sEPE ** 2, p.sI ** 2,
                                   p.sS ** 2)
        assert np.abs(gU - tU).max() <= FLOW_TOL and np.abs(gV - tV).max() <= FLOW_TOL, k
        prev = eo
