"""Learned PatchMatch proposal: dataset, training and evaluation drivers.

The network and its kernels are in csrc/bm_net.hip; this module only drives
them through the C ABI (briefmatch_amd.Net / Context):

* ``build_dataset`` generates seeded synthetic pairs on the host cores, runs
  the plain method on each (its output flow, with four propagation
  iterations and no up-sampling, is the label: the random-search radius in
  input pixels then reaches as far, in 8-pixel cells, as at 1920x1080 with
  up-sampling 2.9) and stores, per
  pair, the proposal feature map and the 8x8 block means of that flow (in
  coarse pixels) in device memory;
* ``train`` runs Adam on random mini-batches of that dataset;
* ``evaluate`` compares, on pairs with known flow, the plain method with its
  default iterations, the plain method with fewer iterations and the
  proposal-seeded method with fewer iterations.

    python -m briefmatch_amd.learned train --pairs 20000 --out proposal.bnet
    python -m briefmatch_amd.learned eval --weights proposal.bnet
"""
import argparse
import ctypes as C
import math
import multiprocessing as mp
import os
import sys
import time

import numpy as np

from . import Context, DeviceBuffer, Net, derive_params, epe, NET_PARAMS, H2D, _download_raw, _check, lib
from . import synth

_HERE = os.path.dirname(os.path.abspath(__file__))
# trained on 20,000 640x480 pairs, half moving rectangles and half large-motion
# regions, labelled by the method started from the block-matching proposal (8
# propagation iterations):
#   python -m briefmatch_amd.learned train --pairs 20000 --epochs 100 --lr 1e-3 --labels seeded \
#       --label-iterations 8 --family mixed --regions-share 0.5 --out proposal_640x480.bnet
DEFAULT_WEIGHTS = os.path.join(_HERE, "weights", "proposal_640x480.bnet")

# (cin, cout) per layer, as in the C ABI's parameter layout
_LAYERS = ((128, 64), (64, 64), (64, 64), (64, 64), (64, 16))


def init_weights(seed=0):
    """He-uniform weights for the ReLU layers, a small uniform last layer with
    only the two flow channels non-zero, zero biases."""
    rng = np.random.default_rng(seed)
    parts = []
    for l, (ci, co) in enumerate(_LAYERS):
        bound = math.sqrt(6.0 / (9 * ci))
        w = rng.uniform(-bound, bound, size=(9, co, ci)).astype(np.float32)
        if l == len(_LAYERS) - 1:
            w *= 0.1
            w[:, 2:, :] = 0.0
        parts += [w.ravel(), np.zeros(co, np.float32)]
    w = np.concatenate(parts)
    assert w.size == NET_PARAMS
    return w


def _offset(buf, nbytes):
    return C.c_void_p(buf.ptr.value + int(nbytes))


def _upload_raw(ptr, array):
    a = np.ascontiguousarray(array)
    _check(lib().bm_memcpy(ptr, a.ctypes.data_as(C.c_void_p), a.nbytes, H2D))


FAMILIES = ("rects", "mixed")


def train_pair(seed, W=640, H=480, with_gt=False, family="rects", regions_share=0.5):
    """Training pair `seed`.  family "rects": moving rectangles over an affine
    background (the family of the 1920x1080 benchmark pair), motion up to
    16..40 px, the frames a window of a scene 1 to 3 times their size (so
    texture and rectangle scales span those of 640x480 and 1920x1080
    renderings).  family "mixed": a share `regions_share` of the seeds
    (drawn per seed) smoothly blended affine regions (the family of the
    1280x720 benchmark pair) with 3..8 regions and motion up to 20..60 px,
    the others as "rects".
    Returns (frame1, frame2) or, with_gt, (frame1, frame2, gt flow (H, W, 2))."""
    if family not in FAMILIES:
        raise ValueError("family must be one of %s" % (FAMILIES,))
    rng = np.random.default_rng([seed, 7])
    if family == "mixed" and np.random.default_rng([seed, 11]).random() < regions_share:
        n_regions = int(rng.integers(3, 9))
        max_motion = float(rng.uniform(20.0, 60.0))
        p = synth.make_pair(W, H, motion="regions", n_regions=n_regions, max_motion=max_motion, seed=seed)
    else:
        n_rects = int(rng.integers(3, 11))
        max_motion = float(rng.uniform(16.0, 40.0))
        k = float(rng.uniform(1.0, 3.0))
        p = synth.make_pair(W, H, motion="rects", n_rects=n_rects, max_motion=max_motion, seed=seed,
                            scene=(int(W * k), int(H * k)))
    return (p.frame1, p.frame2, p.gt) if with_gt else (p.frame1, p.frame2)


def _gen(args):
    seed, W, H, with_gt, family, share = args
    return train_pair(seed, W, H, with_gt, family, share)


class Dataset:
    """Device-resident feature maps [rows][ch][cw][128] BF16 and labels
    [rows][ch][cw][2] FP32."""

    def __init__(self, rows, ch, cw):
        self.rows, self.ch, self.cw = rows, ch, cw
        self.feat_row = ch * cw * 128 * 2
        self.label_row = ch * cw * 2 * 4
        self.feat = DeviceBuffer(rows * self.feat_row)
        self.label = DeviceBuffer(rows * self.label_row)

    def labels(self):
        return self.label.download((self.rows, self.ch, self.cw, 2), np.float32)


LABELS = ("plain", "seeded", "gt")


def _complement(label, feat_row, tau):
    """Complement target: in cells whose cost-volume argmin (feature channels
    123, 124, in eighths of a cell) is within tau cells of the label in both
    axes, the argmin itself, elsewhere the label.  The network, whose output
    is that argmin plus a residual, then learns where the argmin is wrong and
    repeats it elsewhere, so its candidates add options only there."""
    am = (feat_row[..., 123:125].astype(np.uint32) << 16).view(np.float32) * 8.0
    ok = np.isfinite(label).all(-1) & (np.abs(am - label) <= tau).all(-1)
    out = label.copy()
    out[ok] = am[ok]
    return out, float(ok.mean())


def build_dataset(pairs, seed0=1000, W=640, H=480, up=1.0, N=128, iterations=4, workers=16, labels="plain",
                  family="rects", regions_share=0.5, complement=None, log=print):
    """Generate `pairs` seeded pairs (seeds seed0..), label them and store
    features and labels on the device.  labels="plain": the plain method's
    flow (`iterations` propagation iterations); "seeded": the same method
    started from the block-matching proposal alone (the learned path with an
    all-zero network: cost-volume argmin and far match, then `iterations`
    iterations); "gt": the pairs' known flow.  Each is reduced to 8x8 block
    means in coarse pixels; with complement = tau the targets become
    _complement's."""
    if labels not in LABELS:
        raise ValueError("labels must be one of %s" % (LABELS,))
    with_gt = labels == "gt"
    params = derive_params(W, H, up, up, N=N)
    params.iterations = iterations
    ctx = Context(params, seed=0)
    zero = Net() if labels == "seeded" else None
    kept = []
    cw, ch = ctx.proposal_dims()
    ds = Dataset(pairs, ch, cw)
    r1, r2 = ctx.rgb_buffer(), ctx.rgb_buffer()
    U, V = ctx.image_buffer(), ctx.image_buffer()
    g8, F = ctx.frames()
    t0 = time.time()
    jobs = [(seed0 + i, W, H, with_gt, family, regions_share) for i in range(pairs)]
    pool = mp.get_context("spawn").Pool(workers) if workers > 1 else None
    it = pool.imap(_gen, jobs, chunksize=4) if pool else map(_gen, jobs)
    try:
        for i, job in enumerate(it):
            ctx.sync()
            r1.upload(job[0])
            r2.upload(job[1])
            if with_gt:
                ctx.gray(r1, g8[0])
                ctx.gray(r2, g8[1])
                U.upload(np.ascontiguousarray(job[2][..., 0]))
                V.upload(np.ascontiguousarray(job[2][..., 1]))
            elif zero is not None:
                ctx.flow_pair_learned(zero, r1, r2, U, V, pair=0)
            else:
                ctx.flow_pair(r1, r2, U, V, pair=0)
            ctx.coarse_flow(U, V, _offset(ds.label, i * ds.label_row))
            ctx.proposal_features(g8[0], g8[1], _offset(ds.feat, i * ds.feat_row))
            if complement is not None:
                ctx.sync()
                lab = _download_raw(_offset(ds.label, i * ds.label_row), (ch, cw, 2), np.float32)
                feat = _download_raw(_offset(ds.feat, i * ds.feat_row), (ch, cw, 128), np.uint16)
                lab, frac = _complement(lab, feat, complement)
                kept.append(frac)
                _upload_raw(_offset(ds.label, i * ds.label_row), lab)
            if log and ((i + 1) % 1000 == 0 or i + 1 == pairs):
                extra = " (argmin kept in %.1f %% of the cells)" % (100 * np.mean(kept)) if kept else ""
                log("[data] %d/%d pairs, %.1f s%s" % (i + 1, pairs, time.time() - t0, extra))
        ctx.sync()
    finally:
        if pool:
            pool.close()
            pool.join()
    ctx.close()
    if zero is not None:
        zero.close()
    return ds


def train(net, ds, rows, epochs=20, batch=32, lr=2e-3, lr_min=5e-5, seed=0, log=print, log_every=200):
    """Adam over random mini-batches of rows `rows` of the dataset with a
    cosine learning-rate decay; returns the logged (step, loss) pairs."""
    rng = np.random.default_rng(seed)
    rows = np.asarray(rows, np.int32)
    steps_per_epoch = len(rows) // batch
    total = epochs * steps_per_epoch
    hist = []
    step = 0
    t0 = time.time()
    for ep in range(epochs):
        perm = rows[rng.permutation(len(rows))]
        for b in range(steps_per_epoch):
            a = lr_min + 0.5 * (lr - lr_min) * (1.0 + math.cos(math.pi * step / max(total, 1)))
            want = (step % log_every == 0) or step == total - 1
            loss = net.train_step(ds.feat, ds.label, ds.rows, perm[b * batch:(b + 1) * batch], ds.ch, ds.cw, a,
                                  want_loss=want)
            if want:
                hist.append((step, loss))
                if log:
                    log("[train] epoch %d step %d/%d lr %.2e loss %.4f (%.1f s)"
                        % (ep, step, total, a, loss, time.time() - t0))
            step += 1
    net.sync()
    return hist


def validation_loss(net, ds, rows, batch=32):
    """Mean loss over `rows` (no update): the gradient pass with lr 0 would
    change the Adam state, so the forward output is compared on the host."""
    rows = list(rows)
    out = DeviceBuffer(batch * ds.ch * ds.cw * 16 * 4)
    labels = ds.labels()
    tot, n = 0.0, 0
    for b in range(0, len(rows), batch):
        idx = rows[b:b + batch]
        for j, r in enumerate(idx):
            net.forward(_offset(ds.feat, r * ds.feat_row), 1, ds.ch, ds.cw, _offset(out, j * ds.ch * ds.cw * 64))
        o = out.download((len(idx), ds.ch, ds.cw, 16), np.float32)[..., :2]
        lab = labels[idx]
        tot += float(np.abs(o - lab).sum())
        n += lab.size
    return tot / max(n, 1)


def local_search_radius(params, radius_px):
    """A local-search radius in input pixels as matched pixels: times the
    larger up-sampling factor, rounded half up, at least 1 (as the command
    line's --local-search)."""
    return max(1, int(math.floor(radius_px * max(params.sx / params.sx_in, params.sy / params.sy_in) + 0.5)))


def evaluate(net, workload="c2_1920x1080_u2.9_n128", seeds=(0, 1, 2, 3), fewer=1, jf_max=None, local_search=None,
             log=print):
    """EPE (all pixels) of plain(default iterations), plain(fewer),
    learned(fewer), learned(default iterations) and, as the method without
    the trained network, the seeded path with an all-zero network (its
    proposal is then the cost-volume argmin alone) with fewer iterations, on
    seeded pairs of a benchmark workload.  fewer = 0 runs the seeded
    initialisation and its random search only: the flow is then the seeds'
    alone.  jf_max (default: the workload's) is the learned and zero-network
    runs' jump range; local_search = (radius in input pixels, samples) turns
    on the local search around their seeds."""
    wl = synth.WORKLOADS[workload]
    p = derive_params(wl["W"], wl["H"], wl["up"], wl["up"], N=wl["N"])
    default_it = p.iterations
    pf = p.copy()
    pf.iterations = fewer
    ps = p.copy()
    if jf_max is not None:
        pf.jfMax = jf_max
        ps.jfMax = jf_max
    pp = p.copy()
    pp.iterations = fewer
    c_plain = Context(p, seed=0)
    c_plain_few = Context(pp, seed=0)
    c_few = Context(pf, seed=0)
    c_learn = Context(ps, seed=0)
    if local_search:
        r = local_search_radius(pf, local_search[0])
        c_few.set_local_search(r, local_search[1])
        c_learn.set_local_search(r, local_search[1])
    zero = Net()
    res = {"plain": [], "plain_fewer": [], "learned_fewer": [], "learned": [], "argmin_fewer": []}
    for s in seeds:
        pair = synth.workload_pair(workload, seed=s)
        for key, ctx, learned in (("plain", c_plain, None), ("plain_fewer", c_plain_few, None),
                                  ("learned_fewer", c_few, net), ("learned", c_learn, net),
                                  ("argmin_fewer", c_few, zero)):
            if learned is not None:
                U, V = ctx.flow_numpy_learned(learned, pair.frame1, pair.frame2)
            else:
                U, V = ctx.flow_numpy(pair.frame1, pair.frame2)
            res[key].append(epe(U, V, pair.gt)[0])
        if log:
            log("[eval] seed %d: plain(%d) %.4f  plain(%d) %.4f  learned(%d) %.4f  learned(%d) %.4f  "
                "argmin only(%d) %.4f"
                % (s, default_it, res["plain"][-1], fewer, res["plain_fewer"][-1], fewer, res["learned_fewer"][-1],
                   default_it, res["learned"][-1], fewer, res["argmin_fewer"][-1]))
    for c in (c_plain, c_plain_few, c_few, c_learn):
        c.close()
    zero.close()
    out = {k: float(np.mean(v)) for k, v in res.items()}
    out.update(iterations=default_it, fewer=fewer, seeds=list(seeds), workload=workload, jf_max=pf.jfMax,
               local_search=list(local_search) if local_search else None)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    sub = ap.add_subparsers(dest="cmd", required=True)
    t = sub.add_parser("train")
    t.add_argument("--pairs", type=int, default=20000)
    t.add_argument("--val", type=int, default=500, help="held-out pairs for the validation loss")
    t.add_argument("--epochs", type=int, default=20)
    t.add_argument("--batch", type=int, default=32)
    t.add_argument("--lr", type=float, default=2e-3)
    t.add_argument("--workers", type=int, default=16)
    t.add_argument("--seed", type=int, default=0)
    t.add_argument("--labels", choices=LABELS, default="plain",
                   help="label flow: the plain method's output, the same started from the block-matching "
                        "proposal, or the pairs' known flow")
    t.add_argument("--complement", type=float, default=None, metavar="TAU",
                   help="targets: the argmin where it is within TAU cells of the label, else the label")
    t.add_argument("--family", choices=FAMILIES, default="rects",
                   help="training pairs: moving rectangles, or rectangles and large-motion regions")
    t.add_argument("--regions-share", type=float, default=0.5, help="share of region pairs in the mixed family")
    t.add_argument("--label-iterations", type=int, default=4,
                   help="propagation iterations of the method that labels the pairs (--labels plain, seeded)")
    t.add_argument("--out", required=True)
    t.add_argument("--eval", action="store_true", help="evaluate on the 1920x1080 and 1280x720 workloads afterwards")
    t.add_argument("--eval-seeds", type=int, default=8)
    e = sub.add_parser("eval")
    e.add_argument("--weights", default=DEFAULT_WEIGHTS)
    e.add_argument("--seeds", type=int, default=4)
    e.add_argument("--workload", default="c2_1920x1080_u2.9_n128", choices=sorted(synth.WORKLOADS))
    e.add_argument("--fewer", type=int, default=1,
                   help="propagation iterations of the learned, zero-network and shortened plain runs (0: the "
                        "seeded initialisation and its random search only)")
    e.add_argument("--jf-max", type=int, default=None, help="jump range of the learned and zero-network runs")
    e.add_argument("--local-search", default=None, metavar="RADIUS:SAMPLES",
                   help="local search around the seeds (radius in input pixels)")
    a = ap.parse_args(argv)
    log = lambda m: print(m, flush=True)
    if a.cmd == "train":
        ds = build_dataset(a.pairs, workers=a.workers, labels=a.labels, family=a.family,
                           regions_share=a.regions_share, iterations=a.label_iterations,
                           complement=a.complement, log=log)
        net = Net(weights=init_weights(a.seed))
        rows = np.arange(a.pairs - a.val)
        val = np.arange(a.pairs - a.val, a.pairs)
        log("[train] validation loss before: %.4f" % validation_loss(net, ds, val))
        train(net, ds, rows, epochs=a.epochs, batch=a.batch, lr=a.lr, seed=a.seed, log=log)
        log("[train] validation loss after: %.4f" % validation_loss(net, ds, val))
        net.save(a.out)
        log("[train] saved %s" % a.out)
        if a.eval:
            log(str(evaluate(net, seeds=tuple(range(a.eval_seeds)), log=log)))
            log(str(evaluate(net, workload="c3_1280x720_u4.9_n512", seeds=tuple(range(max(1, a.eval_seeds // 2))),
                             log=log)))
    else:
        net = Net(path=a.weights)
        ls = None
        if a.local_search:
            r, n = a.local_search.split(":")
            ls = (float(r), int(n))
        log(str(evaluate(net, workload=a.workload, seeds=tuple(range(a.seeds)), fewer=a.fewer, jf_max=a.jf_max,
                         local_search=ls, log=log)))
    return 0


if __name__ == "__main__":
    sys.exit(main())
