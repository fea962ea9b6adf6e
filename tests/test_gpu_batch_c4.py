"""The uneven-batch workload at its full size on the GPU: 64 pairs with sizes
uniform in [320x240, 1920x1080] (random aspect), 32- and 128-bit descriptors
and up-sampling factors in [1.5, 2.9] (synth.batch_specs(64, 0), 45 input and
236 matched megapixels).  The per-pair launches, the hipGraph of per-pair
branches and the multi-pair launches (each graph captured, then replayed)
give the same fields and flows bit for bit; every pair's match field equals
the oracle run on its own descriptors, pattern and random-search stream
(pair index i), and its flow is within 1e-3 px of the oracle's on bands of
rows at the top, middle and bottom."""
import numpy as np
import pytest

import briefmatch_amd as bm
import oracle
from briefmatch_amd import synth
from conftest import say

pytestmark = [pytest.mark.gpu, pytest.mark.slow]

N_PAIRS, SEED, CHUNK = 64, 1, 8
FLOW_TOL = 1e-3


@pytest.fixture(scope="module")
def c4():
    say("c4 workload: generating", N_PAIRS, "pairs")
    specs = synth.batch_specs(N_PAIRS, 0)
    frames = synth.batch_pairs(specs)
    params = [bm.derive_params(s["W"], s["H"], s["up"], s["up"], N=s["N"]) for s in specs]
    say("c4 workload: batch of", N_PAIRS, "pairs,", round(sum(p.sx * p.sy for p in params) / 1e6, 1),
        "matched Mpix")
    batch = bm.Batch(params, seed=SEED, streams=8)
    r1 = [bm.DeviceBuffer(array=f.frame1) for f in frames]
    r2 = [bm.DeviceBuffer(array=f.frame2) for f in frames]
    Us = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p in params]
    Vs = [bm.DeviceBuffer(p.sx_in * p.sy_in * 4) for p in params]
    runs = {}
    for name, mode, sched in (("per_pair", bm.BATCH_PER_PAIR, {}), ("graph", bm.BATCH_GRAPH, {}),
                              ("graph_replay", bm.BATCH_GRAPH, {}), ("multi", bm.BATCH_MULTI, {"branches": 1}),
                              ("multi_replay", bm.BATCH_MULTI, {"branches": 1}),
                              ("multi_lanes", bm.BATCH_MULTI, {"branches": 4, "order": 1, "persist": 4})):
        say("c4 workload: mode", name, sched)
        if sched:
            batch.set_schedule(**sched)
        if name != "per_pair":   # the swapped pairs first: no mode can pass on a buffer left behind
            batch.flow(r2, r1, Us, Vs, mode=bm.BATCH_PER_PAIR)
        batch.flow(r1, r2, Us, Vs, mode=mode)
        batch.sync()
        runs[name] = [(batch.ctx[i].field(), Us[i].download((p.sy_in, p.sx_in), np.float32),
                       Vs[i].download((p.sy_in, p.sx_in), np.float32)) for i, p in enumerate(params)]
    params = [c.params for c in batch.ctx]
    batch.close()
    return dict(specs=specs, frames=frames, params=params, runs=runs)


def test_c4_modes_bit_exact(c4):
    base = c4["runs"]["per_pair"]
    for name in ("graph", "graph_replay", "multi", "multi_replay", "multi_lanes"):
        for i, (a, b) in enumerate(zip(base, c4["runs"][name])):
            assert np.array_equal(a[0][0], b[0][0]) and np.array_equal(a[0][1], b[0][1]), (name, i)
            assert np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2]), (name, i)
    assert sorted(set(int(p.N) for p in c4["params"])) == [32, 128]


@pytest.mark.parametrize("chunk", range(N_PAIRS // CHUNK))
def test_c4_against_oracle(c4, chunk):
    for i in range(chunk * CHUNK, (chunk + 1) * CHUNK):
        p, fr = c4["params"][i], c4["frames"][i]
        (off, cost), U, V = c4["runs"]["multi_replay"][i]
        say("c4 oracle pair", i, "%dx%d N%d u%.2f" % (p.sx_in, p.sy_in, p.N, c4["specs"][i]["up"]))
        pat = oracle.make_pattern(SEED, p.N, p.patchArea)
        kern = oracle.gauss_kernel(p.sigma)
        g1, g2 = oracle.gray(fr.frame1), oracle.gray(fr.frame2)
        F1 = oracle.brief(oracle.upsample(g1, p.sx, p.sy), pat, p.N)
        F2 = oracle.brief(oracle.upsample(g2, p.sx, p.sy), pat, p.N)
        eo, ec = oracle.match(F1, F2, p.iterations, p.jfMax, bool(p.doubleProp), True, None, i, SEED)
        del F1, F2
        assert np.array_equal(off, eo), i
        assert np.array_equal(cost, ec), i
        W, H = p.sx_in, p.sy_in
        oU, oV = oracle.downsample(eo, W, H, kern)
        oUm = oracle.median1d(oracle.median1d(oU, 1, p.fsize), 0, p.fsize)
        oVm = oracle.median1d(oracle.median1d(oV, 1, p.fsize), 0, p.fsize)
        for r0, r1 in ((0, 8), (H // 2 - 4, H // 2 + 4), (H - 8, H)):
            tU, tV = oracle.trilateral(oU, oV, oUm, oVm, g1.astype(np.float32), p.sEPE ** 2, p.sI ** 2,
                                       p.sS ** 2, r0, r1)
            assert np.abs(U[r0:r1] - tU[r0:r1]).max() <= FLOW_TOL, (i, r0)
            assert np.abs(V[r0:r1] - tV[r0:r1]).max() <= FLOW_TOL, (i, r0)
