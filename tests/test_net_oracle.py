"""Pins the learned-proposal oracle (oracle/net.py) on the CPU: BF16 rounding
against exact nearest-value search, the conv against its textbook definition,
the data and weight gradients against the adjoint identities, the whole
backward pass against central finite differences, and the feature map,
labels, offsets and seeded initialisation against element-by-element loops."""
import numpy as np
import pytest

from oracle import net as onet


def test_bf16_known_values_and_ties():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 3 * 2 ** -8, -2.5, 0.0, 1.0 + 2 ** -8 + 2 ** -20], np.float32)
    bits = onet.bf16_bits(x)
    assert bits[0] == 0x3F80
    assert bits[1] == 0x3F80          # tie -> even
    assert bits[2] == 0x3F82          # tie -> even (up)
    assert bits[3] == 0xC020
    assert bits[4] == 0
    assert bits[5] == 0x3F81          # just above the tie -> up


def test_bf16_is_nearest():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(20000) * 10.0 ** rng.uniform(-6, 6, 20000)).astype(np.float32)
    r = onet.bf16_float(onet.bf16_bits(x)).astype(np.float64)
    # the two BF16 neighbours of x: truncation and one step away from zero
    t = (x.view(np.uint32) & 0xFFFF0000).view(np.float32).astype(np.float64)
    step = ((x.view(np.uint32) & 0xFFFF0000) + 0x10000).view(np.float32).astype(np.float64)
    xd = x.astype(np.float64)
    best = np.where(np.abs(xd - t) <= np.abs(step - xd), t, step)
    tie = np.abs(xd - t) == np.abs(step - xd)
    assert np.array_equal(r[~tie], best[~tie])


def _conv_loops(x, Wl, b, dil):
    n, H, W, ci = x.shape
    co = Wl.shape[1]
    y = np.zeros((n, H, W, co))
    for i in range(n):
        for r in range(H):
            for c in range(W):
                acc = b.astype(np.float64).copy()
                for tap in range(9):
                    rr, cc = r + (tap // 3 - 1) * dil, c + (tap % 3 - 1) * dil
                    if 0 <= rr < H and 0 <= cc < W:
                        acc += Wl[tap] @ x[i, rr, cc]
                y[i, r, c] = acc
    return y


@pytest.mark.parametrize("dil", [1, 2, 4])
def test_conv_matches_definition(dil):
    rng = np.random.default_rng(dil)
    x = rng.standard_normal((2, 7, 9, 5))
    Wl = rng.standard_normal((9, 3, 5))
    b = rng.standard_normal(3)
    assert np.allclose(onet.conv(x, Wl, b, dil), _conv_loops(x, Wl, b, dil), atol=1e-12)


@pytest.mark.parametrize("dil", [1, 2, 4])
def test_conv_gradients_are_adjoint(dil):
    """<conv(x, W), y> = <x, dgrad(y, W)> = <W, wgrad(y, x)>."""
    rng = np.random.default_rng(10 + dil)
    x = rng.standard_normal((2, 6, 11, 4))
    Wl = rng.standard_normal((9, 3, 4))
    y = rng.standard_normal((2, 6, 11, 3))
    lhs = np.sum(onet.conv(x, Wl, np.zeros(3), dil) * y)
    assert np.isclose(lhs, np.sum(x * onet.conv_data_grad(y, Wl, dil)), rtol=1e-12)
    gW, gb = onet.conv_weight_grad(y, x, dil)
    assert np.isclose(lhs, np.sum(Wl * gW), rtol=1e-12)
    assert np.allclose(gb, y.reshape(-1, 3).sum(0))


def test_backward_matches_finite_differences():
    rng = np.random.default_rng(3)
    from briefmatch_amd.learned import init_weights
    w = init_weights(5).astype(np.float64)
    w[-16:] = rng.standard_normal(16) * 0.1          # non-zero biases of the last layer
    feat = rng.uniform(-1, 1, (1, 6, 7, 128))
    label = rng.standard_normal((1, 6, 7, 2)) * 0.05
    label[0, 0, 0, 1] = np.nan                        # ignored label
    loss, g = onet.loss_and_grad(w, feat, label, quant=False)
    idx = np.concatenate([rng.choice(onet.N_PARAMS - 9 * 16 * 64 - 16, 50, replace=False),
                          onet.N_PARAMS - 16 + np.arange(2)])
    h = 1e-6
    for i in idx:
        wp, wm = w.copy(), w.copy()
        wp[i] += h
        wm[i] -= h
        fd = (onet.loss_and_grad(wp, feat, label, quant=False)[0] - onet.loss_and_grad(wm, feat, label, quant=False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-6 + 1e-4 * abs(g[i]), (i, fd, g[i])
    # channels 2..15 of the last layer receive no gradient
    last = onet.split_params(g)[-1]
    assert not last[0][:, 2:].any() and not last[1][2:].any()


def test_features_match_loops():
    rng = np.random.default_rng(4)
    W, H = 56, 44                        # 7 x 5 blocks, remainder dropped
    g1 = rng.integers(0, 256, (H, W), dtype=np.uint8)
    g2 = rng.integers(0, 256, (H, W), dtype=np.uint8)
    g2[:, 8:16] = g1[:, 0:8]             # some exact matches (zero SAD)
    got = onet.bf16_float(onet.features(g1, g2))
    cw, ch = W // 8, H // 8
    S = [[[int(g[8 * y:8 * y + 8, 8 * x:8 * x + 8].astype(np.int64).sum()) for x in range(cw)] for y in range(ch)]
         for g in (g1, g2)]
    cl = lambda v, n: min(max(v, 0), n - 1)
    for cy in range(ch):
        for cx in range(cw):
            best = None
            for c in range(128):
                if c < 121:
                    dx, dy = c % 11 - 5, c // 11 - 5
                    sad = sum(abs(S[0][cl(cy + j, ch)][cl(cx + i, cw)] - S[1][cl(cy + j + dy, ch)][cl(cx + i + dx, cw)])
                              for j in (-1, 0, 1) for i in (-1, 0, 1))
                    want = min(sad, 36864) / 18432.0 - 1.0
                    if best is None or sad < best[0]:
                        best = (sad, dx, dy)
                elif c < 123:
                    want = S[c - 121][cy][cx] / (64 * 255.0) - 0.5
                elif c < 126:
                    want = (best[1] / 8.0, best[2] / 8.0, min(best[0], 36864) / 18432.0 - 1.0)[c - 123]
                else:
                    want = 0.0
                assert abs(got[cy, cx, c] - want) <= 2 ** -8 * abs(want) + 1e-12, (cy, cx, c)


def test_coarse_flow_is_block_mean():
    rng = np.random.default_rng(5)
    U = rng.standard_normal((20, 27)).astype(np.float32) * 10
    V = rng.standard_normal((20, 27)).astype(np.float32) * 10
    got = onet.coarse_flow(U, V)
    assert got.shape == (2, 3, 2)
    want = U[:16, :24].astype(np.float64).reshape(2, 8, 3, 8).mean(axis=(1, 3)) / 8
    assert np.allclose(got[..., 0], want, atol=1e-6)


def test_proposal_offsets_known_values():
    out = np.zeros((1, 2, 16), np.float32)
    out[0, 0, :2] = [10.0, -3.0]
    out[0, 1, :2] = [np.nan, 1e30]
    feat = np.zeros((1, 2, 128), np.float32)
    feat[0, 0, 123:125] = [3 / 8, -5 / 8]
    got = onet.proposal_offsets(out, onet.bf16_bits(feat), W=100, H=50, sx=290, sy=145)
    assert got[0, 0, 0] == round(10.0 * 2048 * 290 / 100)
    assert got[0, 0, 1] == round(-3.0 * 2048 * 145 / 50)
    assert got[0, 1, 0] == -256 * 290 and got[0, 1, 1] == 256 * 145
    assert got[0, 0, 2] == round(3 * 2048 * 290 / 100) and got[0, 0, 3] == round(-5 * 2048 * 145 / 50)
    assert got[0, 1, 2] == 0 and got[0, 1, 3] == 0


def _textured(rng, H, W):
    return rng.integers(0, 256, (H, W)).astype(np.uint8)


@pytest.mark.parametrize("sx,sy", [(4, 0), (0, -3), (-5, 2), (3, 3)])
def test_far_search_translation(sx, sy):
    """Frame 2 = frame 1 moved by (16 sx, 16 sy) pixels: the interior blocks,
    whose windows never clamp, find exactly (sx, sy) (a wrong sign or a
    swapped axis fails here)."""
    rng = np.random.default_rng(100 + sx + 11 * sy)
    g1 = _textured(rng, 256, 272)
    g2 = _textured(rng, 256, 272)
    dy, dx = 16 * sy, 16 * sx
    ys, xs = slice(max(dy, 0), 256 + min(dy, 0)), slice(max(dx, 0), 272 + min(dx, 0))
    g2[ys, xs] = g1[max(-dy, 0):256 - max(dy, 0), max(-dx, 0):272 - max(dx, 0)]
    got = onet.far_search(g1, g2)
    assert got.shape == (16, 17, 2) and got.dtype == np.int32
    inner = got[6:10, 6:11]
    assert (inner[..., 0] == sx).all() and (inner[..., 1] == sy).all()


def test_far_search_ties():
    """Equal costs resolve to the lowest ring, then the lowest index:
    flat frames give (0, 0) everywhere (not the first displacement (-5, -5));
    a texture that repeats every 32 pixels in x, moved by 16 pixels, matches
    at dx = -5, -3, -1, 1, 3, 5 with equal cost and gives (-1, 0)."""
    flat = np.full((96, 128), 77, np.uint8)
    assert (onet.far_search(flat, flat) == 0).all()
    rng = np.random.default_rng(3)
    col = rng.integers(0, 256, (224, 32)).astype(np.uint8)
    g1 = np.tile(col, (1, 7))
    g2 = np.roll(g1, 16, axis=1)
    got = onet.far_search(g1, g2)
    inner = got[6:8, 6:8]
    assert (inner[..., 0] == -1).all() and (inner[..., 1] == 0).all()
    # the same texture, not moved: rings 0, 2, 4 tie, ring 0 wins
    assert (onet.far_search(g1, g1)[6:8, 6:8] == 0).all()


def test_far_search_brute_force():
    """Per-block loops over the displacements, with 16x16 sums taken from the
    image directly, on a pair with regions moving up to 60 pixels: the far
    match where it beats 3/4 of the best ring-0..2 sum, else that nearer
    displacement; both outcomes occur."""
    import oracle
    from briefmatch_amd import synth
    pr = synth.make_pair(160, 128, motion="regions", n_regions=4, max_motion=60.0, seed=4)
    g1, g2 = oracle.gray(pr.frame1).astype(np.int64), oracle.gray(pr.frame2).astype(np.int64)
    ch, cw = 128 // 16, 160 // 16
    bsum = lambda g, by, bx: int(g[16 * by:16 * by + 16, 16 * bx:16 * bx + 16].sum())
    S1 = [[bsum(g1, y, x) for x in range(cw)] for y in range(ch)]
    S2 = [[bsum(g2, y, x) for x in range(cw)] for y in range(ch)]
    cl = lambda v, n: min(max(v, 0), n - 1)
    got = onet.far_search(g1.astype(np.uint8), g2.astype(np.uint8))
    fell_back = took_far = 0
    for y in range(ch):
        for x in range(cw):
            scored = []
            for dy in range(-5, 6):
                for dx in range(-5, 6):
                    sad = sum(abs(S1[cl(y + j, ch)][cl(x + i, cw)] - S2[cl(y + j + dy, ch)][cl(x + i + dx, cw)])
                              for j in (-1, 0, 1) for i in (-1, 0, 1))
                    scored.append((sad, max(abs(dx), abs(dy)), dy, dx))
            best = min(scored)
            near = min(t for t in scored if t[1] <= 2)
            pick = best if 4 * best[0] < 3 * near[0] else near
            assert tuple(got[y, x]) == (pick[3], pick[2]), (y, x)
            fell_back += best[1] >= 3 and pick is near
            took_far += pick[1] >= 3
    assert fell_back > 0 and took_far > 0


def test_far_search_small_frames():
    assert onet.far_search(np.zeros((15, 40), np.uint8), np.zeros((15, 40), np.uint8)).shape == (0, 2, 2)


def test_proposal_offsets_far_override():
    """A far match at ring >= 3 replaces the network's flow of the cells under
    its block (cells 2X, 2X + 1; the odd last column uses the last block) by
    (2 dx, 2 dy) coarse pixels; rings below 3 leave it."""
    W, H, sx, sy = 104, 24, 151, 35
    ch, cw = H // 8, W // 8                              # 3 x 13 cells, 1 x 6 blocks
    out = np.full((ch, cw, 16), 0.5, np.float32)
    feat = onet.bf16_bits(np.zeros((ch, cw, 128), np.float32))
    far = np.array([[[3, -1], [2, 2], [0, 0], [1, -2], [-2, 1], [-4, 0]]], np.int32)
    got = onet.proposal_offsets(out, feat, W, H, sx, sy, far=far)
    base = onet.proposal_offsets(out, feat, W, H, sx, sy)
    for cy in range(ch):
        for cx in range(cw):
            X = min(cx // 2, 5)
            if X in (0, 5):
                dx, dy = far[0, X]
                assert got[cy, cx, 0] == round(2 * dx * 2048 * sx / W)
                assert got[cy, cx, 1] == round(2 * dy * 2048 * sy / H)
            else:
                assert (got[cy, cx, :2] == base[cy, cx, :2]).all()
            assert (got[cy, cx, 2:] == base[cy, cx, 2:]).all()
    assert base[0, 0, 0] == round(0.5 * 2048 * sx / W)


def _init_loops(F1, F2, coff, W, H, first, prev):
    sy, sx = F1.shape[:2]
    ch, cw = coff.shape[:2]
    pop = lambda a: int(sum(bin(int(v)).count("1") for v in a))
    unp = lambda o: (int(np.int16(np.uint16(o & 0xFFFF))), int(np.int16(np.uint16(o >> 16))))
    pk = lambda dx, dy: (dx & 0xFFFF) | ((dy & 0xFFFF) << 16)
    off = np.zeros((sy, sx), np.uint32)
    cost = np.zeros((sy, sx), np.uint16)
    for y in range(sy):
        for x in range(sx):
            nx = ny = 0
            if not first:
                dx0, dy0 = unp(int(prev[y, x]))
                dx0, dy0 = min(max(dx0, -x), sx - 1 - x), min(max(dy0, -y), sy - 1 - y)
                dx1, dy1 = unp(int(prev[y + dy0, x + dx0]))
                nx, ny = min(max(dx1, -x), sx - 1 - x), min(max(dy1, -y), sy - 1 - y)
            best, bo = pop(F1[y, x] ^ F2[y + ny, x + nx]), pk(nx, ny)
            fx = min(max(x * W * 32 // sx - 112, 0), (cw - 1) * 256)
            fy = min(max(y * H * 32 // sy - 112, 0), (ch - 1) * 256)
            x0, y0 = fx >> 8, fy >> 8
            x1, y1 = min(x0 + 1, cw - 1), min(y0 + 1, ch - 1)
            ax, ay = (fx & 255) / 256.0, (fy & 255) / 256.0
            cells = [(y0, x0), (y0, x1), (y1, x0), (y1, x1)]
            cands = []
            for src in range(2):
                o = [coff[r, c, 2 * src:2 * src + 2].astype(np.float64) / 256.0 for r, c in cells]
                bil = ((o[0] * (1 - ax) + o[1] * ax) * (1 - ay) + (o[2] * (1 - ax) + o[3] * ax) * ay)
                cands += [tuple(int(np.floor(v + 0.5)) for v in bil)]
                cands += [tuple(int(np.floor(v + 0.5)) for v in oc) for oc in o]
            for cx, cy in cands:
                if not (0 <= x + cx < sx and 0 <= y + cy < sy) or pk(cx, cy) == bo:
                    continue
                c = pop(F1[y, x] ^ F2[y + cy, x + cx])
                if c < best:
                    best, bo = c, pk(cx, cy)
            off[y, x], cost[y, x] = bo, best
    return off, cost


@pytest.mark.parametrize("first", [True, False])
def test_init_proposal_matches_loops(first):
    rng = np.random.default_rng(6 + first)
    W, H, sx, sy = 32, 24, 47, 35
    # few distinct descriptors so that candidates tie and collide
    pal = rng.integers(0, 2 ** 32, (6, 2), dtype=np.uint64).astype(np.uint32)
    F1 = pal[rng.integers(0, 6, (sy, sx))]
    F2 = pal[rng.integers(0, 6, (sy, sx))]
    coff = rng.integers(-12 * 256, 12 * 256, (H // 8, W // 8, 4)).astype(np.int32)
    coff[0, 0] = [0, 0, 0, 0]
    coff[1, 1, 2:] = coff[1, 1, :2]                 # the two proposals agree in one cell
    prev = None
    if not first:
        dx = rng.integers(-6, 7, (sy, sx))
        dy = rng.integers(-6, 7, (sy, sx))
        prev = ((dx & 0xFFFF) | ((dy & 0xFFFF) << 16)).astype(np.uint32)
    got = onet.init_proposal(F1, F2, coff, W, H, first=first, prev_off=prev)
    want = _init_loops(F1, F2, coff, W, H, first, prev)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_oracle_plain_candidate_is_orc_init():
    """With a proposal that never wins, the seeded initialisation is the plain one."""
    import oracle
    rng = np.random.default_rng(8)
    sx, sy = 40, 30
    F1 = rng.integers(0, 2 ** 32, (sy, sx, 1), dtype=np.uint64).astype(np.uint32)
    F2 = F1.copy()                                  # zero offset is optimal (cost 0)
    coff = rng.integers(-2000, 2000, (2, 3, 4)).astype(np.int32)
    off, cost = onet.init_proposal(F1, F2, coff, 24, 16)
    o2, c2 = oracle.init(F1, F2, first=True)
    assert np.array_equal(off, o2) and np.array_equal(cost, c2)


def test_training_families():
    """The mixed training family: a drawn share of the seeds are region pairs
    with motion of 20..60 px, the others the rectangle pairs of the default
    family; the known flow comes with each pair."""
    from briefmatch_amd import learned
    regions = [s for s in range(1000, 1040)
               if not np.array_equal(learned.train_pair(s, 96, 64, True, "mixed", 0.25)[2],
                                     learned.train_pair(s, 96, 64, True, "rects")[2])]
    assert 3 <= len(regions) <= 20
    f1, f2, gt = learned.train_pair(regions[0], 160, 120, with_gt=True, family="mixed", regions_share=0.25)
    assert f1.shape == (120, 160, 3) and gt.shape == (120, 160, 2)
    assert 20.0 - 1e-3 <= np.sqrt((gt.astype(np.float64) ** 2).sum(-1)).max() <= 60.0 + 1e-3
    assert all(np.array_equal(learned.train_pair(s, 96, 64, True, "mixed", 0.0)[2],
                              learned.train_pair(s, 96, 64, True, "rects")[2]) for s in range(1000, 1004))
    with pytest.raises(ValueError):
        learned.train_pair(1, family="other")


def test_philox_np_equals_oracle_philox():
    """The vectorised Philox of the local search against the oracle's scalar
    one (itself pinned by the Random123 known-answer vectors)."""
    import oracle
    rng = np.random.default_rng(12)
    ctr = rng.integers(0, 2 ** 32, (64, 4), dtype=np.uint64).astype(np.uint32)
    key = rng.integers(0, 2 ** 32, (64, 2), dtype=np.uint64).astype(np.uint32)
    ctr[0] = [0, 0, 0, 0]
    ctr[1] = [0xFFFFFFFF] * 4
    for c, k in zip(ctr, key):
        got = onet.philox_np(c[0], c[1], c[2], c[3], k[0], k[1])
        assert [int(v) for v in got] == [int(v) for v in oracle.philox(c, k)]


def _local_search_loops(F1, F2, off, cost, radius, samples, iteration, pair, seed):
    import oracle
    sy, sx = F1.shape[:2]
    key = np.array([seed & 0xFFFFFFFF, seed >> 32], np.uint32)
    off, cost = off.copy(), cost.astype(np.int64)
    pop = lambda a: int(sum(bin(int(v)).count("1") for v in a))
    for y in range(sy):
        for x in range(sx):
            bo, best = int(off[y, x]), int(cost[y, x])
            for s in range(samples):
                r = max(radius >> s, 1)
                w = oracle.philox(np.array([y * sx + x, iteration, pair, 1 + s], np.uint32), key)
                dx = (int(w[0]) * (2 * r + 1) >> 32) - r
                dy = (int(w[1]) * (2 * r + 1) >> 32) - r
                cx = ((bo & 0xFFFF) ^ 0x8000) - 0x8000
                cy = ((bo >> 16) ^ 0x8000) - 0x8000
                nx = min(max(cx + dx, -x), sx - 1 - x)
                ny = min(max(cy + dy, -y), sy - 1 - y)
                po = (nx & 0xFFFF) | ((ny & 0xFFFF) << 16)
                if po == bo:
                    continue
                c = pop(F1[y, x] ^ F2[y + ny, x + nx])
                if c < best:
                    best, bo = c, po
            off[y, x], cost[y, x] = bo, best
    return off, cost.astype(np.uint16)


@pytest.mark.parametrize("radius,samples", [(6, 4), (1, 3), (40, 7)])
def test_local_search_matches_loops(radius, samples):
    rng = np.random.default_rng(radius)
    sx, sy = 37, 23
    # few distinct descriptors so that candidates tie; offsets near the border
    pal = rng.integers(0, 2 ** 32, (5, 2), dtype=np.uint64).astype(np.uint32)
    F1 = pal[rng.integers(0, 5, (sy, sx))]
    F2 = pal[rng.integers(0, 5, (sy, sx))]
    yy, xx = np.mgrid[0:sy, 0:sx]
    dx = np.clip(rng.integers(-9, 10, (sy, sx)), -xx, sx - 1 - xx)
    dy = np.clip(rng.integers(-9, 10, (sy, sx)), -yy, sy - 1 - yy)
    off = ((dx & 0xFFFF) | ((dy & 0xFFFF) << 16)).astype(np.uint32)
    pop = np.vectorize(lambda v: bin(int(v)).count("1"))
    cost = pop(F1 ^ F2[yy + dy, xx + dx]).sum(-1).astype(np.uint16)
    seed = 0x1234567890AB
    got = onet.local_search(F1, F2, off, cost, radius, samples, 3, 5, seed)
    want = _local_search_loops(F1, F2, off, cost, radius, samples, 3, 5, seed)
    assert np.array_equal(got[0], want[0]) and np.array_equal(got[1], want[1])


def test_local_search_hit_rate_closed_form():
    """Random 64-bit descriptors, frame 2 = frame 1 moved by (3, -2) (a torus,
    so every pixel has its match), zero start offsets given cost 1: only the
    exact offset (cost 0) is strictly cheaper, so a pixel's offset changes
    only when a sample hits (3, -2) exactly.  From zero, sample s hits it with
    probability 1 / (2 r_s + 1)^2 when max(|3|, |-2|) <= r_s, else 0; after a
    hit the best is the translation and later samples cannot beat cost 0.
    The share of pixels that find it matches 1 - prod(1 - p_s) within four
    standard errors; the offsets found are exactly the translation."""
    rng = np.random.default_rng(21)
    sx, sy = 200, 150
    F2 = rng.integers(0, 2 ** 32, (sy, sx, 2), dtype=np.uint64).astype(np.uint32)
    F1 = np.roll(F2, shift=(2, -3), axis=(0, 1))      # F1[y, x] = F2[y - 2, x + 3]
    off = np.zeros((sy, sx), np.uint32)
    cost = np.ones((sy, sx), np.uint16)
    radius, samples = 8, 5                           # r = 8, 4, 2, 1, 1
    o, c = onet.local_search(F1, F2, off, cost, radius, samples, 0, 0, 77)
    # interior pixels only: near the border the clamp changes the odds
    inner = (slice(10, sy - 10), slice(10, sx - 10))
    hit = o[inner] == ((3 & 0xFFFF) | ((-2 & 0xFFFF) << 16))
    assert np.all(c[inner][hit] == 0)
    p_miss = 1.0
    for s in range(samples):
        r = max(radius >> s, 1)
        if r >= 3:
            p_miss *= 1.0 - 1.0 / (2 * r + 1) ** 2
    p = 1.0 - p_miss
    n = hit.size
    assert abs(hit.mean() - p) <= 4.0 * np.sqrt(p * (1 - p) / n)
    assert np.all(o[inner][~hit] == 0) and np.all(c[inner][~hit] == 1)
