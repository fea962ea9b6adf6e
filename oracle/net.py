"""CPU oracle of the learned proposal -- TEST INFRASTRUCTURE ONLY.

numpy restatement of the learned-proposal path: the feature map (a block
matching cost volume at 1/8 resolution), the coarse
labels, the network (forward and backward, float64, optionally with the BF16
rounding points of the device kernels), the conversion of its output to
matched-pixel offsets (with the far block match that overrides the network
where it finds motion beyond the cost volume's reach) and the
proposal-seeded initialisation.  Tests and the
smoke check use it as the checker; the product never imports it.

The reference has no learned component.  The seeded initialisation extends the
reference's initialisation (kernels/briefmatch_cuda.cu:149-184: zero offset on
the first frame, else the previous field followed once); its plain candidate
is that rule, as in oracle_bm.cpp's orc_init.

Layouts: descriptors [sy, sx, K] uint32 (K >= ceil(N / 32), extra words zero),
gray frames [H, W] uint8, feature maps [n, ch, cw, 128] as BF16 bit patterns
(uint16), network outputs [n, ch, cw, 16] float, labels [n, ch, cw, 2]; flows
at the coarse grid are in coarse pixels (8 input pixels).
"""
import numpy as np

RADIUS = 5
COST = (2 * RADIUS + 1) ** 2
FEAT_C = 128
OUT_C = 16
CELL = 8
SKIP = COST + 2          # feature channels of the lowest-cost displacement / 8
FAR_CELL = 16            # block size of the far search
FAR_RADIUS = 5           # far search displacements -5..5 blocks (+-80 input pixels)
FAR_MIN = 3              # a far match replaces the network proposal from this ring on
# (cin, cout, dilation, relu) per layer
LAYERS = ((128, 64, 1, True), (64, 64, 2, True), (64, 64, 4, True), (64, 64, 1, True), (64, 16, 1, False))
N_PARAMS = sum(9 * co * ci + co for ci, co, _, _ in LAYERS)


# ---------------------------------------------------------------------------
# BF16

def bf16_bits(x):
    """float32 -> BF16 bit patterns, round to nearest even."""
    u = np.ascontiguousarray(x, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) >> 16
    return u.astype(np.uint16)


def bf16_float(bits):
    return (np.asarray(bits, np.uint16).astype(np.uint32) << 16).view(np.float32)


def bf16_round(x):
    """Round float values to the nearest BF16 value (returned as float64)."""
    return bf16_float(bf16_bits(np.asarray(x, np.float32))).astype(np.float64)


# ---------------------------------------------------------------------------
# features, labels

def _popcount_rows(a):
    return np.bitwise_count(a).sum(axis=-1, dtype=np.int64)


SAD_CLAMP = 9 * 64 * 64


def block_sums(g):
    """8x8 block sums [ch, cw] of a gray frame (remainder rows / columns dropped)."""
    H, W = g.shape
    cw, ch = W // CELL, H // CELL
    return g[:ch * CELL, :cw * CELL].astype(np.int64).reshape(ch, CELL, cw, CELL).sum(axis=(1, 3))


def features(g1, g2):
    """Feature map [ch, cw, 128] (BF16 bits) of one pair of gray frames.

    Channel (dy + 5) * 11 + dx + 5 for dx, dy in -5..5: the sum of absolute
    differences between the 8x8 block sums of frame 1 over the 3x3 block
    window of the coarse pixel and those of frame 2 displaced by (dx, dy)
    blocks (coordinates clamped to the grid), as min(SAD, 36864) / 18432 - 1.
    Then the block means of both frames / 255 - 1/2, the lowest-SAD
    displacement (first channel on ties) as dx / 8, dy / 8, its scaled SAD,
    and zeros."""
    S1, S2 = block_sums(g1), block_sums(g2)
    ch, cw = S1.shape
    out = np.zeros((ch, cw, FEAT_C), np.float32)
    costs = np.zeros((ch, cw, COST), np.int64)
    yy, xx = np.mgrid[0:ch, 0:cw]
    half = SAD_CLAMP // 2
    for c in range(COST):
        dx, dy = c % (2 * RADIUS + 1) - RADIUS, c // (2 * RADIUS + 1) - RADIUS
        cost = np.zeros((ch, cw), np.int64)
        for j in (-1, 0, 1):
            for i in (-1, 0, 1):
                a = S1[np.clip(yy + j, 0, ch - 1), np.clip(xx + i, 0, cw - 1)]
                b = S2[np.clip(yy + j + dy, 0, ch - 1), np.clip(xx + i + dx, 0, cw - 1)]
                cost += np.abs(a - b)
        costs[..., c] = cost
        out[..., c] = (np.minimum(cost, SAD_CLAMP) - half).astype(np.float32) / np.float32(half)
    out[..., COST] = (S1 - 8160).astype(np.float32) / np.float32(16320)
    out[..., COST + 1] = (S2 - 8160).astype(np.float32) / np.float32(16320)
    bc = np.argmin(costs, axis=-1)
    bcost = np.take_along_axis(costs, bc[..., None], -1)[..., 0]
    out[..., COST + 2] = (bc % (2 * RADIUS + 1) - RADIUS).astype(np.float32) / np.float32(8)
    out[..., COST + 3] = (bc // (2 * RADIUS + 1) - RADIUS).astype(np.float32) / np.float32(8)
    out[..., COST + 4] = (np.minimum(bcost, SAD_CLAMP) - half).astype(np.float32) / np.float32(half)
    return bf16_bits(out)


def coarse_flow(U, V):
    """8x8 block means divided by 8 (coarse pixels) [ch, cw, 2] in float32,
    summed row by row in order."""
    H, W = U.shape
    cw, ch = W // CELL, H // CELL
    su = np.zeros((ch, cw), np.float32)
    sv = np.zeros((ch, cw), np.float32)
    for yy in range(CELL):
        for xx in range(CELL):
            su = su + U[yy:ch * CELL:CELL, xx:cw * CELL:CELL][:ch, :cw]
            sv = sv + V[yy:ch * CELL:CELL, xx:cw * CELL:CELL][:ch, :cw]
    return np.stack([su * np.float32(1 / 512), sv * np.float32(1 / 512)], -1)


# ---------------------------------------------------------------------------
# proposal offsets and the seeded initialisation

def far_search(g1, g2):
    """Far block match [ch16, cw16, 2] int32 (dx, dy in 16-pixel blocks) of one
    pair of gray frames, ch16 = H // 16, cw16 = W // 16.

    For dx, dy in -5..5: the sum of absolute differences between the 16x16
    block sums of frame 1 over the 3x3 block window and those of frame 2
    displaced by (dx, dy) blocks (coordinates clamped to the grid).  The best
    displacement minimises the sum, ties going to the lower ring
    max(|dx|, |dy|) and then to the lower (dy + 5) * 11 + dx + 5 (so that
    equal costs never favour a far displacement); the block reports it when
    its sum is below 3/4 of the best sum over rings 0..2, else the best
    displacement of rings 0..2.  No counterpart in the reference."""
    H, W = g1.shape
    cw, ch = W // FAR_CELL, H // FAR_CELL
    res = np.zeros((ch, cw, 2), np.int32)
    if cw < 1 or ch < 1:
        return res
    S1, S2 = [g[:ch * FAR_CELL, :cw * FAR_CELL].astype(np.int64).reshape(ch, FAR_CELL, cw, FAR_CELL).sum(axis=(1, 3))
              for g in (g1, g2)]
    yy, xx = np.mgrid[0:ch, 0:cw]
    n = 2 * FAR_RADIUS + 1
    big = np.iinfo(np.int64).max
    best = [np.full((ch, cw), big), np.full((ch, cw), big)]      # all rings, rings < FAR_MIN
    cost_of = [np.zeros((ch, cw), np.int64), np.zeros((ch, cw), np.int64)]
    disp = [np.zeros((ch, cw, 2), np.int32), np.zeros((ch, cw, 2), np.int32)]
    for c in range(n * n):
        dx, dy = c % n - FAR_RADIUS, c // n - FAR_RADIUS
        ring = max(abs(dx), abs(dy))
        cost = np.zeros((ch, cw), np.int64)
        for j in (-1, 0, 1):
            for i in (-1, 0, 1):
                a = S1[np.clip(yy + j, 0, ch - 1), np.clip(xx + i, 0, cw - 1)]
                b = S2[np.clip(yy + j + dy, 0, ch - 1), np.clip(xx + i + dx, 0, cw - 1)]
                cost += np.abs(a - b)
        key = (cost * (FAR_RADIUS + 1) + ring) * (n * n) + c
        for k in (0, 1) if ring < FAR_MIN else (0,):
            take = key < best[k]
            best[k] = np.where(take, key, best[k])
            cost_of[k] = np.where(take, cost, cost_of[k])
            disp[k][take] = (dx, dy)
    far = 4 * cost_of[0] < 3 * cost_of[1]
    return np.where(far[..., None], disp[0], disp[1]).astype(np.int32)


def proposal_offsets(net_out, feat, W, H, sx, sy, far=None):
    """Per cell [.., 4] int32 Q8 matched-pixel offsets: the network output
    (channels 0, 1, coarse pixels) and the feature map's lowest-cost
    displacement (8 x channels 123, 124), each rint(clamp(u * 2048 sx / W,
    +-256 sx)) (v with sy, H).  With a far match (far_search), a cell whose
    16-pixel block (cx // 2, cy // 2, clamped to that grid) matched at ring
    max(|dx|, |dy|) >= 3 takes (2 dx, 2 dy) coarse pixels in place of the
    network output."""
    out = np.asarray(net_out, np.float32)
    am = bf16_float(feat)[..., SKIP:SKIP + 2] * np.float32(8)
    u0, v0 = out[..., 0], out[..., 1]
    if far is not None and far.size:
        ch, cw = out.shape[-3:-1]
        fy = np.minimum(np.arange(ch) // 2, far.shape[0] - 1)
        fx = np.minimum(np.arange(cw) // 2, far.shape[1] - 1)
        d = far[fy][:, fx]
        sel = np.abs(d).max(axis=-1) >= FAR_MIN
        u0 = np.where(sel, (2 * d[..., 0]).astype(np.float32), u0)
        v0 = np.where(sel, (2 * d[..., 1]).astype(np.float32), v0)
    srcs = (u0, v0, am[..., 0], am[..., 1])
    res = []
    for k, u in enumerate(srcs):
        s, d = (sy, H) if k & 1 else (sx, W)
        scale = np.float32(256 * CELL * s) / np.float32(d)
        lim = np.float32(256 * s)
        v = np.fmin(np.fmax(u.astype(np.float32) * scale, -lim), lim)
        res.append(np.rint(v).astype(np.int32))
    return np.stack(res, -1)


def _unpack(o):
    o = np.asarray(o, np.uint32)
    return ((o & 0xFFFF).astype(np.uint16).view(np.int16).astype(np.int64),
            (o >> 16).astype(np.uint16).view(np.int16).astype(np.int64))


def _pack(dx, dy):
    return ((np.asarray(dx, np.int64) & 0xFFFF) | ((np.asarray(dy, np.int64) & 0xFFFF) << 16)).astype(np.uint32)


def init_proposal(F1, F2, coff, W, H, first=True, prev_off=None):
    """Proposal-seeded initialisation; returns (off [sy, sx] uint32, cost uint16).

    Candidates in order: the plain initialisation (zero, or prev followed once
    with clamping, briefmatch_cuda.cu:149-184); then for the network's
    proposal (coff[..., 0:2]) and then the argmin proposal (coff[..., 2:4]):
    the bilinear interpolation of the Q8 offsets of the four cells around the
    pixel (coarse coordinate x * W * 32 / sx - 112 in Q8, clamped to the grid;
    sum rounded half up at 2^-24), then the four cells' offsets rounded half
    up.  Skip out-of-frame candidates and candidates equal to the running
    best; take strictly lower costs."""
    sy, sx = F1.shape[:2]
    ch, cw = coff.shape[:2]
    yy, xx = np.mgrid[0:sy, 0:sx].astype(np.int64)
    if first:
        nx = np.zeros((sy, sx), np.int64)
        ny = np.zeros((sy, sx), np.int64)
    else:
        dx0, dy0 = _unpack(prev_off)
        dx0 = np.clip(dx0, -xx, sx - 1 - xx)
        dy0 = np.clip(dy0, -yy, sy - 1 - yy)
        dx1, dy1 = _unpack(np.asarray(prev_off)[yy + dy0, xx + dx0])
        nx = np.clip(dx1, -xx, sx - 1 - xx)
        ny = np.clip(dy1, -yy, sy - 1 - yy)
    best = _popcount_rows(F1 ^ F2[yy + ny, xx + nx])
    bo = _pack(nx, ny)

    qx = np.clip(xx * W * 32 // sx - 112, 0, (cw - 1) * 256)
    qy = np.clip(yy * H * 32 // sy - 112, 0, (ch - 1) * 256)
    x0, y0 = qx >> 8, qy >> 8
    x1, y1 = np.minimum(x0 + 1, cw - 1), np.minimum(y0 + 1, ch - 1)
    ax, ay = qx & 255, qy & 255
    co = np.asarray(coff, np.int64)
    cands = []
    for src in range(2):
        comp = []
        for d in range(2):
            e = 2 * src + d
            o00, o01, o10, o11 = co[y0, x0, e], co[y0, x1, e], co[y1, x0, e], co[y1, x1, e]
            v = (o00 * (256 - ax) + o01 * ax) * (256 - ay) + (o10 * (256 - ax) + o11 * ax) * ay
            comp.append([(v + (1 << 23)) >> 24, (o00 + 128) >> 8, (o01 + 128) >> 8, (o10 + 128) >> 8,
                         (o11 + 128) >> 8])
        for k in range(5):
            cands.append((comp[0][k], comp[1][k]))
    for cx, cy in cands:
        inside = (xx + cx >= 0) & (xx + cx < sx) & (yy + cy >= 0) & (yy + cy < sy)
        po = _pack(cx, cy)
        tx = np.where(inside, xx + cx, xx)
        ty = np.where(inside, yy + cy, yy)
        c = _popcount_rows(F1 ^ F2[ty, tx])
        take = inside & (po != bo) & (c < best)
        best = np.where(take, c, best)
        bo = np.where(take, po, bo)
    return bo.astype(np.uint32), best.astype(np.uint16)


def philox_np(c0, c1, c2, c3, k0, k1):
    """Philox4x32-10 (Salmon et al., SC'11) on arrays of counters: ten rounds
    of the two 32x32 -> 64-bit products with the published multipliers and
    Weyl key increments; the element-wise equal of oracle.philox (checked
    against it on random counters in tests/test_net_oracle.py)."""
    M0, M1 = np.uint64(0xD2511F53), np.uint64(0xCD9E8D57)
    W0, W1 = 0x9E3779B9, 0xBB67AE85
    mask = np.uint64(0xFFFFFFFF)
    c = [np.asarray(v, np.uint64) & mask for v in (c0, c1, c2, c3)]
    c = np.broadcast_arrays(*c)
    c0, c1, c2, c3 = (v.copy() for v in c)
    k0, k1 = int(k0) & 0xFFFFFFFF, int(k1) & 0xFFFFFFFF
    for _ in range(10):
        p0 = M0 * c0
        p1 = M1 * c2
        n0 = (p1 >> np.uint64(32)) ^ c1 ^ np.uint64(k0)
        n2 = (p0 >> np.uint64(32)) ^ c3 ^ np.uint64(k1)
        c1 = p1 & mask
        c3 = p0 & mask
        c0, c2 = n0, n2
        k0 = (k0 + W0) & 0xFFFFFFFF
        k1 = (k1 + W1) & 0xFFFFFFFF
    return tuple(v.astype(np.uint32) for v in (c0, c1, c2, c3))


def local_search(F1, F2, off, cost, radius, samples, iteration, pair, seed):
    """Local search around each pixel's offset (the learned path's refinement
    of its seeds); returns (off, cost).  Sample s = 0 .. samples - 1: r =
    max(radius >> s, 1); Philox4x32-10(ctr = {pixel, iteration, pair, 1 + s},
    key = (seed low, seed high)) words 0 and 1 give d = (word * (2r + 1)) >>
    32 - r; the candidate is the running best plus (dx, dy), clamped to the
    frame per axis; a repeat of the running best is skipped, a strictly lower
    Hamming cost is taken."""
    sy, sx = F1.shape[:2]
    yy, xx = np.mgrid[0:sy, 0:sx].astype(np.int64)
    idx = (yy * sx + xx).astype(np.uint64)
    bo = np.asarray(off, np.uint32).copy()
    best = np.asarray(cost, np.int64).copy()
    k0, k1 = int(seed) & 0xFFFFFFFF, (int(seed) >> 32) & 0xFFFFFFFF
    for s in range(samples):
        r = max(radius >> s, 1)
        w0, w1, _, _ = philox_np(idx, iteration, pair, 1 + s, k0, k1)
        n = np.uint64(2 * r + 1)
        dx = ((w0.astype(np.uint64) * n) >> np.uint64(32)).astype(np.int64) - r
        dy = ((w1.astype(np.uint64) * n) >> np.uint64(32)).astype(np.int64) - r
        bx, by = _unpack(bo)
        nx = np.clip(bx + dx, -xx, sx - 1 - xx)
        ny = np.clip(by + dy, -yy, sy - 1 - yy)
        po = _pack(nx, ny)
        c = _popcount_rows(F1 ^ F2[yy + ny, xx + nx])
        take = (po != bo) & (c < best)
        best = np.where(take, c, best)
        bo = np.where(take, po, bo)
    return bo.astype(np.uint32), best.astype(np.uint16)


# ---------------------------------------------------------------------------
# the network

def split_params(w):
    """Flat parameter vector -> [(W [9, cout, cin], b [cout]) per layer]."""
    w = np.asarray(w)
    out, o = [], 0
    for ci, co, _, _ in LAYERS:
        Wl = w[o:o + 9 * co * ci].reshape(9, co, ci)
        o += 9 * co * ci
        out.append((Wl, w[o:o + co]))
        o += co
    assert o == N_PARAMS
    return out


def _shift(x, oy, ox):
    """y[:, r, c] = x[:, r + oy, c + ox] (zero outside)."""
    n, H, W, C = x.shape
    y = np.zeros_like(x)
    r0, r1 = max(0, -oy), min(H, H - oy)
    c0, c1 = max(0, -ox), min(W, W - ox)
    if r0 < r1 and c0 < c1:
        y[:, r0:r1, c0:c1] = x[:, r0 + oy:r1 + oy, c0 + ox:c1 + ox]
    return y


def conv(x, Wl, b, dil):
    """3x3 conv, zero padding: y[p, co] = b[co] + sum_tap,ci W[tap, co, ci] x[p + off(tap), ci],
    off(tap) = ((tap // 3 - 1) dil, (tap % 3 - 1) dil)."""
    y = np.zeros(x.shape[:3] + (Wl.shape[1],), np.float64)
    for tap in range(9):
        y += _shift(x, (tap // 3 - 1) * dil, (tap % 3 - 1) * dil) @ Wl[tap].T
    return y + b


def conv_data_grad(dz, Wl, dil):
    """Gradient of conv with respect to its input."""
    dx = np.zeros(dz.shape[:3] + (Wl.shape[2],), np.float64)
    for tap in range(9):
        dx += _shift(dz @ Wl[tap], -(tap // 3 - 1) * dil, -(tap % 3 - 1) * dil)
    return dx


def conv_weight_grad(dz, x, dil):
    """Gradient of conv with respect to W [9, cout, cin] and b."""
    co, ci = dz.shape[-1], x.shape[-1]
    gW = np.zeros((9, co, ci), np.float64)
    dzf = dz.reshape(-1, co)
    for tap in range(9):
        xs = _shift(x, (tap // 3 - 1) * dil, (tap % 3 - 1) * dil).reshape(-1, ci)
        gW[tap] = dzf.T @ xs
    return gW, dzf.sum(axis=0)


def forward(w, feat, quant=True):
    """Network output [n, ch, cw, 16] (float64) and the activations.

    feat: BF16 bits [n, ch, cw, 128] or floats.  With quant, weights and the
    ReLU activations are rounded to BF16 where the device kernels store them.
    Output channels 0, 1 carry a fixed skip connection: 8 x feature channels
    123, 124 (the lowest-cost displacement) are added to them."""
    x = bf16_float(feat).astype(np.float64) if np.asarray(feat).dtype == np.uint16 else np.asarray(feat, np.float64)
    acts = [x]
    for l, ((Wl, b), (_, _, dil, relu)) in enumerate(zip(split_params(w), LAYERS)):
        if quant:
            Wl = bf16_round(Wl)
        z = conv(acts[-1], Wl, np.asarray(b, np.float64), dil)
        if relu:
            z = np.maximum(z, 0.0)
            if quant:
                z = bf16_round(z)
        acts.append(z)
    out = acts[-1].copy()
    out[..., :2] += 8.0 * x[..., SKIP:SKIP + 2]
    return out, acts


def loss_and_grad(w, feat, label, quant=True):
    """Mean absolute error of output channels 0, 1 against label (non-finite
    labels contribute nothing; the mean is over 2 x all pixels) and its
    gradient with respect to the flat parameter vector."""
    out, acts = forward(w, feat, quant)
    label = np.asarray(label, np.float64)
    npix = int(np.prod(out.shape[:3]))
    ok = np.isfinite(label)
    d = np.where(ok, out[..., :2] - np.where(ok, label, 0.0), 0.0)
    loss = np.abs(d).sum() / (2.0 * npix)
    inv = float(np.float32(1.0) / (np.float32(2.0) * np.float32(npix))) if quant else 1.0 / (2.0 * npix)
    dz = np.zeros(out.shape, np.float64)
    dz[..., :2] = np.sign(d) * inv
    if quant:
        dz = bf16_round(dz)
    params = split_params(w)
    grads = [None] * len(LAYERS)
    for l in range(len(LAYERS) - 1, -1, -1):
        Wl, _ = params[l]
        dil = LAYERS[l][2]
        gW, gb = conv_weight_grad(dz, acts[l], dil)
        grads[l] = (gW, gb)
        if l > 0:
            Wq = bf16_round(Wl) if quant else np.asarray(Wl, np.float64)
            dx = conv_data_grad(dz, Wq, dil) * (acts[l] > 0)
            dz = bf16_round(dx) if quant else dx
    flat = np.concatenate([np.concatenate([g.ravel(), b.ravel()]) for g, b in grads])
    return loss, flat
