"""Pins the CPU oracle (the checker of the HIP path) without trusting its own
formulas: published known-answer vectors, closed forms, invariants, special
cases that reduce to library routines, and brute force on tiny inputs.  Runs on
the CPU only."""
import json
import os
import shutil

import numpy as np
import pytest
from scipy.special import ndtri

import oracle

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def _unpack(off):
    return oracle.unpack_off(off)


def _pack(dx, dy):
    return (np.asarray(dx, np.int16).view(np.uint16).astype(np.uint32) |
            (np.asarray(dy, np.int16).view(np.uint16).astype(np.uint32) << 16))


def _popcount(a):
    return np.unpackbits(np.ascontiguousarray(a, np.uint32).view(np.uint8), axis=-1).sum(-1)


# --------------------------------------------------------------- Philox / table

def test_philox_known_answers():
    """Random123 (Salmon et al., SC'11) Philox4x32-10 known-answer vectors."""
    with open(os.path.join(GOLDEN, "philox_kat.json")) as f:
        kats = json.load(f)
    for k in kats:
        got = oracle.philox([int(c, 16) for c in k["ctr"]], [int(c, 16) for c in k["key"]])
        assert ["%08x" % v for v in got] == k["out"]


def test_ndtri_table_matches_scipy():
    path = os.path.join(os.path.dirname(GOLDEN), "..", "briefmatch_amd", "csrc", "bm_ndtri_q16.inc")
    txt = open(path).read().split("\n", 1)[1]
    q = np.array([int(v) for v in txt.replace("\n", "").split(",") if v.strip()], np.int64)
    assert q.size == 4096
    k = np.arange(4096)
    ref = np.rint(ndtri((k + 0.5) / 4096.0) * 65536.0).astype(np.int64)
    assert np.array_equal(q, ref)
    assert np.array_equal(q[::-1], -q)          # odd symmetry about 1/2
    assert (np.diff(q) > 0).all()               # strictly increasing


# --------------------------------------------------------------- parameters

def test_derived_params_fixture():
    """of_pipeline.cpp:178-190 derivations against the committed values."""
    with open(os.path.join(GOLDEN, "derived_params.json")) as f:
        fx = json.load(f)
    for case in fx:
        p = oracle.derive_params(case["W"], case["H"], case["u"], case["u"])
        for key in ("sx", "sy", "P", "fsize"):
            assert getattr(p, key) == case[key], (case["name"], key)
        assert abs(p.patchArea - case["patchArea"]) < 5e-3
        assert abs(p.sS - case["sS"]) < 5e-3
        s = int(np.ceil(2.5 * float(np.sqrt(np.float32(p.sS) * np.float32(p.sS)))))
        assert s == case["tri_half_window"]
        assert abs(oracle.lib().orc_rs_sigma(p.sx) - case["rs_sigma"]) < 1e-3


def test_derived_params_defaults():
    """Absent up-sampling factors select 2.740 / 3.052 (of_pipeline.cpp:178-179)."""
    p = oracle.derive_params(584, 388)
    assert p.sx == int(np.float32(2.740) * np.float32(584)) and p.sy == int(np.float32(3.052) * np.float32(388))
    assert p.N == 128 and p.iterations == 2 and p.jfMax == 256


# --------------------------------------------------------------- gray / bilinear

def test_gray_exhaustive():
    """All 2^24 RGB triples: (77R + 150G + 29B + 128) >> 8 is within 1 of the
    BT.601 luma 0.299R + 0.587G + 0.114B and equals the rounded 8-bit weighted sum."""
    v = np.arange(1 << 24, dtype=np.uint32)
    rgb = np.stack([(v >> 16) & 255, (v >> 8) & 255, v & 255], -1).astype(np.uint8).reshape(4096, 4096, 3)
    g = oracle.gray(rgb).reshape(-1).astype(np.int64)
    r, gg, b = [rgb.reshape(-1, 3)[:, i].astype(np.int64) for i in range(3)]
    assert np.array_equal(g, np.floor((77 * r + 150 * gg + 29 * b) / 256.0 + 0.5).astype(np.int64))
    luma = 0.299 * r + 0.587 * gg + 0.114 * b
    assert np.abs(g - luma).max() <= 1.0


def test_upsample_properties():
    rng = np.random.default_rng(0)
    # constant image -> constant
    g = np.full((13, 17), 77, np.uint8)
    assert (oracle.upsample(g, 40, 31) == 77.0).all()
    # identity at factor 1
    g = rng.integers(0, 256, (13, 17)).astype(np.uint8)
    assert np.array_equal(oracle.upsample(g, 17, 13), g.astype(np.float32))
    # factor 2: even samples are the input
    L = oracle.upsample(g, 34, 26)
    assert np.array_equal(L[::2, ::2], g.astype(np.float32))
    # linear ramps are reproduced exactly at the quantised positions
    W, H, sx, sy = 23, 19, 61, 53
    xx, yy = np.meshgrid(np.arange(W), np.arange(H))
    g = (3 * xx + 5 * yy + 7).astype(np.uint8)
    L = oracle.upsample(g, sx, sy)
    X = (np.arange(sx, dtype=np.int64) * W * 256) // sx
    Y = (np.arange(sy, dtype=np.int64) * H * 256) // sy
    ex = 3 * np.minimum(X / 256.0, W - 1) + 5 * np.minimum(Y / 256.0, H - 1)[:, None] + 7
    # positions past the last column are clamped: the ramp holds where x0 + 1 < W
    inner = (X // 256 < W - 1)[None, :] & (Y // 256 < H - 1)[:, None]
    assert np.array_equal(L[inner], ex[inner].astype(np.float32))


def _lanczos_formula(g, sx, sy):
    """The reference's 36-tap sum v / w (util_cuda.cu:126-158) in float64,
    non-separable, with the source coordinates computed in float as there."""
    H, W = g.shape
    out = np.zeros((sy, sx))

    def lz(t):
        return 1.0 if t == 0 else 3.0 * np.sin(t) * np.sin(t / 3.0) / (t * t)
    for y in range(sy):
        yp = float(np.float32(y) * np.float32(H) / np.float32(sy))
        y1 = int(np.floor(yp))
        for x in range(sx):
            xp = float(np.float32(x) * np.float32(W) / np.float32(sx))
            x1 = int(np.floor(xp))
            v = w = 0.0
            for i in range(x1 - 2, x1 + 4):
                Lx = lz(np.pi * (xp - i))
                for j in range(y1 - 2, y1 + 4):
                    Ly = lz(np.pi * (yp - j))
                    w += Lx * Ly
                    v += float(g[min(max(j, 0), H - 1), min(max(i, 0), W - 1)]) * Lx * Ly
            out[y, x] = v / w
    return out


@pytest.mark.parametrize("W,H,sx,sy", [(9, 7, 23, 17), (12, 10, 12, 29), (6, 5, 35, 5)])
def test_lanczos_matches_reference_formula(W, H, sx, sy):
    g = np.random.default_rng(W).integers(0, 256, (H, W)).astype(np.uint8)
    got = oracle.upsample_lanczos(g, sx, sy)
    assert np.abs(got - _lanczos_formula(g, sx, sy)).max() <= 2e-4


def test_lanczos_tables_and_properties():
    first, w = oracle.lanczos_table(17, 49)
    xp = np.arange(49, dtype=np.float32) * np.float32(17) / np.float32(49)
    assert np.array_equal(first, np.floor(xp).astype(np.int32) - 2)
    assert np.abs(w.astype(np.float64).sum(1) - 1.0).max() <= 1e-6
    rng = np.random.default_rng(2)
    g = np.full((13, 17), 77, np.uint8)
    assert np.abs(oracle.upsample_lanczos(g, 40, 31) - 77.0).max() <= 1e-4
    g = rng.integers(0, 256, (13, 17)).astype(np.uint8)
    assert np.abs(oracle.upsample_lanczos(g, 17, 13) - g).max() <= 1e-4        # factor 1: interpolating
    assert np.abs(oracle.upsample_lanczos(g, 34, 26)[::2, ::2] - g).max() <= 1e-4


# --------------------------------------------------------------- BRIEF

def _tiny_image(W=11, H=9, seed=0):
    rng = np.random.default_rng(seed)
    return rng.permutation(W * H).reshape(H, W).astype(np.float32)  # distinct values: no ties


def test_brief_bit_order_and_clamp():
    L = _tiny_image()
    H, W = L.shape
    N = 37  # two words, the second partial
    for k in (0, 5, 31, 32, 36):
        pat = np.zeros((N, 4), np.int16)                       # o1 == o2: bit 0
        pat[k] = (3, -2, -4, 1)
        F = oracle.brief(L, pat.reshape(-1), N)
        xx, yy = np.meshgrid(np.arange(W), np.arange(H))
        a = L[np.clip(yy - 2, 0, H - 1), np.clip(xx + 3, 0, W - 1)]
        b = L[np.clip(yy + 1, 0, H - 1), np.clip(xx - 4, 0, W - 1)]
        expect = (a > b).astype(np.uint32) << np.uint32(k % 32)
        assert np.array_equal(F[..., k // 32], expect)
        assert (F[..., 1 - k // 32] == 0).all()
        assert (F[..., 1] >> np.uint32(N - 32) == 0).all()    # bits >= N stay 0


def test_brief_constant_and_swap():
    rng = np.random.default_rng(1)
    N = 64
    pat = rng.integers(-4, 5, (N, 4)).astype(np.int16)
    assert (oracle.brief(np.full((9, 12), 3.0, np.float32), pat.reshape(-1), N) == 0).all()
    L = _tiny_image(12, 9, 2)
    F = oracle.brief(L, pat.reshape(-1), N)
    sw = pat[:, [2, 3, 0, 1]].copy()
    Fs = oracle.brief(L, sw.reshape(-1), N)
    H, W = L.shape
    xx, yy = np.meshgrid(np.arange(W), np.arange(H))
    for s in range(N):
        same = (np.clip(xx + pat[s, 0], 0, W - 1) == np.clip(xx + pat[s, 2], 0, W - 1)) & \
               (np.clip(yy + pat[s, 1], 0, H - 1) == np.clip(yy + pat[s, 3], 0, H - 1))
        b1 = (F[..., s // 32] >> np.uint32(s % 32)) & 1
        b2 = (Fs[..., s // 32] >> np.uint32(s % 32)) & 1
        assert np.array_equal(b1[~same] ^ b2[~same], np.ones((~same).sum(), np.uint32))
        assert (b1[same] == 0).all() and (b2[same] == 0).all()


def test_pattern_law():
    """Seeded pattern: sign(v) v^2 with v ~ N(0, 0.2), clamped, times patchArea,
    truncated (briefmatch.cpp:124-148)."""
    area = 10000.0
    p1 = oracle.make_pattern(123, 512, area).astype(np.float64)
    assert np.array_equal(p1, oracle.make_pattern(123, 512, area))
    assert not np.array_equal(p1, oracle.make_pattern(124, 512, area))
    val = p1 / area
    # E|val| = E[v^2] = 0.04; P(|val| > 0.154) = P(|z| > 1.962) = 0.0498
    assert abs(np.abs(val).mean() - 0.04) < 0.004
    assert abs((np.abs(val) > 0.154).mean() - 0.0498) < 0.015
    assert abs(np.sign(val).mean()) < 0.06
    assert np.abs(val).max() <= 1.0
    # truncation toward zero: |offset| <= patchArea * max|v|^2, mostly 0 at small areas
    small = oracle.make_pattern(123, 512, 2.5)
    assert np.abs(small).max() <= 1 and (small == 0).mean() > 0.9


# --------------------------------------------------------------- matching

def _rand_desc(sy, sx, NS, seed):
    return np.random.default_rng(seed).integers(0, 1 << 32, (sy, sx, NS), dtype=np.uint64).astype(np.uint32)


def test_init_cost_is_hamming():
    F1, F2 = _rand_desc(7, 9, 3, 0), _rand_desc(7, 9, 3, 1)
    off, cost = oracle.init(F1, F2, first=True)
    assert (off == 0).all()
    assert np.array_equal(cost, _popcount(F1 ^ F2))


def _ref_accepts(a, b, N, words, div_ulps=0):
    """The reference's candidate test (featDist + insert, briefmatch_cuda.cu:60-80,
    131-144) in float32: current distance D = b / N, early exit once a partial
    Hamming sum exceeds N * D, accept iff a / N < D.  `words` are the per-word
    popcounts summing to a (the order the partial sums arrive in); div_ulps
    moves every quotient by that many ulps (fast-math division is within 2 ulps),
    a < 0 towards zero for the candidate and away for D, > 0 the opposite.
    Equal costs keep equal quotients: the division is a function of its inputs."""
    f32 = np.float32
    D = f32(b) / f32(N)
    q = f32(a) / f32(N)
    for _ in range(abs(div_ulps) if a != b else 0):
        q = np.nextafter(q, f32(np.inf) if div_ulps > 0 else f32(0))
        D = np.nextafter(D, f32(0) if div_ulps > 0 else f32(np.inf))
    ddmax = f32(N) * D
    dist = 0
    for w in words:
        dist += int(w)
        if f32(dist) > ddmax:
            return False
    return bool(q < D)


def test_float_distance_rule_equals_integer_strict_compare():
    """The integer rule of this build (accept a candidate of Hamming cost a iff
    a < b, the current cost) accepts exactly what the reference's float rule
    accepts, for every descriptor length N in [2, 512] and every pair of costs
    0 <= a, b <= N, with the partial sums arriving in the worst order for the
    early exit (all of a in the first word) and spread over words, and with
    each quotient moved by up to 2 ulps against the comparison (fast-math
    division)."""
    f32 = np.float32
    for N in range(2, 513):
        a = np.arange(N + 1)
        b = np.arange(N + 1)
        A, B = np.meshgrid(a, b, indexing="ij")
        D = b.astype(f32) / f32(N)
        Q = a.astype(f32) / f32(N)
        ddmax = f32(N) * D
        # early exit never rejects a cheaper candidate: every partial sum of a
        # cost a < b is at most a, and a <= N * D
        cheaper = A < B
        assert (A[cheaper].astype(f32) <= ddmax[B[cheaper]]).all(), N
        # the quotients order exactly like the integers, even 2 ulps apart
        lo, hi = Q.copy(), D.copy()
        for _ in range(2):
            lo = np.nextafter(lo, f32(np.inf))   # candidate quotient rounded up
            hi = np.nextafter(hi, f32(0))        # current quotient rounded down
        assert (lo[:-1] < hi[1:]).all(), N      # a = b - 1 still accepted
        acc = (Q[A] < D[B]) & (A.astype(f32) <= ddmax[B])
        assert np.array_equal(acc, cheaper), N  # a == b and a > b rejected
    # spot checks through the sequential form with spread partial sums
    rng = np.random.default_rng(6)
    for N in (2, 3, 31, 32, 33, 100, 128, 255, 512):
        NS = (N + 31) // 32
        for _ in range(200):
            a, b = int(rng.integers(0, N + 1)), int(rng.integers(0, N + 1))
            cuts = np.sort(rng.integers(0, a + 1, NS - 1)) if NS > 1 else np.array([], int)
            words = np.diff(np.concatenate([[0], cuts, [a]]))
            for ulps in (0, -2, 2):
                assert _ref_accepts(a, b, N, words, ulps) == (a < b), (N, a, b, ulps)


def test_level_selection_under_the_float_rule():
    """One oracle propagation level on random descriptors, replayed pixel by
    pixel with the reference's float distance and early exit (sequential
    insert over the four candidates of the level-start snapshot, equal offsets
    skipped): every pixel ends with the oracle's offset and cost."""
    sy, sx, N = 20, 24, 72
    NS = (N + 31) // 32
    rng = np.random.default_rng(8)
    F1 = _rand_desc(sy, sx, NS, 20)
    F2 = _rand_desc(sy, sx, NS, 21)
    last = (1 << (N - 32 * (NS - 1))) - 1
    F1[..., -1] &= np.uint32(last)
    F2[..., -1] &= np.uint32(last)
    xx, yy = np.meshgrid(np.arange(sx), np.arange(sy))
    dx = np.clip(rng.integers(-6, 7, (sy, sx)), -xx, sx - 1 - xx).astype(np.int16)
    dy = np.clip(rng.integers(-6, 7, (sy, sx)), -yy, sy - 1 - yy).astype(np.int16)
    off0 = _pack(dx, dy)
    cost0 = _popcount(F1 ^ F2[yy + dy, xx + dx]).astype(np.uint16)
    f32 = np.float32
    for p, pd in ((4, -1), (1, 1), (8, 1)):
        eo, ec = oracle.propagate_level(F1, F2, p, pd, True, off0, cost0)
        edx, edy = _unpack(eo)
        for y in range(sy):
            for x in range(sx):
                idx = y * sx + x
                bx, by, D = int(dx[y, x]), int(dy[y, x]), f32(int(cost0[y, x])) / f32(N)
                best = int(cost0[y, x])
                for d in (-1, 1):
                    for xy in (0, 1):
                        n = idx + pd * p * d * (1 if xy else sx)
                        if n < 0 or n >= sx * sy:
                            continue
                        nx = min(max(int(dx.flat[n]), -x), sx - 1 - x)
                        ny = min(max(int(dy.flat[n]), -y), sy - 1 - y)
                        if (nx, ny) == (bx, by):
                            continue
                        words = [bin(int(v)).count("1") for v in F1[y, x] ^ F2[y + ny, x + nx]]
                        a = sum(words)
                        if _ref_accepts(a, best, N, words) and f32(a) / f32(N) < D:
                            bx, by, D, best = nx, ny, f32(a) / f32(N), a
                assert (edx[y, x], edy[y, x], ec[y, x]) == (bx, by, best), (p, pd, x, y)


def test_temporal_init_follows_previous_field():
    sy, sx = 6, 8
    F1, F2 = _rand_desc(sy, sx, 1, 2), _rand_desc(sy, sx, 1, 3)
    dx = np.zeros((sy, sx), np.int16)
    dy = np.zeros((sy, sx), np.int16)
    dx[2, 3], dy[2, 3] = 2, 1          # pixel (3,2) points at (5,3)
    dx[3, 5], dy[3, 5] = -1, 2         # whose offset is (-1, 2)
    off, cost = oracle.init(F1, F2, first=False, prev_off=_pack(dx, dy))
    odx, ody = _unpack(off)
    assert (odx[2, 3], ody[2, 3]) == (-1, 2)
    assert cost[2, 3] == _popcount(F1[2, 3] ^ F2[4, 2])


def test_level_candidate_order_and_snapshot():
    """Four equal-cost candidates: the reference order (d = -1 first, vertical
    before horizontal, neighbour index idx + pd*p*d*(xy ? 1 : sx)) picks the first."""
    sy, sx, NS, p = 9, 9, 1, 2
    F1 = np.zeros((sy, sx, NS), np.uint32)
    F2 = np.full((sy, sx, NS), 0xFFFFFFFF, np.uint32)
    y0, x0 = 4, 4
    F1[y0, x0] = 0
    # targets of the four candidate offsets, all at Hamming distance 1
    offs = {(0, 1): (1, 1), (0, -1): (-1, 1), (1, 0): (1, -1), (-1, 0): (-1, -1)}
    dx = np.zeros((sy, sx), np.int16)
    dy = np.zeros((sy, sx), np.int16)
    for (ny, nx), (ox, oy) in offs.items():
        dy[y0 + p * ny, x0 + p * nx], dx[y0 + p * ny, x0 + p * nx] = oy, ox
        F2[y0 + oy, x0 + ox] = 1
    off = _pack(dx, dy)
    cost = np.full((sy, sx), 20, np.uint16)
    for pd, first in ((-1, (1, 0)), (1, (-1, 0))):
        o2, c2 = oracle.propagate_level(F1, F2, p, pd, True, off, cost)
        ddx, ddy = _unpack(o2)
        assert c2[y0, x0] == 1
        assert (ddx[y0, x0], ddy[y0, x0]) == offs[first], pd
        # without doubleProp only d = +1 candidates exist: vertical at idx + pd*p*sx
        o3, _ = oracle.propagate_level(F1, F2, p, pd, False, off, cost)
        ddx, ddy = _unpack(o3)
        assert (ddx[y0, x0], ddy[y0, x0]) == offs[(-first[0], 0)], pd
    # snapshot: a pixel reads neighbours' level-start offsets, never this level's
    o2, _ = oracle.propagate_level(F1, F2, p, -1, True, off, cost)
    o4, _ = oracle.propagate_level(F1, F2, p, -1, True, off, cost)
    assert np.array_equal(o2, o4)


def test_match_translation_known_answer():
    """Unique descriptors shifted by an integer translation: the interior field
    converges to that translation with zero cost."""
    sy, sx, NS = 60, 80, 4
    tx, ty = 7, -5
    F1 = _rand_desc(sy, sx, NS, 10)
    F2 = _rand_desc(sy, sx, NS, 11)
    ys, xs = np.mgrid[0:sy, 0:sx]
    ok = (xs + tx >= 0) & (xs + tx < sx) & (ys + ty >= 0) & (ys + ty < sy)
    F2[(ys + ty)[ok], (xs + tx)[ok]] = F1[ok]
    off, cost = oracle.match(F1, F2, iterations=4, jf_max=64, seed=3)
    dx, dy = _unpack(off)
    hit = (dx == tx) & (dy == ty)
    assert hit[ok].mean() > 0.99
    assert (cost[hit] == 0).all()
    assert np.array_equal(cost == 0, hit)           # zero cost only at the true offset


def test_match_invariants_and_monotone_levels():
    sy, sx, NS = 24, 30, 2
    F1, F2 = _rand_desc(sy, sx, NS, 20), _rand_desc(sy, sx, NS, 21)
    off, cost = oracle.init(F1, F2, first=True)
    off, cost = oracle.random_search(F1, F2, 0, 0, 9, off, cost)
    for p in (16, 8, 4, 2, 1):
        o2, c2 = oracle.propagate_level(F1, F2, p, -1, True, off, cost)
        assert (c2 <= cost).all()
        off, cost = o2, c2
    dx, dy = _unpack(off)
    ys, xs = np.mgrid[0:sy, 0:sx]
    tx, ty = xs + dx, ys + dy
    assert (tx >= 0).all() and (tx < sx).all() and (ty >= 0).all() and (ty < sy).all()
    assert np.array_equal(cost, _popcount(F1 ^ F2[ty, tx]))
    # exhaustive minimum is a lower bound
    best = np.full((sy, sx), 10 ** 9)
    for oy in range(-sy + 1, sy):
        for ox in range(-sx + 1, sx):
            vy, vx = ys + oy, xs + ox
            v = (vy >= 0) & (vy < sy) & (vx >= 0) & (vx < sx)
            c = np.full((sy, sx), 10 ** 9)
            c[v] = _popcount(F1[v] ^ F2[vy[v], vx[v]])
            best = np.minimum(best, c)
    assert (cost >= best).all()


def test_random_search_distribution():
    """Offsets are round(sigma z), z ~ N(0, 1) via the table; sigma = sqrt(ceil(sx/6))."""
    sy, sx, NS = 200, 600, 1
    F1 = np.zeros((sy, sx, NS), np.uint32)
    F2 = np.zeros((sy, sx, NS), np.uint32)
    off = np.zeros((sy, sx), np.uint32)
    cost = np.full((sy, sx), 65535, np.uint16)     # every new candidate is accepted
    o2, c2 = oracle.random_search(F1, F2, 1, 0, 77, off, cost)
    dx, dy = _unpack(o2)
    sigma = np.sqrt(np.ceil(sx * 0.5 / 3.0))
    inner = (slice(40, 160), slice(60, 540))
    for d in (dx[inner].astype(np.float64), dy[inner].astype(np.float64)):
        se = sigma / np.sqrt(d.size)                # standard error of the mean
        assert abs(d.mean()) < 4 * se
        assert abs(d.std() / sigma - 1.0) < 0.02
    # different iteration or pair -> different stream
    o3, _ = oracle.random_search(F1, F2, 2, 0, 77, off, cost)
    o4, _ = oracle.random_search(F1, F2, 1, 1, 77, off, cost)
    assert (o3 != o2).mean() > 0.9 and (o4 != o2).mean() > 0.9


# --------------------------------------------------------------- refinement

def test_gauss_kernel():
    for sigma in (0.15, 1.597, 3.0):
        k = oracle.gauss_kernel(sigma).astype(np.float64)
        assert len(k) == int(np.ceil(np.float32(sigma) * 3) * 2 + 1)
        assert abs(k.sum() - 1.0) < 1e-6
        assert np.allclose(k, k[::-1])
        r = np.arange(len(k)) - len(k) // 2
        e = np.exp(-r * r / (2.0 * sigma * sigma))
        assert np.allclose(k, e / e.sum(), rtol=1e-5)


def test_downsample_closed_forms():
    sx, sy, W, H = 87, 64, 30, 22
    k = oracle.gauss_kernel(1.3)
    P2 = len(k) // 2
    # constant field -> constant times the scale
    off = _pack(np.full((sy, sx), 5), np.full((sy, sx), -3))
    U, V = oracle.downsample(off, W, H, k)
    assert np.allclose(U, 5 * np.float32(W) / np.float32(sx), atol=1e-5)
    assert np.allclose(V, -3 * np.float32(H) / np.float32(sy), atol=1e-5)
    # a ramp dx = x survives the symmetric Gaussian away from the borders, so U
    # samples it at the nearest-neighbour column round(x sx / W)
    xx = np.broadcast_to(np.arange(sx), (sy, sx))
    U, _ = oracle.downsample(_pack(xx, np.zeros((sy, sx))), W, H, k)
    x1 = np.floor(np.arange(W) * sx / W + 0.5)
    inner = (x1 >= P2) & (x1 < sx - P2)
    assert np.allclose(U[5, inner], x1[inner] * np.float32(W) / np.float32(sx), atol=1e-4)
    # identity when not up-sampled
    off = _pack(np.arange(W * H).reshape(H, W) % 7 - 3, np.arange(W * H).reshape(H, W) % 5 - 2)
    U, V = oracle.downsample(off, W, H, k)
    dx, dy = _unpack(off)
    assert np.array_equal(U, dx.astype(np.float32)) and np.array_equal(V, dy.astype(np.float32))


@pytest.mark.parametrize("fsize", [1, 5, 23, 47])
def test_median_matches_numpy(fsize):
    rng = np.random.default_rng(fsize)
    a = rng.normal(size=(31, 53)).astype(np.float32)
    ks = fsize // 2
    for d in (0, 1):
        got = oracle.median1d(a, d, fsize)
        if d:
            idx = np.clip(np.arange(a.shape[1])[:, None] + np.arange(-ks, ks + 1)[None, :], 0, a.shape[1] - 1)
            ref = np.median(a[:, idx], axis=-1)
        else:
            idx = np.clip(np.arange(a.shape[0])[:, None] + np.arange(-ks, ks + 1)[None, :], 0, a.shape[0] - 1)
            ref = np.median(a[idx, :], axis=1)
        assert np.array_equal(got, ref.astype(np.float32))


def _trilateral_f64(Ui, Vi, Um, Vm, I, sE, sI, sS):
    H, W = Ui.shape
    s = int(np.ceil(2.5 * float(np.sqrt(np.float32(sS)))))
    out_u = np.zeros((H, W))
    out_v = np.zeros((H, W))
    r = np.arange(-s, s)
    for y in range(H):
        for x in range(W):
            yy = np.clip(y + r, 0, H - 1)[None, :]   # j (rows) inner
            xx = np.clip(x + r, 0, W - 1)[:, None]   # i (cols) outer
            u1, v1, i1 = Ui[yy, xx].astype(np.float64), Vi[yy, xx].astype(np.float64), I[yy, xx] / 255.0
            w = (np.exp(-((u1 - Um[y, x]) ** 2 + (v1 - Vm[y, x]) ** 2) / (2 * sE)) *
                 np.exp(-(i1 - I[y, x] / 255.0) ** 2 / (2 * sI)) *
                 np.exp(-(r[:, None] ** 2 + r[None, :] ** 2) / (2 * sS)))
            den = max(1e-5, w.sum())
            out_u[y, x] = (w * u1).sum() / den
            out_v[y, x] = (w * v1).sum() / den
    return out_u, out_v


def test_trilateral_against_double_and_limits():
    rng = np.random.default_rng(5)
    H, W = 10, 13
    Ui = rng.normal(0, 2, (H, W)).astype(np.float32)
    Vi = rng.normal(0, 2, (H, W)).astype(np.float32)
    Um = (Ui + rng.normal(0, 0.3, (H, W))).astype(np.float32)
    Vm = (Vi + rng.normal(0, 0.3, (H, W))).astype(np.float32)
    I = rng.integers(0, 256, (H, W)).astype(np.float32)
    sE, sI, sS = 0.4 ** 2, 0.1 ** 2, 2.2 ** 2
    u, v = oracle.trilateral(Ui, Vi, Um, Vm, I, sE, sI, sS)
    ru, rv = _trilateral_f64(Ui, Vi, Um, Vm, I, sE, sI, sS)
    assert np.abs(u - ru).max() < 1e-4 and np.abs(v - rv).max() < 1e-4
    # constant flow is a fixed point
    c = np.full((H, W), 1.75, np.float32)
    u, v = oracle.trilateral(c, -c, c, -c, I, sE, sI, sS)
    assert np.allclose(u, 1.75, atol=1e-6) and np.allclose(v, -1.75, atol=1e-6)
    # huge range sigmas: a spatial Gaussian mean over the clamped window
    u, _ = oracle.trilateral(Ui, Vi, Um, Vm, I, 1e12, 1e12, sS)
    s = int(np.ceil(2.5 * float(np.sqrt(np.float32(sS)))))
    r = np.arange(-s, s)
    g = np.exp(-r * r / (2 * sS))
    y, x = 4, 6
    wgt = g[:, None] * g[None, :]
    win = Ui[np.clip(y + r, 0, H - 1)[:, None], np.clip(x + r, 0, W - 1)[None, :]]
    assert abs(u[y, x] - (wgt * win).sum() / wgt.sum()) < 1e-5


def test_epe_closed_form():
    H, W = 5, 7
    gt = np.zeros((H, W, 2), np.float32)
    U = np.full((H, W), 3.0, np.float32)
    V = np.full((H, W), 4.0, np.float32)
    occ = np.zeros((H, W), np.uint8)
    gt[0, 0] = (1e10, 0)          # unknown flow is skipped
    gt[0, 1] = (np.nan, 0)
    occ[1, :] = 1                 # occluded rows only leave EPE-noc
    U[1, :] = 103.0
    a, n = oracle.epe(U, V, gt, occ)
    known = H * W - 2
    assert abs(a - (5.0 * (known - W) + W * np.hypot(103, 4)) / known) < 1e-9
    assert abs(n - 5.0) < 1e-12


# --------------------------------------------------------------- golden fixtures

def test_golden_stage_vectors():
    """The committed per-stage vectors (tests/golden/make_golden.py) are reproduced."""
    data = np.load(os.path.join(GOLDEN, "stages_small.npz"))
    for tag in sorted({k.split("/")[0] for k in data.files}):
        d = {k.split("/", 1)[1]: data[k] for k in data.files if k.startswith(tag + "/")}
        prm = json.loads(str(d["params"]))
        p = oracle.Params()
        for k, v in prm.items():
            setattr(p, k, v)
        seed = int(d["seed"])
        assert np.array_equal(oracle.make_pattern(seed, p.N, p.patchArea), d["pattern"])
        g1 = oracle.gray(d["frame1"])
        assert np.array_equal(g1, d["gray1"])
        L1 = oracle.upsample(g1, p.sx, p.sy)
        assert np.array_equal(L1, d["L1"])
        F1 = oracle.brief(L1, d["pattern"], p.N)
        assert np.array_equal(F1, d["F1"])
        off, cost, U, V = oracle.flow_pair(p, d["frame1"], d["frame2"], seed=seed, pair=int(d["pair"]))
        assert np.array_equal(off, d["off"]) and np.array_equal(cost, d["cost"])
        assert np.array_equal(U, d["U"]) and np.array_equal(V, d["V"])


def test_median_networks(tmp_path):
    """The kernels' pruned median-selection networks (bm_median_net.h) select
    the middle element: exhaustive on 0/1 inputs up to 19 taps, random floats
    with ties up to 47 taps, checked by a host build of the same header."""
    import shutil
    import subprocess
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("no host C++ compiler")
    here = os.path.dirname(os.path.abspath(__file__))
    csrc = os.path.join(os.path.dirname(here), "briefmatch_amd", "csrc")
    exe = str(tmp_path / "median_net_check")
    subprocess.run([gxx, "-O2", "-std=c++17", "-I" + csrc, os.path.join(here, "median_net_check.cpp"), "-o", exe],
                   check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("mismatches 0") == 24


@pytest.mark.skipif(shutil.which("make") is None or shutil.which("g++") is None, reason="no host toolchain")
def test_oracle_under_sanitizers():
    """The oracle under AddressSanitizer and UBSan (oracle/san_main.cpp): the
    whole path with both up-samplers on odd frame sizes and descriptor lengths
    2..512, the stages alone (temporal init, levels in both directions, random
    search), median windows wider than the frame, partial trilateral row
    ranges and EPE with unknown ground truth; any report fails the run."""
    import subprocess
    odir = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")
    b = subprocess.run(["make", "-C", odir, "san"], capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stdout + b.stderr
    r = subprocess.run([os.path.join(odir, "oracle_san")], capture_output=True, text=True, timeout=600,
                       env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1"))
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count(" ok") == 6
