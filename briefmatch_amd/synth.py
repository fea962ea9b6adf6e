"""Seeded synthetic frame pairs with exact ground-truth flow.

Frames are RGB uint8 (H, W, 3).  Textures are sums of random Gaussian blobs plus
band-limited noise (white noise low-passed by a Gaussian).  Two motion models:

* ``"rects"``: a background textured layer moved by a small affine map around the
  image centre, and ``n_rects`` textured rectangles on top translating by up to
  ``max_motion`` pixels (later rectangles occlude earlier ones).
* ``"regions"``: one textured layer moved by a smooth flow field, a soft blend of
  ``n_regions`` random affine motions, scaled so max |flow| = ``max_motion``.

Ground truth is the forward flow on the frame-1 grid (frame 1 -> frame 2), the
direction BriefMatch estimates.  ``occ`` flags frame-1 pixels whose destination is
covered by a higher layer in frame 2 or leaves the frame.  Both frames get
independent N(0, noise_sigma^2) noise, then are rounded and clipped.

All randomness comes from ``numpy.random.default_rng(seed)``, so a seed fixes the
pair exactly on any machine with this numpy/scipy.
"""
from dataclasses import dataclass

import numpy as np
from scipy import ndimage


@dataclass
class Pair:
    frame1: np.ndarray  # (H, W, 3) uint8
    frame2: np.ndarray  # (H, W, 3) uint8
    gt: np.ndarray      # (H, W, 2) float32, forward flow on the frame-1 grid
    occ: np.ndarray     # (H, W) uint8, 1 = occluded / leaves the frame


def _texture(rng, h, w, n_blobs=None, noise_sigma=1.2, noise_amp=22.0, scale=None):
    """Blobs + band-limited noise, float32 (h, w, 3) roughly in [0, 255].
    Blob radii are relative to `scale` (default max(h, w))."""
    img = np.empty((h, w, 3), np.float32)
    img[:] = rng.uniform(60.0, 190.0, size=3).astype(np.float32)
    if n_blobs is None:
        n_blobs = int(np.clip(h * w / 6000.0, 24, 420))
    scale = float(max(h, w)) if scale is None else float(scale)
    for _ in range(n_blobs):
        cy, cx = rng.uniform(0, h), rng.uniform(0, w)
        sig = scale * rng.uniform(0.006, 0.05)
        amp = rng.uniform(-90.0, 90.0, size=3).astype(np.float32)
        r = int(3 * sig) + 1
        y0, y1 = max(0, int(cy) - r), min(h, int(cy) + r + 1)
        x0, x1 = max(0, int(cx) - r), min(w, int(cx) + r + 1)
        if y0 >= y1 or x0 >= x1:
            continue
        gy = np.exp(-((np.arange(y0, y1, dtype=np.float32) - cy) ** 2) / (2 * sig * sig))
        gx = np.exp(-((np.arange(x0, x1, dtype=np.float32) - cx) ** 2) / (2 * sig * sig))
        img[y0:y1, x0:x1, :] += (gy[:, None, None] * gx[None, :, None]) * amp[None, None, :]
    noise = rng.standard_normal((h, w, 3)).astype(np.float32)
    noise = ndimage.gaussian_filter(noise, sigma=(noise_sigma, noise_sigma, 0))
    noise *= noise_amp / max(float(noise.std()), 1e-6)
    img += noise
    return img


def _sample(tex, ys, xs, order=1):
    """Sample a (h, w, 3) texture at float coordinates (nearest at the border)."""
    out = np.empty(ys.shape + (3,), np.float32)
    for c in range(3):
        out[..., c] = ndimage.map_coordinates(tex[..., c], [ys, xs], order=order, mode="nearest")
    return out


def _finish(rng, f1, f2, noise_sigma):
    if noise_sigma > 0:
        f1 = f1 + rng.normal(0.0, noise_sigma, f1.shape).astype(np.float32)
        f2 = f2 + rng.normal(0.0, noise_sigma, f2.shape).astype(np.float32)
    return (np.clip(np.rint(f1), 0, 255).astype(np.uint8),
            np.clip(np.rint(f2), 0, 255).astype(np.uint8))


def _rects_pair(rng, W, H, n_rects, max_motion, noise_sigma, scene=None):
    # scene = (Ws, Hs): the frames are a W x H window of a Ws x Hs scene, whose
    # size sets the blob and rectangle sizes (default: the frame itself)
    Ws, Hs = (W, H) if scene is None else scene
    m = int(np.ceil(max_motion)) + 8
    bg = _texture(rng, H + 2 * m, W + 2 * m,
                  scale=None if scene is None else max(Hs, Ws) * max(H + 2 * m, W + 2 * m) / max(H, W))
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float32)
    cy, cx = (H - 1) / 2.0, (W - 1) / 2.0

    # background affine q = c + s R(theta) (p - c) + t, kept within max_motion / 2
    theta = rng.uniform(-1.0, 1.0) * 0.25 * max_motion / max(W, H)
    s = 1.0 + rng.uniform(-1.0, 1.0) * 0.25 * max_motion / max(W, H)
    t = rng.uniform(-1.0, 1.0, size=2) * 0.2 * max_motion
    A = s * np.array([[np.cos(theta), -np.sin(theta)], [np.sin(theta), np.cos(theta)]])
    bx = A[0, 0] * (xx - cx) + A[0, 1] * (yy - cy) + cx + t[0] - xx
    by = A[1, 0] * (xx - cx) + A[1, 1] * (yy - cy) + cy + t[1] - yy
    Ainv = np.linalg.inv(A)
    # frame 2 background: sample the frame-1 texture at A^-1 (q - c - t) + c
    qx, qy = xx - cx - t[0], yy - cy - t[1]
    px = Ainv[0, 0] * qx + Ainv[0, 1] * qy + cx
    py = Ainv[1, 0] * qx + Ainv[1, 1] * qy + cy
    f1 = bg[m:m + H, m:m + W].copy()
    f2 = _sample(bg, py + m, px + m)

    gtx, gty = bx.astype(np.float32), by.astype(np.float32)
    layer1 = np.zeros((H, W), np.int32)   # topmost layer index per pixel, frame 1
    layer2 = np.zeros((H, W), np.int32)   # same in frame 2
    for k in range(1, n_rects + 1):
        rw = min(int(rng.uniform(0.08, 0.28) * Ws), W - 1)
        rh = min(int(rng.uniform(0.08, 0.28) * Hs), H - 1)
        x0 = int(rng.uniform(0, W - rw))
        y0 = int(rng.uniform(0, H - rh))
        ang = rng.uniform(0, 2 * np.pi)
        mag = max_motion * np.sqrt(rng.uniform(0.05, 1.0))
        tx, ty = mag * np.cos(ang), mag * np.sin(ang)
        tex = _texture(rng, rh + 4, rw + 4)
        # frame 1
        f1[y0:y0 + rh, x0:x0 + rw] = tex[2:2 + rh, 2:2 + rw]
        gtx[y0:y0 + rh, x0:x0 + rw] = tx
        gty[y0:y0 + rh, x0:x0 + rw] = ty
        layer1[y0:y0 + rh, x0:x0 + rw] = k
        # frame 2: rectangle moved by (tx, ty), sub-pixel via texture sampling
        inside = (xx >= x0 + tx) & (xx < x0 + rw + tx) & (yy >= y0 + ty) & (yy < y0 + rh + ty)
        if inside.any():
            sy_ = yy[inside] - ty - y0 + 2
            sx_ = xx[inside] - tx - x0 + 2
            f2[inside] = _sample(tex, sy_, sx_)
            layer2[inside] = k

    # occlusion: destination outside the frame or under a higher layer in frame 2
    dx, dy = xx + gtx, yy + gty
    outside = (dx < 0) | (dx > W - 1) | (dy < 0) | (dy > H - 1)
    ix = np.clip(np.rint(dx), 0, W - 1).astype(np.int64)
    iy = np.clip(np.rint(dy), 0, H - 1).astype(np.int64)
    covered = layer2[iy, ix] > layer1
    occ = (outside | covered).astype(np.uint8)
    f1, f2 = _finish(rng, f1, f2, noise_sigma)
    gt = np.stack([gtx, gty], axis=-1).astype(np.float32)
    return Pair(f1, f2, gt, occ)


def _regions_pair(rng, W, H, n_regions, max_motion, noise_sigma):
    m = int(np.ceil(max_motion)) + 8
    tex = _texture(rng, H + 2 * m, W + 2 * m)
    yy, xx = np.mgrid[0:H, 0:W].astype(np.float64)
    cs = np.stack([rng.uniform(0, W, n_regions), rng.uniform(0, H, n_regions)], axis=1)
    tau = 0.18 * np.hypot(W, H)
    logits = np.stack([-((xx - c[0]) ** 2 + (yy - c[1]) ** 2) / (2 * tau * tau) * 8.0 for c in cs])
    logits -= logits.max(axis=0, keepdims=True)
    wts = np.exp(logits)
    wts /= wts.sum(axis=0, keepdims=True)

    def flow_at(px, py, w_):
        fx = np.zeros_like(px)
        fy = np.zeros_like(py)
        for k in range(n_regions):
            a = aff[k]
            ux = a[0] + a[1] * (px - cs[k, 0]) / W + a[2] * (py - cs[k, 1]) / H
            uy = a[3] + a[4] * (px - cs[k, 0]) / W + a[5] * (py - cs[k, 1]) / H
            fx += w_[k] * ux
            fy += w_[k] * uy
        return fx, fy

    aff = [np.concatenate([rng.uniform(-1, 1, 1), rng.uniform(-0.3, 0.3, 2),
                           rng.uniform(-1, 1, 1), rng.uniform(-0.3, 0.3, 2)]) for _ in range(n_regions)]
    fx, fy = flow_at(xx, yy, wts)
    scale = max_motion / max(float(np.sqrt(fx * fx + fy * fy).max()), 1e-6)
    aff = [a * scale for a in aff]
    fx, fy = fx * scale, fy * scale

    # frame 2 by inverting p + f(p) = q with fixed-point iterations p <- q - f(p)
    px, py = xx.copy(), yy.copy()
    for _ in range(12):
        lg = np.stack([-((px - c[0]) ** 2 + (py - c[1]) ** 2) / (2 * tau * tau) * 8.0 for c in cs])
        lg -= lg.max(axis=0, keepdims=True)
        w2 = np.exp(lg)
        w2 /= w2.sum(axis=0, keepdims=True)
        gx, gy = flow_at(px, py, w2)
        px, py = xx - gx, yy - gy
    f1 = tex[m:m + H, m:m + W].copy()
    f2 = _sample(tex, (py + m).astype(np.float32), (px + m).astype(np.float32))
    dx, dy = xx + fx, yy + fy
    occ = ((dx < 0) | (dx > W - 1) | (dy < 0) | (dy > H - 1)).astype(np.uint8)
    f1, f2 = _finish(rng, f1, f2, noise_sigma)
    gt = np.stack([fx, fy], axis=-1).astype(np.float32)
    return Pair(f1, f2, gt, occ)


def make_pair(W, H, motion="rects", n_rects=3, n_regions=6, max_motion=8.0,
              noise_sigma=2.0, seed=0, scene=None):
    """One seeded synthetic pair; see the module docstring.  For "rects",
    scene = (Ws, Hs) renders a W x H window of a larger scene (blob and
    rectangle sizes follow the scene size)."""
    rng = np.random.default_rng(seed)
    if motion == "rects":
        return _rects_pair(rng, int(W), int(H), int(n_rects), float(max_motion), float(noise_sigma), scene)
    if motion == "regions":
        return _regions_pair(rng, int(W), int(H), int(n_regions), float(max_motion), float(noise_sigma))
    raise ValueError("motion must be 'rects' or 'regions'")


# The benchmark workloads: (W, H, generator kwargs, up-sampling, feature length).
WORKLOADS = {
    "c1_320x240_u1.5_n32": dict(W=320, H=240, gen=dict(motion="rects", n_rects=3, max_motion=8.0),
                                up=1.5, N=32),
    "c2_1920x1080_u2.9_n128": dict(W=1920, H=1080, gen=dict(motion="rects", n_rects=8, max_motion=40.0),
                                   up=2.9, N=128),
    "c3_1280x720_u4.9_n512": dict(W=1280, H=720, gen=dict(motion="regions", n_regions=6, max_motion=60.0),
                                  up=4.9, N=512),
}


def workload_pair(name, seed=0):
    wl = WORKLOADS[name]
    return make_pair(wl["W"], wl["H"], seed=seed, **wl["gen"])


def batch_specs(n, seed=0):
    """The uneven-batch workload: n pairs with sizes uniform in [320x240,
    1920x1080] (width and height drawn independently, so the aspect varies),
    feature length 32 or 128 bits, equal up-sampling factors uniform in
    [1.5, 2.9], rectangles moving up to 8 px at 320 wide and 40 px at 1920.
    Returns one dict per pair: W, H, N, up, and the generator arguments."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        W, H = int(rng.integers(320, 1921)), int(rng.integers(240, 1081))
        N = int(rng.choice([32, 128]))
        u = float(rng.uniform(1.5, 2.9))
        mm = 8.0 + 32.0 * (max(W, H) - 320) / 1600.0
        out.append(dict(W=W, H=H, N=N, up=u,
                        gen=dict(motion="rects", n_rects=3 + (i % 6), max_motion=mm, seed=1000 + i)))
    return out


def _batch_pair(spec):
    return make_pair(spec["W"], spec["H"], **spec["gen"])


def batch_pairs(specs, workers=None):
    """The pairs of batch_specs, generated in a pool of processes (the pairs
    are independent and each is fixed by its own seed).  The workers are
    spawned, not forked, so a caller that already uses the GPU can call this."""
    import concurrent.futures
    import multiprocessing
    import os
    if workers is None:
        workers = min(16, len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else 4)
    if workers <= 1 or len(specs) <= 1:
        return [_batch_pair(s) for s in specs]
    # largest first so the pool's tail is short; results in spec order
    order = sorted(range(len(specs)), key=lambda i: -specs[i]["W"] * specs[i]["H"])
    ctx = multiprocessing.get_context("spawn")
    with concurrent.futures.ProcessPoolExecutor(max_workers=workers, mp_context=ctx) as pool:
        done = dict(zip(order, pool.map(_batch_pair, [specs[i] for i in order])))
    return [done[i] for i in range(len(specs))]
