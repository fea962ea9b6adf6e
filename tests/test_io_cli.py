"""Flow / image file formats and the command-line front end on the CPU.

The .flo reader is pinned by the format description (flowIO.cpp:5-20) applied
with numpy to the reference's own data/RubberWhale/flow10.flo; the PNG reader
and writer by PIL; the CLI's parameter errors (OFPipeline::setParams,
of_pipeline.cpp:90-195) run before any GPU work."""
import os
import subprocess

import numpy as np
import pytest

import briefmatch_amd as bm
from briefmatch_amd import io as bio

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rubberwhale")


def test_read_flo_matches_format():
    path = os.path.join(GOLD, "flow10.flo")
    raw = np.fromfile(path, np.uint8)
    assert raw[:4].tobytes() == b"PIEH"
    assert np.frombuffer(raw[:4].tobytes(), np.float32)[0] == np.float32(202021.25)
    W, H = np.frombuffer(raw[4:12].tobytes(), np.int32)
    data = np.frombuffer(raw[12:].tobytes(), np.float32).reshape(H, W, 2)
    U, V = bio.read_flo(path)
    assert U.shape == (388, 584) == (H, W)
    assert np.array_equal(U, data[..., 0], equal_nan=True) and np.array_equal(V, data[..., 1], equal_nan=True)
    # the fixture marks unknown pixels with 1e9+ values
    unknown = (np.abs(U) > 1e9) | (np.abs(V) > 1e9)
    assert 0 < unknown.mean() < 0.5


def test_flo_roundtrip_and_errors(tmp_path):
    rng = np.random.default_rng(0)
    U = rng.normal(size=(7, 11)).astype(np.float32)
    V = rng.normal(size=(7, 11)).astype(np.float32)
    p = str(tmp_path / "a.flo")
    bio.write_flo(p, U, V)
    U2, V2 = bio.read_flo(p)
    assert np.array_equal(U, U2) and np.array_equal(V, V2)
    raw = open(p, "rb").read()
    for bad, msg in [(b"XXXX" + raw[4:], "wrong tag"), (raw[:-4], "too short"), (raw + b"\0", "too long")]:
        q = str(tmp_path / "b.flo")
        open(q, "wb").write(bad)
        with pytest.raises(bm.BriefMatchError, match=msg):
            bio.read_flo(q)
    with pytest.raises(bm.BriefMatchError, match="flo"):
        bio.write_flo(str(tmp_path / "c.txt"), U, V)
    r = str(tmp_path / "d.bin")
    bio.write_raw(r, U, V)
    d = np.fromfile(r, np.float32)
    assert np.array_equal(d[:U.size].reshape(U.shape), U) and np.array_equal(d[U.size:].reshape(V.shape), V)


def test_png_reader_against_pil(tmp_path):
    from PIL import Image
    g = np.asarray(Image.open(os.path.join(GOLD, "frame10.png")))
    img = bio.read_image(os.path.join(GOLD, "frame10.png"))
    assert img.shape == g.shape + (3,)
    assert all(np.array_equal(img[..., c], g) for c in range(3))
    rng = np.random.default_rng(1)
    rgb = rng.integers(0, 256, (13, 17, 3)).astype(np.uint8)
    for mode, arr in [("RGB", rgb), ("RGBA", np.concatenate([rgb, rgb[..., :1]], -1)), ("L", rgb[..., 0])]:
        p = str(tmp_path / ("x_%s.png" % mode))
        Image.fromarray(arr, mode).save(p)
        got = bio.read_image(p)
        exp = arr[..., :3] if arr.ndim == 3 else np.stack([arr] * 3, -1)
        assert np.array_equal(got, exp), mode
    p16 = str(tmp_path / "x16.png")
    Image.fromarray((rgb[..., 0].astype(np.uint16) * 257), "I;16").save(p16)
    assert np.array_equal(bio.read_image(p16)[..., 0], rgb[..., 0])
    # PNG writer read back by PIL; binary PPM input
    w = str(tmp_path / "w.png")
    bio.write_png(w, rgb)
    assert np.array_equal(np.asarray(Image.open(w).convert("RGB")), rgb)
    ppm = str(tmp_path / "x.ppm")
    with open(ppm, "wb") as f:
        f.write(b"P6\n# c\n17 13\n255\n" + rgb.tobytes())
    assert np.array_equal(bio.read_image(ppm), rgb)


def test_color_coding():
    """Middlebury wheel (55 hues: RY 15, YG 6, GC 4, CB 11, BM 13, MR 6): +x maps
    to hue 0 (red), +y to hue 13.5 of 54 (between (255,221,0) and (255,238,0)),
    -x to hue 27 (CB segment, (0, 255 - 255*2/11, 255)), -y to hue 40.5 (BM,
    between (78,0,255) and (98,0,255)); on the unit circle the colour is the
    pure wheel colour, zero motion is white, unknown flow black."""
    U = np.array([[0.0, 1.0, 0.0, 1e10, -1.0, 0.0]], np.float32)
    V = np.array([[0.0, 0.0, 1.0, 0.0, 0.0, -1.0]], np.float32)
    c = bio.flow_to_rgb(U, V, maxmot=1.0)
    assert [tuple(int(v) for v in px) for px in c[0]] == [(255, 255, 255), (255, 0, 0), (255, 229, 0), (0, 0, 0),
                                                         (0, 209, 255), (87, 0, 255)]
    # half radius: halfway to white; beyond the radius: darkened to 75 %
    h = bio.flow_to_rgb(np.array([[0.5, 2.0]], np.float32), np.zeros((1, 2), np.float32), maxmot=1.0)
    assert tuple(int(v) for v in h[0, 0]) == (255, 127, 127) and tuple(int(v) for v in h[0, 1]) == (191, 0, 0)


def _cli(*args, cwd=None):
    return subprocess.run([bio.CLI_PATH, *args], capture_output=True, text=True, cwd=cwd, timeout=60)


@pytest.mark.skipif(not os.path.exists(bio.CLI_PATH), reason="CLI not built")
def test_cli_parameter_errors(tmp_path):
    r = _cli("--help")
    assert r.returncode == 1 and "--feature-length" in r.stdout and "--range-spatial" in r.stdout
    frames = os.path.join(GOLD, "frame%02d.png")
    cases = [
        (["--bogus", "1"], "not a valid input option"),
        (["-f", "10:11"], "Required input option"),
        (["-i", frames, "-f", "10-11"], "Unable to parse frame range"),
        (["-i", frames, "-f", "11:10"], "End frame should be >= start frame"),
        (["-i", frames, "-f", "10:11", "-o", "x.jpg"], "Unsupported output format"),
        (["-i", frames, "-f", "10:11", "--feature-length", "600"], "out of range"),
        (["-i", frames, "-f", "10:11", "-fl"], "No value provided"),
        (["-i", str(tmp_path / "none%02d.png"), "-f", "1:2"], "Cannot read image file"),
        (["-i", frames, "-f", "10:11", "--upsampler", "cubic"], "expects bilinear or lanczos3"),
        (["-i", os.path.join(GOLD, "frame%s.png"), "-f", "10:11"], "Frame pattern"),
        (["-i", os.path.join(GOLD, "f%02d_%d.png"), "-f", "10:11"], "Frame pattern"),
        (["-i", frames, "-f", "10:11", "-o", "out%n.flo"], "Frame pattern"),
        # the frame number is an unsigned int: long and long long conversions
        # would read an argument that was never passed
        (["-i", os.path.join(GOLD, "frame%02lu.png"), "-f", "10:11"], "Frame pattern"),
        (["-i", os.path.join(GOLD, "frame%02lld.png"), "-f", "10:11"], "Frame pattern"),
        (["-i", os.path.join(GOLD, "frame%02hhhd.png"), "-f", "10:11"], "Frame pattern"),
        (["-i", frames, "-f", "10:11", "-ls", "16:6"], "needs '--proposal'"),
        (["-i", frames, "-f", "10:11", "-pp", "w.bnet", "-ls", "16"], "expects RADIUS:SAMPLES"),
        (["-i", frames, "-f", "10:11", "-pp", "w.bnet", "-ls", "16:0"], "expects RADIUS:SAMPLES"),
        (["-i", frames, "-f", "10:11", "-pp", "w.bnet", "-ls", "0:4"], "expects RADIUS:SAMPLES"),
        (["-i", frames, "-f", "10:11", "-pp", "w.bnet", "-ls", "16:6x"], "expects RADIUS:SAMPLES"),
    ]
    for args, msg in cases:
        r = _cli(*args)
        assert r.returncode == 1, args
        assert msg in r.stderr, (args, r.stderr)


@pytest.mark.skipif(__import__("shutil").which("g++") is None, reason="no host compiler")
def test_readers_under_sanitizers(tmp_path):
    """The PNG / PNM / .flo readers under AddressSanitizer and UBSan
    (tests/io_sanitize.cpp): every truncation and 3,000 seeded mutations of
    valid files each, and headers that promise gigabytes; every input parses
    or raises IOError, and no input may make the reader allocate more than
    2 GiB (ASan's allocation cap turns that into a failure)."""
    root = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    exe = str(tmp_path / "io_sanitize")
    b = subprocess.run(["g++", "-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
                        "-std=c++17", "-I" + os.path.join(root, "include"),
                        os.path.join(root, "tests", "io_sanitize.cpp"),
                        os.path.join(root, "briefmatch_amd", "csrc", "bm_io.cpp"), "-lz", "-o", exe],
                       capture_output=True, text=True, timeout=600)
    assert b.returncode == 0, b.stderr
    env = dict(os.environ, ASAN_OPTIONS="max_allocation_size_mb=2048:allocator_may_return_null=0")
    r = subprocess.run([exe, str(tmp_path)], capture_output=True, text=True, timeout=600, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok" in r.stdout
