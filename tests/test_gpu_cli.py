"""The drop-in C++ layer on the GPU: the `briefmatch` tool (OFPipeline +
BriefMatch over the C ABI) on the reference's RubberWhale frames."""
import os
import re
import shutil
import subprocess

import numpy as np
import pytest

import briefmatch_amd as bm
from briefmatch_amd import io as bio
from conftest import say

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(not os.path.exists(bio.CLI_PATH), reason="CLI not built")]

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "rubberwhale")


def _run(args, cwd):
    say("cli", " ".join(args))
    r = subprocess.run([bio.CLI_PATH, *args], capture_output=True, text=True, cwd=cwd, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r


def test_cli_pair_equals_flow_pair(tmp_path):
    """frames 10 -> 11 through OFPipeline equal bm_flow_pair on the same frames."""
    _run(["-i", os.path.join(GOLD, "frame%02d.png"), "-f", "10:11", "-ux", "2.9", "-uy", "2.9", "-fl", "128",
          "-o", str(tmp_path / "flow%02d.flo")], cwd=tmp_path)
    assert not (tmp_path / "flow10.flo").exists()          # the first frame has no predecessor
    U, V = bio.read_flo(str(tmp_path / "flow11.flo"))
    f10 = bio.read_image(os.path.join(GOLD, "frame10.png"))
    f11 = bio.read_image(os.path.join(GOLD, "frame11.png"))
    p = bm.derive_params(584, 388, 2.9, 2.9, N=128)
    ctx = bm.Context(p, seed=0)
    pU, pV = ctx.flow_numpy(f10, f11, pair=0)
    assert np.array_equal(U, pU) and np.array_equal(V, pV)
    gU, gV = bio.read_flo(os.path.join(GOLD, "flow10.flo"))
    gt = np.stack([gU, gV], -1)
    e, _ = bm.epe(U, V, gt)
    z, _ = bm.epe(np.zeros_like(U), np.zeros_like(V), gt)
    print("[epe] RubberWhale 10->11: %.4f (zero flow %.4f)" % (e, z), flush=True)
    assert e < 0.5 * z


def test_cli_sequence_outputs(tmp_path):
    """A three-frame run with temporal initialisation, .png / .bin outputs and
    the EPE print of --flow-gt."""
    for f in ("frame09.png", "frame10.png", "frame11.png"):
        shutil.copy(os.path.join(GOLD, f), tmp_path / f)
    r = _run(["-i", "frame%02d.png", "-f", "9:11", "-o", "vis%02d.png", "-fl", "64", "-ux", "1.5", "-uy", "1.5",
              "-gt", os.path.join(GOLD, "flow10.flo")], cwd=tmp_path)
    assert (tmp_path / "vis10.png").exists() and (tmp_path / "vis11.png").exists()
    vis = bio.read_image(str(tmp_path / "vis11.png"))
    assert vis.shape == (388, 584, 3)
    errs = [float(x) for x in re.findall(r"Error \(EPE\) = ([0-9.]+)", r.stdout)]
    assert len(errs) == 2 and all(np.isfinite(errs))
    _run(["-i", "frame%02d.png", "-f", "10:11", "-o", "raw%02d.bin", "-fl", "32", "-ux", "1.5", "-uy", "1.5"],
         cwd=tmp_path)
    d = np.fromfile(str(tmp_path / "raw11.bin"), np.float32)
    assert d.size == 2 * 584 * 388 and np.isfinite(d).all()


def test_cli_lanczos_equals_flow_pair(tmp_path):
    """--upsampler lanczos3 through OFPipeline equals bm_flow_pair with the
    Lanczos-3 mode on the same frames."""
    _run(["-i", os.path.join(GOLD, "frame%02d.png"), "-f", "10:11", "-ux", "2.0", "-uy", "2.0", "-fl", "64",
          "--upsampler", "lanczos3", "-o", str(tmp_path / "flow%02d.flo")], cwd=tmp_path)
    U, V = bio.read_flo(str(tmp_path / "flow11.flo"))
    f10 = bio.read_image(os.path.join(GOLD, "frame10.png"))
    f11 = bio.read_image(os.path.join(GOLD, "frame11.png"))
    p = bm.derive_params(584, 388, 2.0, 2.0, N=64)
    ctx = bm.Context(p, seed=0)
    ctx.set_upsampler(bm.UPSAMPLE_LANCZOS3)
    pU, pV = ctx.flow_numpy(f10, f11, pair=0)
    assert np.array_equal(U, pU) and np.array_equal(V, pV)


def test_cli_proposal_equals_flow_pair_learned(tmp_path):
    """--proposal through OFPipeline equals bm_flow_pair_learned on the same
    frames; in a three-frame run the second pair (stream index 1) reads the
    previous frame's RGB kept on the device."""
    from briefmatch_amd import learned
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no trained proposal weights")
    _run(["-i", os.path.join(GOLD, "frame%02d.png"), "-f", "9:11", "-ux", "2.9", "-uy", "2.9", "-fl", "128",
          "-it", "1", "--proposal", learned.DEFAULT_WEIGHTS, "-o", str(tmp_path / "flow%02d.flo")], cwd=tmp_path)
    assert not (tmp_path / "flow09.flo").exists()
    frames = {k: bio.read_image(os.path.join(GOLD, "frame%02d.png" % k)) for k in (9, 10, 11)}
    p = bm.derive_params(584, 388, 2.9, 2.9, N=128)
    p.iterations = 1
    ctx = bm.Context(p, seed=0)
    net = bm.Net(path=learned.DEFAULT_WEIGHTS)
    for pair, (a, b) in enumerate(((9, 10), (10, 11))):
        U, V = bio.read_flo(str(tmp_path / ("flow%02d.flo" % b)))
        pU, pV = ctx.flow_numpy_learned(net, frames[a], frames[b], pair=pair)
        assert np.array_equal(U, pU) and np.array_equal(V, pV), (a, b)


def test_cli_proposal_local_search_equals_flow_pair_learned(tmp_path):
    """--proposal with --local-search 16:6 and --jf-max 8 (the benchmark's
    learned configuration) equals bm_flow_pair_learned with the local search
    radius 16 input pixels times the up-sampling factor."""
    from briefmatch_amd import learned
    if not os.path.exists(learned.DEFAULT_WEIGHTS):
        pytest.skip("no trained proposal weights")
    _run(["-i", os.path.join(GOLD, "frame%02d.png"), "-f", "10:11", "-ux", "2.9", "-uy", "2.9", "-fl", "128",
          "-it", "1", "-jm", "8", "--proposal", learned.DEFAULT_WEIGHTS, "--local-search", "16:6",
          "-o", str(tmp_path / "flow%02d.flo")], cwd=tmp_path)
    p = bm.derive_params(584, 388, 2.9, 2.9, N=128)
    p.iterations = 1
    p.jfMax = 8
    ctx = bm.Context(p, seed=0)
    ctx.set_local_search(learned.local_search_radius(p, 16.0), 6)
    assert ctx.local_search_config() == (46, 6)
    net = bm.Net(path=learned.DEFAULT_WEIGHTS)
    frames = [bio.read_image(os.path.join(GOLD, "frame%02d.png" % k)) for k in (10, 11)]
    U, V = bio.read_flo(str(tmp_path / "flow11.flo"))
    pU, pV = ctx.flow_numpy_learned(net, frames[0], frames[1], pair=0)
    assert np.array_equal(U, pU) and np.array_equal(V, pV)


def test_cli_proposal_missing_weights(tmp_path):
    r = subprocess.run([bio.CLI_PATH, "-i", os.path.join(GOLD, "frame%02d.png"), "-f", "10:11",
                        "--proposal", str(tmp_path / "none.bnet"), "-o", str(tmp_path / "f%02d.flo")],
                       capture_output=True, text=True, cwd=tmp_path, timeout=120)
    assert r.returncode != 0
    assert "proposal weights" in (r.stdout + r.stderr)
