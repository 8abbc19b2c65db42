"""GPU parity of the Middlebury harness: the device bad-pixel counts (Evaluator.h:49-76) against
the oracle, and the benchmark driver bin/bench_middv3 (main_bench_middv3.cu) run end to end on
synthetic scene directories (PNG images and mask, PFM ground truth, calib.txt, an MC-CNN volume
file), its disp0HBP_ISP.pfm compared bit for bit with the oracle's pipeline and its logged rates
with the oracle's counts."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle_ffi as O
from hbp_amd import middlebury as M
from hbp_amd import synth

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "hbp-isp_amd", "bin", "bench_middv3")
ALPHA, TAU_S, MIN_LEVEL = 1000.1, 0.6, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_bits(a, b):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    return a.shape == b.shape and np.array_equal(na, nb) and np.array_equal(a.view(np.uint32)[~na],
                                                                            b.view(np.uint32)[~nb])


# ---------------------------------------------------------------- device counts

def eval_case(seed, H, W):
    rng = np.random.default_rng(seed)
    gt = (rng.random((H, W)) * 200).astype(np.float32)
    disp = gt + rng.normal(0, 1.5, (H, W)).astype(np.float32)
    for frac, v in [(0.05, 0.0), (0.03, np.inf), (0.02, -np.inf), (0.02, np.nan), (0.01, -3.0)]:
        gt[rng.random((H, W)) < frac] = v
    disp[rng.random((H, W)) < 0.03] = np.nan
    disp[rng.random((H, W)) < 0.02] = np.inf
    mask = np.where(rng.random((H, W)) < 0.75, 255, 0).astype(np.uint8)
    return disp, gt, mask


@pytest.mark.parametrize("seed,H,W", [(1, 48, 64), (2, 1, 1), (3, 7, 1031), (4, 984, 1472)])
@pytest.mark.parametrize("thr", [0.5, 1.0, 2.0, 0.3])
def test_eval_counts_match_oracle(ctx, seed, H, W, thr):
    say(f"eval seed={seed} {W}x{H} thr={thr}")
    disp, gt, mask = eval_case(seed, H, W)
    assert ctx.eval_bad_pixels(disp, gt, mask, thr) == O.eval_bad_pixels(disp, gt, mask, thr)


def test_eval_threshold_in_float(ctx):
    t = np.float32(0.3)
    gt = np.float32([[t, 10.0, 10.0, 5.0]])
    disp = np.float32([[0.0, 10.5, np.nan, 5.0]])
    mask = np.uint8([[255, 255, 255, 0]])
    assert ctx.eval_bad_pixels(disp, gt, mask, 0.3) == [(4, 2), (3, 1), (1, 1)]


# ---------------------------------------------------------------- the driver

CALIB = ("cam0 = [120.5 0 40.25; 0 120.5 30.5; 0 0 1]\ncam1 = [120.5 0 44.75; 0 120.5 30.5; 0 0 1]\n"
         "doffs = 4.5\nbaseline = 160.0\nwidth = {W}\nheight = {H}\nndisp = {D}\nisint = 0\nvmin = 1\n"
         "vmax = {D}\ndyavg = 0.1\ndymax = 0.5\n")


def make_scene(root, kind, seed, W, H, D, mccnn=False):
    """a scene directory under root/<kind>/Synth<seed>/ with every file the driver reads; PNG rows
    use all five filter types"""
    left, right, gt = synth.make_pair(seed, W, H, D, (3, 6))
    d = os.path.join(root, kind, f"Synth{seed}") + os.sep
    os.makedirs(d)
    with open(d + "calib.txt", "w") as f:
        f.write(CALIB.format(W=W, H=H, D=D))
    M.write_png(d + "im0.png", left, filters=[(y * 3) % 5 for y in range(H)])
    M.write_png(d + "im1.png", right, filters=[(y * 7 + 1) % 5 for y in range(H)])
    rng = np.random.default_rng(seed)
    gt = gt.astype(np.float32)
    gt[rng.random((H, W)) < 0.03] = np.inf          # unknown pixels
    mask = np.where(rng.random((H, W)) < 0.85, 255, 128).astype(np.uint8)
    M.write_pfm(d + "disp0GT.pfm", gt)
    M.write_png(d + "mask0nocc.png", mask, filters=[y % 5 for y in range(H)])
    raw = None
    if mccnn:
        vol, _ = O.cost_volume(left, right, D)
        raw = (vol.astype(np.float32) / np.float32(25.0) + rng.normal(0, 0.05, vol.shape).astype(np.float32))
        raw = raw.astype(np.float32)
        raw.tofile(d + "im0.acrt")
    return d, left, right, gt, np.where(mask == 255, 255, 0).astype(np.uint8), raw


def run_driver(d, out, cost="Census", lr=False):
    args = [DRIVER, d, out, "2", "4", "0", str(ALPHA), str(TAU_S), cost, "1", str(MIN_LEVEL)] + (["lr"] if lr else [])
    say("driver:", " ".join(args[1:]))
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return r.stdout


def oracle_view(img, vol, wta, D):
    bp = O.BP(O.Pyramid(img), vol, wta, D, alpha=ALPHA, tau_s=TAU_S, min_level=MIN_LEVEL)
    bp.step()
    return bp.map()


def logged_rates(out):
    with open(out + "debug/log_output.txt") as f:
        lines = f.read().splitlines()
    assert lines[0] == "Time\tall\tnonocc"
    return [ln.split("\t")[1:] for ln in lines[1:]]


def expected_rates(disp, gt, mask, thr):
    return ["%f" % (100.0 * (1.0 - g / n)) for n, g in O.eval_bad_pixels(disp, gt, mask, thr)]


@pytest.fixture(scope="module")
def driver_built():
    if not os.path.exists(DRIVER):
        pytest.fail("bin/bench_middv3 is not built (make -C hbp-isp_amd)")


def test_driver_census_matches_oracle(tmp_path, driver_built):
    W, H, D = 96, 72, 16
    d, left, right, gt, mask, _ = make_scene(str(tmp_path), "trainingQ", 21, W, H, D)
    out = str(tmp_path / "out") + os.sep
    stdout = run_driver(d, out)
    assert "ndisp = 16" in stdout and "errorThresh = 0.500000" in stdout
    got = M.read_pfm(out + "disp0HBP_ISP.pfm")
    vol, wta = O.cost_volume(left, right, D)
    disp, _ = oracle_view(left, vol, wta, D)
    assert same_bits(got, disp)
    assert logged_rates(out) == [expected_rates(disp, gt, mask, 0.5)]
    with open(out + "timeHBP_ISP.txt") as f:
        assert float(f.read()) > 0.0


def test_driver_lr_matches_oracle(tmp_path, driver_built):
    W, H, D = 80, 64, 16
    d, left, right, gt, mask, _ = make_scene(str(tmp_path), "trainingF", 22, W, H, D)
    out = str(tmp_path / "out") + os.sep
    run_driver(d, out, lr=True)
    vol, wta = O.cost_volume(left, right, D)
    dl, pl = oracle_view(left, vol, wta, D)
    volr, wtar = O.cost_volume(left, right, D, right_view=True)
    dr, pr = oracle_view(right, volr, wtar, D)
    fl, fr, f2 = (np.zeros((H, W), np.uint8) for _ in range(3))
    dl0 = dl.copy()
    O.lib.orc_left_right_check(dl, dr, pl, pr, W, H, fl, fr, f2)
    checked = np.zeros((H, W), np.float32)
    O.lib.orc_plane_to_disp(pl, W, H, checked)
    O.lib.orc_weighted_median(left, fl, pl, W, H)
    final = np.zeros((H, W), np.float32)
    O.lib.orc_plane_to_disp(pl, W, H, final)
    assert same_bits(M.read_pfm(out + "disp0HBP_ISP.pfm"), final)
    # MidV3_LR evaluates three times: BP alone, after the check and fill, after the refinement
    assert logged_rates(out) == [expected_rates(x, gt, mask, 2.0) for x in (dl0, checked, final)]
    with open(out + "timeHBP_ISP.txt") as f:
        assert f.read().startswith("runtime ")


def test_driver_mccnn_volume_file(tmp_path, driver_built):
    W, H, D = 64, 48, 16
    d, left, right, gt, mask, raw = make_scene(str(tmp_path), "testH", 23, W, H, D, mccnn=True)
    out = str(tmp_path / "out") + os.sep
    run_driver(d, out, cost="MC-CNN")
    vol, wta = O.cost_volume_f32(raw, 0.5)
    disp, _ = oracle_view(left, vol, wta, D)
    assert same_bits(M.read_pfm(out + "disp0HBP_ISP.pfm"), disp)
    assert logged_rates(out) == []          # a test scene is not evaluated


def test_driver_reports_missing_inputs(tmp_path, driver_built):
    os.makedirs(tmp_path / "empty")
    r = subprocess.run([DRIVER, str(tmp_path / "empty") + "/", str(tmp_path / "o") + "/", "2", "4", "0", "1000.1",
                        "0.6", "Census", "1", "2"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "ndisp is not speficied." in r.stdout
    r = subprocess.run([DRIVER, "a", "b"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "usage" in r.stderr
