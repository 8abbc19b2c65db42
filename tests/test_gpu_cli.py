"""The single-pair driver bin/stereo_hbp_isp (main_stereo_hbp_isp.cu's argument order and call
sequence) run on binary PPM images: its PFM outputs equal, bit for bit, the oracle's pipeline:
BP, then the CLI's 7-tap median; with "lr", the right view, the left-right check and fill and
the weighted median.  Also the AD cost through the driver."""
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
CLI = os.path.join(ROOT, "hbp-isp_amd", "bin", "stereo_hbp_isp")
ALPHA, TAU_S, MIN_LEVEL = 1000.1, 0.6, 2


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_bits(a, b):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    return a.shape == b.shape and np.array_equal(na, nb) and np.array_equal(a.view(np.uint32)[~na],
                                                                            b.view(np.uint32)[~nb])


def write_ppm(path, bgr):
    h, w = bgr.shape[:2]
    with open(path, "wb") as f:
        f.write(b"P6\n# synthetic\n%d %d\n255\n" % (w, h))
        f.write(np.ascontiguousarray(bgr[..., ::-1]).tobytes())


def run_cli(tmp_path, left, right, D, cost="Census", lr=False, sigma="0"):
    if not os.path.exists(CLI):
        pytest.fail("bin/stereo_hbp_isp is not built (make -C hbp-isp_amd)")
    lp, rp, out = tmp_path / "l.ppm", tmp_path / "r.ppm", str(tmp_path / "res")
    write_ppm(lp, left)
    write_ppm(rp, right)
    args = [CLI, str(lp), str(rp), str(D), "2", "4", sigma, str(ALPHA), str(TAU_S), out, "1", str(MIN_LEVEL), cost]
    args += ["lr"] if lr else []
    say("cli:", " ".join(args[3:]))
    r = subprocess.run(args, capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    return out


def oracle_view(img, vol, wta, D, tau=None, sigma=0.0):
    bp = O.BP(O.Pyramid(img, sigma=sigma), vol, wta, D, alpha=ALPHA, tau_s=TAU_S, min_level=MIN_LEVEL, tau=tau)
    bp.step()
    return bp.map()


def test_cli_census_with_median7(tmp_path):
    W, H, D = 96, 72, 16
    left, right, _ = synth.make_pair(71, W, H, D, (3, 6))
    out = run_cli(tmp_path, left, right, D)
    vol, wta = O.cost_volume(left, right, D)
    disp, _ = oracle_view(left, vol, wta, D)
    O.lib.orc_median7(disp, W, H)
    assert same_bits(M.read_pfm(out + "_hbp_isp_disp.pfm"), disp)


def test_cli_census_with_gaussian_preblur(tmp_path):
    W, H, D = 90, 70, 16
    left, right, _ = synth.make_pair(74, W, H, D, (3, 6))
    out = run_cli(tmp_path, left, right, D, sigma="0.8")
    vol, wta = O.cost_volume(left, right, D)
    disp, _ = oracle_view(left, vol, wta, D, sigma=0.8)
    O.lib.orc_median7(disp, W, H)
    assert same_bits(M.read_pfm(out + "_hbp_isp_disp.pfm"), disp)


@pytest.mark.parametrize("seed,W,H,D", [(72, 80, 64, 16), (75, 101, 67, 19)])
def test_cli_lr_refinement(tmp_path, seed, W, H, D):
    left, right, _ = synth.make_pair(seed, W, H, D, (3, 6))
    out = run_cli(tmp_path, left, right, D, lr=True)
    vol, wta = O.cost_volume(left, right, D)
    dl, pl = oracle_view(left, vol, wta, D)
    volr, wtar = O.cost_volume(left, right, D, right_view=True)
    dr, pr = oracle_view(right, volr, wtar, D)
    fl, fr, f2 = (np.zeros((H, W), np.uint8) for _ in range(3))
    O.lib.orc_left_right_check(dl, dr, pl, pr, W, H, fl, fr, f2)
    O.lib.orc_weighted_median(left, fl, pl, W, H)
    final = np.zeros((H, W), np.float32)
    O.lib.orc_plane_to_disp(pl, W, H, final)
    assert same_bits(M.read_pfm(out + "_hbp_isp_disp_lr.pfm"), final)


def test_cli_ad_cost(tmp_path):
    W, H, D = 96, 72, 16
    left, right, _ = synth.make_pair(73, W, H, D, (3, 6))
    out = run_cli(tmp_path, left, right, D, cost="AD")
    vol, wta = O.cost_volume_ad(left, right, D)
    disp, _ = oracle_view(left, vol, wta, D, tau=30.0)
    O.lib.orc_median7(disp, W, H)
    assert same_bits(M.read_pfm(out + "_hbp_isp_disp.pfm"), disp)
