"""The HIP path (C ABI of libhbp_hip.so) against the committed golden fixtures of tests/golden:
the 64x48 pair's cost volumes, pyramid levels, BP state after the first sweep and after the first
level, final disparity and planes of both views, and the left-right check.  Bit for bit."""
import json
import os
import sys

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
META = json.load(open(os.path.join(GOLD, "meta.json")))


def load(name):
    return np.load(os.path.join(GOLD, name + ".npy"), allow_pickle=False)


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_bits(a, b):
    """equal bit for bit; for floats any NaN equals any NaN"""
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    if a.dtype == np.float32:
        na, nb = np.isnan(a), np.isnan(b)
        return np.array_equal(na, nb) and np.array_equal(a.view(np.uint32)[~na], b.view(np.uint32)[~nb])
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def same_state(st, tag):
    # slots 0..2P-1; the candidate slots 2P..3P-1 hold the next sweep's proposals, which this
    # library draws at the end of a sweep and the oracle at the start of the next
    return (same_bits(st["planes"][:8], load(f"bp_{tag}_planes")[:8]) and
            same_bits(st["belief"], load(f"bp_{tag}_belief")) and same_bits(st["msg"], load(f"bp_{tag}_msg")))


@pytest.mark.parametrize("D", [16, 64])
def test_golden_volume(ctx, D):
    say(f"golden: volume D={D}")
    vol, wta = ctx.cost_volume(load("left"), load("right"), D).download()
    assert same_bits(vol, load(f"vol_d{D}"))
    assert same_bits(wta, load(f"wta_d{D}"))


def test_golden_right_volume(ctx):
    say("golden: right view volume")
    vol, wta = ctx.cost_volume(load("left"), load("right"), 16, right_view=True).download()
    assert same_bits(vol, load("vol_right_d16"))
    assert same_bits(wta, load("wta_right_d16"))


def test_golden_pyramid(ctx):
    say("golden: pyramid")
    pyr = ctx.pyramid(load("left"))
    assert pyr.levels == META["levels"]
    for l in range(pyr.levels):
        lv = pyr.level(l)
        assert (lv["V"], lv["E"]) == (META["V"][l], META["E"][l]), l
        if l < 3:
            assert same_bits(lv["clus"], load(f"pyr_l{l}_clus")), l
            assert same_bits(lv["bd"], load(f"pyr_l{l}_bd")), l
            assert same_bits(lv["bdl"], load(f"pyr_l{l}_bdl")), l


def test_golden_bp_and_lr(ctx):
    say("golden: BP states, MAP, right view, LR check")
    left, right = load("left"), load("right")
    D, ml = META["D_bp"], META["min_level"]
    pyr = ctx.pyramid(left)
    vol = ctx.cost_volume(left, right, D)
    bp = ctx.bp(pyr, vol, alpha=META["alpha"], tau_s=META["tau_s"], min_level=ml)
    assert bp.total_sweeps == META["total_sweeps"]
    assert bp.step(1) == 1
    st = bp.state()
    assert st["level"] == META["bp_first_level"]
    assert same_state(st, "first")
    n = META["bp_first_level_sweeps"] - 1
    assert bp.step(n) == n
    st = bp.state()
    assert st["level"] == META["bp_first_level"]
    assert same_state(st, "level")
    disp, plane = bp.run()
    assert same_bits(disp, load("disp")) and same_bits(plane, load("plane3"))
    pyr_r = ctx.pyramid(right)
    vol_r = ctx.cost_volume(left, right, D, right_view=True)
    dr, pr = ctx.bp(pyr_r, vol_r, alpha=META["alpha"], tau_s=META["tau_s"], min_level=ml).run()
    assert same_bits(dr, load("disp_right")) and same_bits(pr, load("plane3_right"))
    filled, fl, fr, f2 = ctx.lr_check_fill(disp, dr, plane)
    assert same_bits(fl, load("lr_fail_l")) and same_bits(fr, load("lr_fail_r"))
    assert same_bits(f2, load("lr_fail2_l"))
    assert same_bits(filled, load("lr_plane3"))
    assert same_bits(ctx.weighted_median(left, fl, filled), load("wm_plane3"))
    assert same_bits(ctx.median7(disp), load("disp_median7"))
