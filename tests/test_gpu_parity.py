"""GPU parity: the HIP path (through the C ABI of libhbp_hip.so) against the oracle on the same
seeded synthetic pairs.  Integer/byte outputs must be bit-exact; float outputs are compared bit
for bit as well (the pinned orders make them deterministic), NaNs matching NaNs."""
import ctypes as C
import os
import sys
import time

import numpy as np
import pytest

import oracle_ffi as O
from hbp_amd import lib as H
from hbp_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_bits(a, b):
    """float arrays equal bit for bit, any NaN equal to any NaN"""
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    if a.shape != b.shape:
        return False
    na, nb = np.isnan(a), np.isnan(b)
    if not np.array_equal(na, nb):
        return False
    return np.array_equal(a.view(np.uint32)[~na], b.view(np.uint32)[~nb])


def mismatch(a, b):
    a = np.asarray(a, np.float32)
    b = np.asarray(b, np.float32)
    both = np.isnan(a) & np.isnan(b)
    return (a.view(np.uint32) != b.view(np.uint32)) & ~both


# ---------------------------------------------------------------- (a)
COST_CASES = [(1, 64, 48, 16), (2, 37, 23, 5), (3, 450, 375, 64), (4, 130, 70, 100), (5, 1030, 40, 256)]


@pytest.mark.parametrize("seed,W,H,D", COST_CASES)
@pytest.mark.parametrize("right_view", [False, True])
def test_cost_volume_bit_exact(ctx, seed, W, H, D, right_view):
    say(f"cost volume seed={seed} {W}x{H} D={D} right={right_view}")
    left, right, _ = synth.make_pair(seed, W, H, D, (3, 6))
    vo, wo = O.cost_volume(left, right, D, right_view)
    v = ctx.cost_volume(left, right, D, right_view)
    vg, wg = v.download()
    assert np.array_equal(vg, vo)
    assert np.array_equal(wg, wo)


def test_cost_volume_large_pair(ctx):
    say("cost volume 1472x984 D=256")
    left, right, _ = synth.make_pair(11, 1472, 984, 256, (20, 40))
    vo, wo = O.cost_volume(left, right, 256)
    v = ctx.cost_volume(left, right, 256)
    vg, wg = v.download()
    assert np.array_equal(vg, vo)
    assert np.array_equal(wg, wo)


def test_cost_volume_errors(ctx):
    img = np.zeros((10, 8, 3), np.uint8)
    with pytest.raises(H.HbpError):
        ctx.cost_volume(img, img, 4)          # narrower than the census window
    img = np.zeros((20, 20, 3), np.uint8)
    with pytest.raises(H.HbpError):
        ctx.cost_volume(img, img, 21)         # D > W


# ---------------------------------------------------------------- (b)
PYR_CASES = [(1, 64, 48), (2, 37, 23), (6, 120, 90), (3, 450, 375)]


@pytest.mark.parametrize("seed,W,H", PYR_CASES)
@pytest.mark.parametrize("link", [2, 0, 1, 3])
def test_pyramid_levels_bit_exact(ctx, seed, W, H, link):
    say(f"pyramid seed={seed} {W}x{H} link={link}")
    left, _, _ = synth.make_pair(seed, W, H, 16, (3, 8))
    po = O.Pyramid(left, link=link)
    pg = ctx.pyramid(left, link=link)
    assert pg.levels == po.levels
    for l in range(po.levels):
        a, b = po.level(l), pg.level(l)
        assert (a["V"], a["E"]) == (b["V"], b["E"]), l
        assert np.array_equal(a["clus"], b["clus"]), l
        assert np.array_equal(a["bd"], b["bd"]), l
        assert np.array_equal(a["bdl"], b["bdl"]), l
        assert same_bits(a["mean"], b["mean"]), l


def test_pyramid_large_pair(ctx):
    say("pyramid 1472x984")
    left, _, _ = synth.make_pair(11, 1472, 984, 256, (20, 40))
    po = O.Pyramid(left)
    pg = ctx.pyramid(left)
    assert pg.levels == po.levels
    for l in range(po.levels):
        a, b = po.level(l), pg.level(l)
        assert (a["V"], a["E"]) == (b["V"], b["E"]), l
        assert np.array_equal(a["clus"], b["clus"]), l
        assert np.array_equal(a["bd"], b["bd"]), l
        assert np.array_equal(a["bdl"], b["bdl"]), l
        assert same_bits(a["mean"], b["mean"]), l


@pytest.mark.parametrize("W,H,kind", [(20, 1000, "flat"), (24, 900, "columns")])
def test_pyramid_level_whose_crtrees_outgrows_the_schedule(ctx, W, H, kind):
    """A level above the one-block size (> 8000 nodes) whose 1-NN chains are longer than the
    scheduled pointer jumping covers (2^8): in a flat image every pixel's first minimum is the
    pixel above it, so each column is one chain of H nodes.  The scheduled kernels overflow and
    the cooperative kernel redoes the level; the labels the compaction reads in between must
    stay in range and the level must equal the oracle's."""
    say(f"pyramid CRTrees overflow {W}x{H} {kind}")
    left = np.full((H, W, 3), 90, np.uint8)
    if kind == "columns":
        left[:, :, 1] = (np.arange(W) * 9 % 256).astype(np.uint8)[None, :]
    assert W * H > 8000 and H > 2 ** 8
    po = O.Pyramid(left)
    pg = ctx.pyramid(left)
    assert pg.levels == po.levels
    for l in range(po.levels):
        a, b = po.level(l), pg.level(l)
        assert (a["V"], a["E"]) == (b["V"], b["E"]), l
        assert np.array_equal(a["clus"], b["clus"]), l
        assert np.array_equal(a["bd"], b["bd"]), l
        assert np.array_equal(a["bdl"], b["bdl"]), l
        assert same_bits(a["mean"], b["mean"]), l


@pytest.mark.parametrize("link", [0, 1, 3])
def test_pyramid_linkages_large_image(ctx, link):
    say(f"pyramid 1280x720 link={link}")
    left, _, _ = synth.make_pair(12, 1280, 720, 128, (12, 24))
    po = O.Pyramid(left, link=link)
    pg = ctx.pyramid(left, link=link)
    assert pg.levels == po.levels
    for l in range(po.levels):
        a, b = po.level(l), pg.level(l)
        assert (a["V"], a["E"]) == (b["V"], b["E"]), l
        assert np.array_equal(a["clus"], b["clus"]), l
        assert np.array_equal(a["bd"], b["bd"]), l
        assert np.array_equal(a["bdl"], b["bdl"]), l
        assert same_bits(a["mean"], b["mean"]), l


@pytest.mark.parametrize("target,link,W,H", [(2, 2, 120, 90), (50, 2, 120, 90), (300, 0, 160, 120),
                                              (5000, 2, 450, 375), (777, 3, 450, 375)])
def test_pyramid_target_clus_bit_exact(ctx, target, link, W, H):
    say(f"pyramid target_clus={target} link={link} {W}x{H}")
    left, _, _ = synth.make_pair(14, W, H, 16, (3, 8))
    po = O.Pyramid(left, link=link, target_clus=target)
    pg = ctx.pyramid(left, link=link, target_clus=target)
    assert pg.levels == po.levels
    for l in range(po.levels):
        a, b = po.level(l), pg.level(l)
        assert (a["V"], a["E"]) == (b["V"], b["E"]), l
        assert np.array_equal(a["clus"], b["clus"]), l
        assert np.array_equal(a["bd"], b["bd"]), l
        assert np.array_equal(a["bdl"], b["bdl"]), l
        assert same_bits(a["mean"], b["mean"]), l
    top = pg.level(pg.levels - 1)["clus"]
    assert top.max() + 1 == target and len(np.unique(top)) == target


def test_bp_on_target_clus_pyramid_bit_exact(ctx):
    say("bp on target_clus=40")
    left, right = _bp_pair(15, 120, 90, 24)
    vo, wo = O.cost_volume(left, right, 24)
    bo = O.BP(O.Pyramid(left, target_clus=40), vo, wo, 24, min_level=1)
    bo.step()
    do, plo = bo.map()
    bg = ctx.bp(ctx.pyramid(left, target_clus=40), ctx.cost_volume(left, right, 24), min_level=1)
    dg, plg = bg.run()
    assert same_bits(do, dg) and same_bits(plo, plg)


@pytest.mark.parametrize("sigma,W,H", [(0.1, 64, 48), (0.7, 157, 93), (1.0, 33, 9), (2.83, 120, 90),
                                         (1.0, 1472, 984)])
def test_pyramid_blurred_levels_bit_exact(ctx, sigma, W, H):
    say(f"pyramid sigma={sigma} {W}x{H}")
    left, _, _ = synth.make_pair(21, W, H, 16, (3, 8))
    po = O.Pyramid(left, sigma=sigma)
    pg = ctx.pyramid(left, sigma=sigma)
    assert pg.levels == po.levels
    for l in range(po.levels):
        a, b = po.level(l), pg.level(l)
        assert (a["V"], a["E"]) == (b["V"], b["E"]), l
        assert np.array_equal(a["clus"], b["clus"]), l
        assert np.array_equal(a["bd"], b["bd"]), l
        assert np.array_equal(a["bdl"], b["bdl"]), l
        assert same_bits(a["mean"], b["mean"]), l


@pytest.mark.parametrize("sigma,link", [(1.0, 0), (0.7, 1), (1.0, 3), (2.0, 3)])
def test_pyramid_blurred_other_linkages_bit_exact(ctx, sigma, link):
    # the blurred weight channel (the fold of the weights, not 1) enters Ward's weights and
    # every level's sums
    say(f"pyramid sigma={sigma} link={link} 157x93")
    left, _, _ = synth.make_pair(22, 157, 93, 16, (3, 8))
    po = O.Pyramid(left, link=link, sigma=sigma)
    pg = ctx.pyramid(left, link=link, sigma=sigma)
    assert pg.levels == po.levels
    for l in range(po.levels):
        a, b = po.level(l), pg.level(l)
        assert (a["V"], a["E"]) == (b["V"], b["E"]), l
        assert np.array_equal(a["clus"], b["clus"]), l
        assert np.array_equal(a["bd"], b["bd"]), l
        assert np.array_equal(a["bdl"], b["bdl"]), l
        assert same_bits(a["mean"], b["mean"]), l


def test_bp_on_blurred_pyramid_bit_exact(ctx):
    say("bp on sigma=1.2")
    left, right = _bp_pair(17, 120, 90, 24)
    vo, wo = O.cost_volume(left, right, 24)
    bo = O.BP(O.Pyramid(left, sigma=1.2), vo, wo, 24, min_level=2)
    bo.step()
    do, plo = bo.map()
    bg = ctx.bp(ctx.pyramid(left, sigma=1.2), ctx.cost_volume(left, right, 24), min_level=2)
    dg, plg = bg.run()
    assert same_bits(do, dg) and same_bits(plo, plg)


def test_sigma_beyond_the_blur_window_is_an_error(ctx):
    left, _, _ = synth.make_pair(16, 32, 24, 8, (3, 6))
    with pytest.raises(H.HbpError, match="sigma"):
        ctx.pyramid(left, sigma=3.0)


def test_target_clus_out_of_range_is_an_error(ctx):
    left, _, _ = synth.make_pair(16, 32, 24, 8, (3, 6))
    for bad in (0, 32 * 24):
        with pytest.raises(H.HbpError, match="target_clus"):
            ctx.pyramid(left, target_clus=bad)


@pytest.mark.parametrize("link", [0, 1, 3])
def test_bp_on_other_linkages_bit_exact(ctx, link):
    say(f"bp on link={link}")
    left, right = _bp_pair(4, 120, 90, 24)
    vo, wo = O.cost_volume(left, right, 24)
    bo = O.BP(O.Pyramid(left, link=link), vo, wo, 24, min_level=2)
    bo.step()
    do, plo = bo.map()
    bg = ctx.bp(ctx.pyramid(left, link=link), ctx.cost_volume(left, right, 24), min_level=2)
    dg, plg = bg.run()
    assert same_bits(do, dg) and same_bits(plo, plg)


def _hip():
    return C.CDLL("libamdhip64.so")


def test_crtrees_device_api_matches_literal_model(ctx):
    """CRTrees::get_clus / compact_clus_label on device arrays (CRTrees.hpp:16-78) against the
    oracle's literal model, on random functional graphs small (one block) and large (grid)."""
    hip = _hip()
    rng = np.random.default_rng(5)
    cases = [(V, "random") for V in (7, 300, 5000, 20000, 200000)]
    # above the one-block size: near-grid graphs (the scheduled kernels) and a chain (its
    # pointer jumping outgrows the schedule: the cooperative kernel redoes it)
    cases += [(60000, "near"), (150000, "near"), (70000, "chain")]
    for V, kind in cases:
        if kind == "random":
            nn = rng.integers(0, V, V).astype(np.int32)
        elif kind == "near":
            nn = np.clip(np.arange(V) + rng.integers(-3, 4, V), 0, V - 1).astype(np.int32)
        else:
            nn = np.arange(1, V + 1, dtype=np.int32)
            nn[-1] = V - 2
        say(f"crtrees V={V} {kind}")
        ref = np.zeros(V, np.int32)
        O.lib.orc_crtrees_literal(nn, V, 0, ref)
        d_nn, d_cl = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(d_nn), C.c_size_t(4 * V)) == 0
        assert hip.hipMalloc(C.byref(d_cl), C.c_size_t(4 * V)) == 0
        try:
            assert hip.hipMemcpy(d_nn, nn.ctypes.data_as(C.c_void_p), C.c_size_t(4 * V), 1) == 0
            assert ctx.lib.hbp_crtrees_get_clus(ctx.h, d_nn, d_cl, V) == 0, ctx.lib.hbp_last_error()
            out = np.zeros(V, np.int32)
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d_cl, C.c_size_t(4 * V), 2) == 0
            assert np.array_equal(out, ref)
            n = C.c_int()
            assert ctx.lib.hbp_crtrees_compact(ctx.h, d_cl, V, C.byref(n)) == 0
            assert hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), d_cl, C.c_size_t(4 * V), 2) == 0
            refc = ref.copy()
            assert n.value == O.lib.orc_compact(refc, V)
            assert np.array_equal(out, refc)
            nn2 = np.zeros(V, np.int32)
            assert hip.hipMemcpy(nn2.ctypes.data_as(C.c_void_p), d_nn, C.c_size_t(4 * V), 2) == 0
            assert np.array_equal(nn2, nn)          # nn is left unchanged
        finally:
            hip.hipFree(d_nn)
            hip.hipFree(d_cl)


# ---------------------------------------------------------------- (c)
def _bp_pair(seed, W, H, D, planes=(3, 8)):
    left, right, _ = synth.make_pair(seed, W, H, D, planes)
    return left, right


@pytest.mark.parametrize("seed,W,H,D,ml", [(1, 64, 48, 16, 2), (2, 96, 72, 24, 1), (6, 120, 90, 32, 2)])
def test_bp_state_after_each_of_the_first_sweeps(ctx, seed, W, H, D, ml):
    say(f"bp state seed={seed} {W}x{H} D={D} min_level={ml}")
    left, right = _bp_pair(seed, W, H, D)
    vo, wo = O.cost_volume(left, right, D)
    po = O.Pyramid(left)
    bo = O.BP(po, vo, wo, D, min_level=ml)
    vg = ctx.cost_volume(left, right, D)
    pg = ctx.pyramid(left)
    bg = ctx.bp(pg, vg, min_level=ml)
    assert bg.total_sweeps == O.lib.orc_bp_total_sweeps(po.levels, ml, D)
    for sweep in range(12):
        assert bo.step(1) == 1 and bg.step(1) == 1
        so, sg = bo.state(), bg.state()
        assert so["level"] == sg["level"]
        assert same_bits(so["cost"], sg["cost"]), f"cost differs after sweep {sweep}"
        assert same_bits(so["belief"], sg["belief"]), f"belief differs after sweep {sweep}"
        assert same_bits(so["planes"][:8], sg["planes"][:8]), f"planes differ after sweep {sweep}"
        assert same_bits(so["msg"], sg["msg"]), f"messages differ after sweep {sweep}"


@pytest.mark.parametrize("seed,W,H,D,ml", [(1, 64, 48, 16, 2), (2, 96, 72, 24, 1), (6, 120, 90, 32, 2),
                                           (8, 160, 120, 48, 3), (11, 101, 67, 20, 1), (12, 203, 131, 37, 2)])
def test_bp_full_run_bit_exact(ctx, seed, W, H, D, ml):
    say(f"bp full seed={seed} {W}x{H} D={D} min_level={ml}")
    left, right = _bp_pair(seed, W, H, D)
    vo, wo = O.cost_volume(left, right, D)
    po = O.Pyramid(left)
    bo = O.BP(po, vo, wo, D, min_level=ml)
    bo.step()
    do, plo = bo.map()
    vg = ctx.cost_volume(left, right, D)
    pg = ctx.pyramid(left)
    bg = ctx.bp(pg, vg, min_level=ml)
    dg, plg = bg.run()
    assert same_bits(do, dg), f"{mismatch(do, dg).mean():.4%} of disparities differ"
    assert same_bits(plo, plg)


@pytest.mark.parametrize("seed,W,H,D,ml", [(21, 48, 36, 2, 1), (22, 53, 41, 3, 1), (23, 80, 60, 1, 1),
                                           (24, 64, 40, 4, 2), (36, 9, 7, 5, 1)])
def test_bp_tiny_disparity_counts_bit_exact(ctx, seed, W, H, D, ml):
    """D of 1-4: the plane cost's clamp into [0, D-2] and the range test at its ends (D < 2
    takes the general form everywhere); and the smallest image the census window allows"""
    say(f"bp tiny D seed={seed} {W}x{H} D={D} min_level={ml}")
    left, right, _ = synth.make_pair(seed, W, H, max(D, 2), (3, 6))
    vo, wo = O.cost_volume(left, right, D)
    po = O.Pyramid(left)
    bo = O.BP(po, vo, wo, D, min_level=ml)
    bo.step()
    do, plo = bo.map()
    vg = ctx.cost_volume(left, right, D)
    pg = ctx.pyramid(left)
    dg, plg = ctx.bp(pg, vg, min_level=ml).run()
    assert same_bits(do, dg), f"{mismatch(do, dg).mean():.4%} of disparities differ"
    assert same_bits(plo, plg)


def test_bp_use_cost_vol_false(ctx):
    say("bp use_cost_vol=0")
    left, right = _bp_pair(3, 80, 60, 16)
    vo, wo = O.cost_volume(left, right, 16)
    po = O.Pyramid(left)
    bo = O.BP(po, vo, wo, 16, min_level=2, use_cost_vol=0, alpha=1.3)
    bo.step()
    do, plo = bo.map()
    bg = ctx.bp(ctx.pyramid(left), ctx.cost_volume(left, right, 16), alpha=1.3, min_level=2, use_cost_vol=0)
    dg, plg = bg.run()
    assert same_bits(do, dg) and same_bits(plo, plg)


def test_bp_negative_alpha(ctx):
    """alpha < 0 (every plane cost term negative): the fixed-point sums keep the sign."""
    say("bp alpha=-3.7")
    left, right = _bp_pair(9, 96, 72, 24)
    vo, wo = O.cost_volume(left, right, 24)
    po = O.Pyramid(left)
    bo = O.BP(po, vo, wo, 24, min_level=2, alpha=-3.7)
    for _ in range(6):
        bo.step(1)
    bg = ctx.bp(ctx.pyramid(left), ctx.cost_volume(left, right, 24), alpha=-3.7, min_level=2)
    for _ in range(6):
        bg.step(1)
    so, sg = bo.state(), bg.state()
    assert same_bits(so["cost"], sg["cost"]) and same_bits(so["belief"], sg["belief"])
    bo.step()
    do, plo = bo.map()
    dg, plg = bg.run()
    assert same_bits(do, dg) and same_bits(plo, plg)


def test_plain_path_cfg1_bit_exact(ctx):
    """450x375, D=64, reference defaults (alpha 1000.1, tau_s 0.6, min_level 2)."""
    say("plain path 450x375 D=64")
    left, right, gt = synth.make_pair(1, **{k: v for k, v in synth.PLAIN_SMALL.items()})
    t = time.time()
    do, plo, L, ml = O.plain_pair(left, right, 64)
    say(f"  oracle {time.time() - t:.1f}s levels={L}")
    vg = ctx.cost_volume(left, right, 64)
    pg = ctx.pyramid(left)
    t = time.time()
    dg, plg = ctx.bp(pg, vg, min_level=2).run()
    say(f"  gpu bp {time.time() - t:.3f}s")
    frac = mismatch(do, dg).mean()
    say(f"  disparity mismatch fraction {frac:.6f}")
    assert frac == 0.0
    assert same_bits(plo, plg)


def test_bp_is_deterministic(ctx):
    left, right = _bp_pair(4, 120, 90, 32)
    vg = ctx.cost_volume(left, right, 32)
    pg = ctx.pyramid(left)
    d1, p1 = ctx.bp(pg, vg, min_level=2).run()
    d2, p2 = ctx.bp(pg, vg, min_level=2).run()
    assert same_bits(d1, d2) and same_bits(p1, p2)


def test_large_pair_bp_state_bit_exact(ctx):
    """1472x984, D=256 (the benchmark pair's shape): BP state after the 530 first sweeps (all of
    level L-2 and the init of level L-3), and a full run twice (repeatable)."""
    say("large pair bp state")
    left, right, _ = synth.make_pair(11, 1472, 984, 256, (20, 40))
    vo, wo = O.cost_volume(left, right, 256)
    po = O.Pyramid(left)
    ml = max(1, po.levels - 5)
    bo = O.BP(po, vo, wo, 256, min_level=ml)
    vg = ctx.cost_volume(left, right, 256)
    pg = ctx.pyramid(left)
    bg = ctx.bp(pg, vg, min_level=ml)
    n = 530
    assert bo.step(n) == n and bg.step(n) == n
    so, sg = bo.state(), bg.state()
    say(f"  level {so['level']} V={so['V']}")
    assert so["level"] == sg["level"]
    for key in ("cost", "belief", "msg"):
        assert same_bits(so[key], sg[key]), key
    assert same_bits(so["planes"][:8], sg["planes"][:8])
    d1, p1 = ctx.bp(pg, vg, min_level=ml).run()
    d2, p2 = ctx.bp(pg, vg, min_level=ml).run()
    assert same_bits(d1, d2) and same_bits(p1, p2)


def test_bp_rejects_bad_min_level(ctx):
    left, right = _bp_pair(4, 64, 48, 16)
    vg = ctx.cost_volume(left, right, 16)
    pg = ctx.pyramid(left)
    with pytest.raises(H.HbpError):
        ctx.bp(pg, vg, min_level=0)
    with pytest.raises(H.HbpError):
        ctx.bp(pg, vg, min_level=pg.levels)


# ---------------------------------------------------------------- (d)
@pytest.mark.parametrize("seed,W,H,D", [(1, 64, 48, 16), (5, 120, 90, 32), (7, 101, 67, 20), (8, 217, 151, 27)])
def test_left_right_check_and_fill_bit_exact(ctx, seed, W, H, D):
    say(f"lr check seed={seed} {W}x{H}")
    left, right = _bp_pair(seed, W, H, D)
    # left view
    vo, wo = O.cost_volume(left, right, D)
    bol = O.BP(O.Pyramid(left), vo, wo, D, min_level=2)
    bol.step()
    dl, pl = bol.map()
    # right view: pyramid of the right image, right-view volume
    vr, wr = O.cost_volume(left, right, D, right_view=True)
    bor = O.BP(O.Pyramid(right), vr, wr, D, min_level=2)
    bor.step()
    dr, pr = bor.map()
    # GPU right view
    gr = ctx.bp(ctx.pyramid(right), ctx.cost_volume(left, right, D, right_view=True), min_level=2)
    dgr, pgr = gr.run()
    assert same_bits(dr, dgr) and same_bits(pr, pgr)
    # oracle check + fill
    plo = pl.copy()
    fl, fr, f2 = (np.zeros((H, W), np.uint8) for _ in range(3))
    O.lib.orc_left_right_check(dl, dr, plo, pr.copy(), W, H, fl, fr, f2)
    dispo = np.zeros((H, W), np.float32)
    O.lib.orc_plane_to_disp(plo, W, H, dispo)
    plg, flg, frg, f2g = ctx.lr_check_fill(dl, dr, pl)
    assert np.array_equal(fl, flg) and np.array_equal(fr, frg) and np.array_equal(f2, f2g)
    assert same_bits(plo, plg)
    assert same_bits(dispo, ctx.plane_to_disp(plg))


# ---------------------------------------------------------------- batched BP
def _batch_inputs(ctx, specs, right_view=False):
    out = []
    for seed, W, H, D, ml in specs:
        left, right = _bp_pair(seed, W, H, D)
        vol = ctx.cost_volume(left, right, D, right_view)
        pyr = ctx.pyramid(right if right_view else left)
        out.append((pyr, vol, ml))
    return out


def _check_batch(ctx, items, pre_steps=None, **kw):
    singles = [ctx.bp(p, v, min_level=ml, **kw).run() for p, v, ml in items]
    bps = [ctx.bp(p, v, min_level=ml, **kw) for p, v, ml in items]
    if pre_steps:
        for b, n in zip(bps, pre_steps):
            b.step(n)
    ctx.bp_run_batch(bps)
    for i, (b, (d1, p1)) in enumerate(zip(bps, singles)):
        d2, p2 = b.run()
        assert same_bits(d1, d2), f"pair {i}: {mismatch(d1, d2).mean():.4%} of disparities differ"
        assert same_bits(p1, p2), f"pair {i}: planes differ"


def test_batch_equals_single_runs(ctx):
    """Sweep i of every pair in the same launches: each pair's output is bit-identical to its
    own run (sizes, disparity ranges and min_level differ, so levels and schedules differ)."""
    say("batch of 6 mixed pairs")
    specs = [(21, 64, 48, 16, 2), (22, 96, 72, 24, 1), (23, 120, 90, 32, 2), (24, 160, 120, 48, 3),
             (25, 80, 200, 32, 2), (26, 200, 64, 40, 1)]
    _check_batch(ctx, _batch_inputs(ctx, specs))


def test_batch_partially_stepped(ctx):
    say("batch of partially stepped pairs")
    specs = [(31, 96, 72, 24, 2), (32, 120, 90, 32, 2), (33, 64, 48, 16, 1)]
    _check_batch(ctx, _batch_inputs(ctx, specs), pre_steps=[0, 7, 30])


def test_batch_right_view_and_wta_costs(ctx):
    say("batch right view / use_cost_vol=0")
    specs = [(41, 96, 72, 24, 2), (42, 120, 90, 32, 2)]
    _check_batch(ctx, _batch_inputs(ctx, specs, right_view=True))
    _check_batch(ctx, _batch_inputs(ctx, specs), use_cost_vol=0, alpha=1.3)


def test_batch_larger_pairs(ctx):
    """Coarse (segment) and fine (chunk) plane-cost forms in the same launches."""
    say("batch of larger pairs")
    specs = [(51, 640, 480, 64, 2), (52, 320, 240, 64, 2), (53, 1280, 720, 128, 4)]
    _check_batch(ctx, _batch_inputs(ctx, specs))


def test_batch_odd_sizes(ctx):
    """Odd widths, heights and disparity counts (tile and segment edges) in the same launches."""
    say("batch of odd-sized pairs")
    specs = [(71, 101, 67, 20, 1), (72, 203, 131, 37, 2), (73, 9, 7, 5, 1), (74, 333, 217, 3, 2)]
    _check_batch(ctx, _batch_inputs(ctx, specs))


def test_batch_rejects_mixed_cost_forms(ctx):
    left, right = _bp_pair(61, 64, 48, 16)
    a = ctx.bp(ctx.pyramid(left), ctx.cost_volume(left, right, 16), min_level=2)
    b = ctx.bp(ctx.pyramid(left), ctx.cost_volume(left, right, 16), min_level=2, use_cost_vol=0)
    with pytest.raises(H.HbpError):
        ctx.bp_run_batch([a, b])


# ---------------------------------------------------------------- request server
def _single(ctx, left, right, D, ml):
    pyr = ctx.pyramid(left)
    ml = max(1, min(ml, pyr.levels - 1))
    return ctx.bp(pyr, ctx.cost_volume(left, right, D), min_level=ml).run()


def test_server_results_equal_single_runs(ctx):
    """Pairs of different sizes submitted at different times, at most 3 on the device at once:
    every result equals the pair's own run (each result the same alone and in the mix)."""
    say("server: 9 pairs, 3 slots, staggered submits")
    specs = [(71, 96, 72, 24, 2), (72, 160, 120, 48, 2), (73, 64, 48, 16, 1), (74, 200, 64, 40, 2),
             (75, 120, 90, 32, 3), (76, 80, 200, 32, 2), (77, 128, 96, 32, 2), (78, 101, 67, 19, 1),
             (79, 9, 7, 3, 1)]
    pairs = [(_bp_pair(s, W, H, D), D, ml) for s, W, H, D, ml in specs]
    srv = H.Server(ctx, max_active=3)
    tickets = {}
    steps = 0
    n_done = 0
    pending = list(range(len(pairs)))
    while pending or any(t is not None for t in tickets.values()):
        # one more submit every 40 steps
        if pending and steps % 40 == 0:
            i = pending.pop(0)
            (left, right), D, ml = pairs[i]
            tickets[i] = srv.submit(left, right, D, ml)
        active, queued, done = srv.step()
        assert active <= 3 and done >= 0
        n_done += done
        steps += 1
        for i, t in list(tickets.items()):
            if t is None:
                continue
            r = srv.poll(t)
            if r is not None:
                (left, right), D, ml = pairs[i]
                d1, p1 = _single(ctx, left, right, D, ml)
                assert same_bits(d1, r[0]), f"pair {i}: disparity differs"
                assert same_bits(p1, r[1]), f"pair {i}: planes differ"
                tickets[i] = None
        assert steps < 200000
    assert n_done == len(pairs)
    srv.close()


def test_server_rejects_what_does_not_fit(ctx):
    say("server: budget")
    (left, right) = _bp_pair(81, 96, 72, 24)
    (l2, r2) = _bp_pair(82, 64, 48, 16)
    srv = H.Server(ctx, max_active=4, hbm_budget=2 << 20)   # 2 MiB: fits 64x48 only
    big = srv.submit(left, right, 24)
    small = srv.submit(l2, r2, 16)
    res = None
    for _ in range(100000):
        srv.step()
        res = srv.poll(small)
        if res is not None:
            break
    assert res is not None
    d1, p1 = _single(ctx, l2, r2, 16, 2)
    assert same_bits(d1, res[0]) and same_bits(p1, res[1])
    with pytest.raises(H.HbpError):
        srv.poll(big)
    srv.close()


@pytest.mark.parametrize("sched", [(1, 1, 1), (2, 3, 2), (5, 8, 8)])
def test_pyramid_under_every_crtrees_schedule(ctx, sched):
    """The scheduled CRTrees phases run the schedule the context learned from its previous levels
    (crt_rounds, crt_jumps, crt_max_jumps); a level that needs more steps overflows it and runs
    again with the full schedule.  Every starting schedule gives the oracle's levels, and the
    context's schedule afterwards stays inside the bounds."""
    rounds, jumps, max_jumps = sched
    say(f"pyramid 640x480 under CRTrees schedule {sched}")
    left, _, _ = synth.make_pair(12, 640, 480, 64, (6, 12))
    po = O.Pyramid(left)
    before = {k: ctx.get_tuning(k) for k in ("crt_rounds", "crt_jumps", "crt_max_jumps")}
    try:
        for _ in range(2):   # the second build starts from the schedule the first one learned
            ctx.set_tuning("crt_rounds", rounds)
            ctx.set_tuning("crt_jumps", jumps)
            ctx.set_tuning("crt_max_jumps", max_jumps)
            pg = ctx.pyramid(left)
            assert pg.levels == po.levels
            for l in range(po.levels):
                a, b = po.level(l), pg.level(l)
                assert (a["V"], a["E"]) == (b["V"], b["E"]), l
                assert np.array_equal(a["clus"], b["clus"]), l
                assert np.array_equal(a["bd"], b["bd"]), l
                assert same_bits(a["mean"], b["mean"]), l
            assert 1 <= ctx.get_tuning("crt_rounds") <= 5
            assert 1 <= ctx.get_tuning("crt_jumps") <= 8
            assert 1 <= ctx.get_tuning("crt_max_jumps") <= 8
            rounds, jumps, max_jumps = (ctx.get_tuning(k) for k in ("crt_rounds", "crt_jumps", "crt_max_jumps"))
    finally:
        for k, v in before.items():
            ctx.set_tuning(k, v)
