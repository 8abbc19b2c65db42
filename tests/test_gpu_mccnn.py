"""GPU parity of the f32 costs, through the C ABI against the oracle, bit for bit: the "MC-CNN"
cost (a given f32 volume: compute_stereo_cost, stereo_cost.cu:399-445; cost_to_right, 471-516),
the "AD" cost (ad_cost, stereo_cost.cu:9-43), and BP on both."""
import sys

import numpy as np
import pytest

import oracle_ffi as O
from hbp_amd import synth

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_bits(a, b):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    return a.shape == b.shape and np.array_equal(na, nb) and np.array_equal(a.view(np.uint32)[~na],
                                                                            b.view(np.uint32)[~nb])


def mccnn_like(seed, W, H, D):
    """a volume in the range of matching-network costs: the census costs scaled to [0, 1] with
    noise, a few NaNs and values above the truncation"""
    left, right, _ = synth.make_pair(seed, W, H, D, (3, 8))
    vol, _ = O.cost_volume(left, right, D)
    rng = np.random.default_rng(seed)
    raw = (vol.astype(np.float32) / np.float32(25.0) + rng.normal(0, 0.05, vol.shape).astype(np.float32))
    raw[rng.random(raw.shape) < 0.001] = np.float32(np.nan)
    return left, right, raw.astype(np.float32)


@pytest.mark.parametrize("right_view", [False, True])
def test_f32_volume_and_wta(ctx, right_view):
    say(f"mc-cnn volume right={right_view}")
    _, _, raw = mccnn_like(3, 96, 64, 32)
    vo, wo = O.cost_volume_f32(raw, 0.5, right_view=right_view)
    vg, wg = ctx.cost_volume_f32(raw, right_view=right_view).download_f32()
    assert same_bits(vg, vo) and same_bits(wg, wo)


@pytest.mark.parametrize("seed,W,H,D,ml,right_view", [(4, 96, 72, 24, 2, False), (5, 128, 96, 32, 1, False),
                                                      (6, 120, 90, 32, 2, True), (9, 103, 71, 21, 1, False)])
def test_bp_on_f32_volume_bit_exact(ctx, seed, W, H, D, ml, right_view):
    say(f"mc-cnn bp seed={seed} {W}x{H} D={D} min_level={ml} right={right_view}")
    left, right, raw = mccnn_like(seed, W, H, D)
    img = right if right_view else left
    vo, wo = O.cost_volume_f32(raw, 0.5, right_view=right_view)
    po = O.Pyramid(img)
    bo = O.BP(po, vo, wo, D, min_level=ml)
    vg = ctx.cost_volume_f32(raw, right_view=right_view)
    pg = ctx.pyramid(img)
    bg = ctx.bp(pg, vg, min_level=ml)
    for sweep in range(6):
        assert bo.step(1) == 1 and bg.step(1) == 1
        so, sg = bo.state(), bg.state()
        assert same_bits(so["belief"], sg["belief"]), f"belief differs after sweep {sweep}"
        assert same_bits(so["planes"][:8], sg["planes"][:8]), f"planes differ after sweep {sweep}"
    bo.step()
    do, plo = bo.map()
    dg, plg = bg.run()
    assert same_bits(do, dg) and same_bits(plo, plg)


def test_batch_on_f32_volumes_equals_single_runs(ctx):
    say("mc-cnn batch")
    singles, bps, keep = [], [], []
    for seed, W, H, D in [(7, 80, 60, 16), (8, 112, 80, 24)]:
        left, _, raw = mccnn_like(seed, W, H, D)
        v, p = ctx.cost_volume_f32(raw), ctx.pyramid(left)
        singles.append(ctx.bp(p, v, min_level=2).run())
        v2, p2 = ctx.cost_volume_f32(raw), ctx.pyramid(left)
        keep += [v2, p2]
        bps.append(ctx.bp(p2, v2, min_level=2))
    ctx.bp_run_batch(bps)
    for (d1, p1), bp in zip(singles, bps):
        d2, p2 = bp.run()
        assert same_bits(d1, d2) and same_bits(p1, p2)


# ---- the "AD" cost ----

@pytest.mark.parametrize("seed,W,H,D,right_view", [(1, 96, 64, 32, False), (2, 37, 23, 5, True),
                                                   (3, 450, 375, 64, False), (4, 1030, 40, 256, True)])
def test_ad_volume_and_wta(ctx, seed, W, H, D, right_view):
    say(f"ad volume seed={seed} {W}x{H} D={D} right={right_view}")
    left, right, _ = synth.make_pair(seed, W, H, D, (3, 6))
    vo, wo = O.cost_volume_ad(left, right, D, right_view=right_view)
    vg, wg = ctx.cost_volume(left, right, D, right_view=right_view, cost="AD").download_f32()
    assert same_bits(vg, vo) and same_bits(wg, wo)


@pytest.mark.parametrize("seed,W,H,D,ml,right_view", [(5, 96, 72, 24, 2, False), (6, 120, 90, 32, 1, True),
                                                      (8, 101, 67, 19, 1, True)])
def test_bp_on_ad_volume_bit_exact(ctx, seed, W, H, D, ml, right_view):
    say(f"ad bp seed={seed} {W}x{H} D={D} min_level={ml} right={right_view}")
    left, right, _ = synth.make_pair(seed, W, H, D, (3, 8))
    img = right if right_view else left
    vo, wo = O.cost_volume_ad(left, right, D, right_view=right_view)
    bo = O.BP(O.Pyramid(img), vo, wo, D, min_level=ml, tau=30.0)
    bo.step()
    do, plo = bo.map()
    vg = ctx.cost_volume(left, right, D, right_view=right_view, cost="AD")
    dg, plg = ctx.bp(ctx.pyramid(img), vg, min_level=ml).run()
    assert same_bits(do, dg) and same_bits(plo, plg)


def test_unknown_cost_name_is_an_error(ctx):
    left, right, _ = synth.make_pair(7, 32, 24, 8, (3, 6))
    with pytest.raises(Exception, match="Census"):
        ctx.cost_volume(left, right, 8, cost="SAD")
