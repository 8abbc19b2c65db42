"""GPU parity of the refinements after the left-right check (C ABI of libhbp_hip.so against the
oracle, bit for bit): the weighted median of the failed pixels' planes (post_process.cu:127-185)
and the CLI's 7-tap median (main_stereo_hbp_isp.cu:120-193)."""
import sys

import numpy as np
import pytest

import oracle_ffi as O

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)
    sys.stdout.flush()


def same_bits(a, b):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    na, nb = np.isnan(a), np.isnan(b)
    return a.shape == b.shape and np.array_equal(na, nb) and np.array_equal(a.view(np.uint32)[~na],
                                                                            b.view(np.uint32)[~nb])


def case(seed, W, H, nlabels=9, frac=0.15, specials=False):
    rng = np.random.default_rng(seed)
    bgr = np.clip(rng.normal(128, 50, (H, W, 3)), 0, 255).astype(np.uint8)
    bgr[:, W // 2:] = bgr[:, W // 2:] // 2      # an edge in colour
    labels = rng.integers(0, nlabels, (H // 5 + 1, W // 5 + 1)).repeat(5, 0).repeat(5, 1)[:H, :W]
    planes = (rng.normal(0, 1, (nlabels, 3)) * [0.05, 0.05, 12] + [0, 0, 40]).astype(np.float32)
    plane = planes[labels].astype(np.float32)
    if specials:
        plane[labels == 0] = np.float32(np.nan)
        plane[labels == 1] = np.float32([-0.0, 0.0, -0.0])
        plane[labels == 2, 2] = np.float32(np.inf)
    fail = (rng.random((H, W)) < frac).astype(np.uint8) * 255
    fail[0, 0] = fail[H - 1, W - 1] = fail[0, W - 1] = 255
    return bgr, fail, plane


@pytest.mark.parametrize("seed,W,H,specials", [(1, 96, 72, False), (2, 23, 17, False), (3, 160, 120, True),
                                               (4, 41, 41, True), (5, 300, 40, False)])
def test_weighted_median_bit_exact(ctx, seed, W, H, specials):
    say(f"weighted median seed={seed} {W}x{H} specials={specials}")
    bgr, fail, plane = case(seed, W, H, specials=specials)
    ref = plane.copy()
    O.lib.orc_weighted_median(bgr, fail, ref, W, H)
    got = ctx.weighted_median(bgr, fail, plane)
    assert same_bits(got, ref)
    assert same_bits(got[fail == 0], plane[fail == 0])


def test_weighted_median_nothing_failed(ctx):
    bgr, fail, plane = case(7, 64, 48)
    fail[:] = 0
    assert same_bits(ctx.weighted_median(bgr, fail, plane), plane)


@pytest.mark.parametrize("seed,W,H", [(1, 96, 72), (2, 7, 7), (3, 6, 9), (4, 333, 211)])
def test_median7_bit_exact(ctx, seed, W, H):
    say(f"median7 seed={seed} {W}x{H}")
    rng = np.random.default_rng(seed)
    D = (rng.random((H, W)) * 60).astype(np.float32)
    D[rng.random((H, W)) < 0.05] = np.float32(np.nan)
    D[rng.random((H, W)) < 0.05] *= np.float32(-1)
    D[rng.random((H, W)) < 0.02] = np.float32(np.inf)
    ref = D.copy()
    O.lib.orc_median7(ref, W, H)
    assert same_bits(ctx.median7(D), ref)
