"""The Gaussian pre-blur of the superpixel pyramid (SegHAC's sigma > FLT_MIN; segment.cu:795-809,
imutil.cu:399-498) and the Lab conversion of blurred, non-integer pixels (imutil.cu:4-40).

Pins, each independent of the oracle's code:
  * t^2.4 (pin_pow24) against libm's pow over the whole linearisation domain;
  * the weights against a Python statement of compute_gauss_kernel with math.exp (libm, as the
    reference's host code), bit for bit, and against the closed form of a normalised Gaussian;
  * the blur against a numpy statement of the f32 fold in the reference's (r, c) order, bit
    for bit, and against scipy's float64 correlation with nearest-edge padding (the textbook
    blur) to 1e-3;
  * the float Lab conversion against the u8 one (libm pow through the 256-entry table) on
    integer-valued pixels;
  * the window limit round(3 sigma) <= 8 of the reference's launch, and sigma <= FLT_MIN = no blur.
"""
import math

import numpy as np
import pytest

import oracle_ffi as O
from hbp_amd import synth

FLT_MIN = 1.17549435e-38


def c_round(x):
    """C's round() for x >= 0 (half away from zero)"""
    return int(math.floor(x + 0.5))


def test_pow24_within_4_ulp_of_libm_over_the_domain():
    lo = (0.04045 + 0.055) / 1.055
    ts = np.concatenate([np.linspace(lo, 1.0, 200001), [np.nextafter(lo, 2.0), 1.0, 1.0 + 2.0 ** -21]])
    worst = 0.0
    for t in ts:
        got = O.lib.orc_pin_pow24(float(t))
        ref = math.pow(float(t), 2.4)
        worst = max(worst, abs(got - ref) / np.spacing(ref))
    assert worst <= 4.0
    assert O.lib.orc_pin_pow24(1.0) == 1.0


def test_pow24_is_monotone_on_a_fine_grid():
    ts = np.linspace(0.5, 0.5 + 1e-9, 2001)
    v = np.array([O.lib.orc_pin_pow24(float(t)) for t in ts])
    assert np.all(np.diff(v) >= 0)


@pytest.mark.parametrize("sigma", [0.1, 0.5, 0.7, 1.0, 1.5, 2.0, 2.83])
def test_weights_match_compute_gauss_kernel(sigma):
    S, w = O.gauss_weights(sigma)
    assert S == c_round(3 * sigma)
    F = 2 * S + 1
    f = [[0.0] * F for _ in range(F)]
    total = 0.0
    for r in range(-S, S + 1):
        for c in range(-S, S + 1):
            v = math.exp(-float(r * r + c * c) / (2 * sigma * sigma))
            f[r + S][c + S] = v
            total += v
    inv = 1.0 / total
    ref = np.array([[np.float32(f[i][j] * inv) for j in range(F)] for i in range(F)], np.float32)
    assert np.array_equal(w.view(np.uint32), ref.view(np.uint32))
    # closed form: a normalised, separable, symmetric Gaussian
    assert abs(float(w.astype(np.float64).sum()) - 1.0) < 1e-6
    assert np.array_equal(w, w.T) and np.array_equal(w, w[::-1, :])
    g = np.exp(-np.arange(-S, S + 1) ** 2 / (2 * sigma * sigma))
    outer = np.outer(g, g) / np.outer(g, g).sum()
    assert np.allclose(w, outer, rtol=1e-6, atol=1e-9)


def numpy_blur(bgr, w):
    """the reference's fold: acc += in(clamp(y+r), clamp(x+c)) * w[r][c], r outer, c inner, f32"""
    F = w.shape[0]
    S = F // 2
    H, W = bgr.shape[:2]
    src = np.concatenate([bgr.astype(np.float32), np.ones((H, W, 1), np.float32)], axis=2)
    pad = np.pad(src, ((S, S), (S, S), (0, 0)), mode="edge")
    acc = np.zeros((H, W, 4), np.float32)
    for r in range(F):
        for c in range(F):
            acc = acc + pad[r:r + H, c:c + W] * w[r, c]
    return acc


@pytest.mark.parametrize("sigma,W,H", [(0.1, 23, 17), (0.7, 45, 31), (1.0, 64, 48), (2.83, 37, 29)])
def test_blur_matches_the_literal_fold_and_the_textbook_blur(sigma, W, H):
    from scipy import ndimage
    left, _, _ = synth.make_pair(3, W, H, 8, (3, 6))
    S, out = O.gauss_blur(left, sigma)
    _, w = O.gauss_weights(sigma)
    ref = numpy_blur(left, w)
    assert np.array_equal(out.view(np.uint32), ref.view(np.uint32))
    for ch in range(3):
        tb = ndimage.correlate(left[:, :, ch].astype(np.float64), w.astype(np.float64), mode="nearest")
        assert np.max(np.abs(out[:, :, ch] - tb)) < 1e-3
    # w channel: the same fold of the weights at every pixel
    wsum = np.float32(0)
    for v in w.ravel():
        wsum = np.float32(wsum + np.float32(1.0) * v)
    assert np.all(out[:, :, 3] == wsum)


def test_sigma_limit_of_the_reference_launch():
    left, _, _ = synth.make_pair(5, 20, 16, 8, (3, 6))
    w = np.zeros(289, np.float32)
    assert O.lib.orc_gauss_weights(2.83, w) == 8
    assert O.lib.orc_gauss_weights(2.84, w) == -1
    out = np.zeros((16, 20, 4), np.float32)
    assert O.lib.orc_gauss_blur(np.ascontiguousarray(left), 20, 16, 3.0, out) == -1
    with pytest.raises(AssertionError, match="sigma"):
        O.Pyramid(left, sigma=3.0)


def test_float_lab_equals_u8_lab_on_integer_pixels():
    rng = np.random.default_rng(7)
    bgr = rng.integers(0, 256, size=(64, 64, 3), dtype=np.uint8)
    bgr[0, :8] = [[v, v, v] for v in (0, 1, 10, 11, 12, 254, 255, 128)]
    H, W = bgr.shape[:2]
    ref = np.zeros((H, W, 4), np.float32)
    O.lib.orc_lab(bgr, W, H, ref)
    f4 = np.concatenate([bgr.astype(np.float32), np.ones((H, W, 1), np.float32)], axis=2)
    got = O.lab_f(f4)
    same = got.view(np.uint32) == ref.view(np.uint32)
    assert same.mean() > 0.999
    # within one f32 ulp; a and b of grey pixels are 0 up to the FP64 rounding of x - y
    tol = np.maximum(np.spacing(np.abs(ref)), 1e-12)
    assert np.all(np.abs(got - ref) <= tol)
    assert np.all(got[:, :, 3] == 1.0)


def test_sigma_at_or_below_flt_min_is_no_blur():
    left, _, _ = synth.make_pair(6, 40, 30, 8, (3, 6))
    a, b = O.Pyramid(left), O.Pyramid(left, sigma=FLT_MIN)
    assert a.levels == b.levels
    for l in range(a.levels):
        x, y = a.level(l), b.level(l)
        assert np.array_equal(x["clus"], y["clus"]) and np.array_equal(x["mean"], y["mean"])


def test_blurred_pyramid_level0_is_the_blurred_lab():
    sigma = 1.0
    left, _, _ = synth.make_pair(8, 48, 36, 8, (3, 6))
    H, W = left.shape[:2]
    _, blur = O.gauss_blur(left, sigma)
    lab = O.lab_f(blur)
    p = O.Pyramid(left, sigma=sigma)
    m = p.level(0)["mean"]
    assert np.array_equal(m[0].view(np.uint32), lab[:, :, 0].ravel().view(np.uint32))
    assert np.array_equal(m[1].view(np.uint32), lab[:, :, 1].ravel().view(np.uint32))
    assert np.array_equal(m[2].view(np.uint32), lab[:, :, 2].ravel().view(np.uint32))
    assert np.array_equal(m[3], np.tile(np.arange(W, dtype=np.float32), H))
    assert np.array_equal(m[5], blur[:, :, 3].ravel())
    # the blur smooths: fewer level-1 clusters than without it on a noisy image
    rng = np.random.default_rng(1)
    noisy = np.clip(left.astype(np.int16) + rng.integers(-40, 41, size=left.shape), 0, 255).astype(np.uint8)
    v_plain = O.Pyramid(noisy).level(1)["V"]
    v_blur = O.Pyramid(noisy, sigma=1.5).level(1)["V"]
    assert v_blur < v_plain
    # the pyramid still ends at one cluster, and the weights sum level to level
    top = p.level(p.levels - 1)
    assert np.all(top["clus"] == 0)
    total = float(blur[:, :, 3].astype(np.float64).sum())
    assert abs(float(top["mean"][5].astype(np.float64).sum()) - total) < 1e-4 * total
