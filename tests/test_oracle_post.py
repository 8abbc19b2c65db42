"""Oracle pins for the refinements after the left-right check: the weighted median of planes
(post_process.cu:127-185) and the CLI's 7-tap separable median (main_stereo_hbp_isp.cu:120-193);
and for the "MC-CNN" cost (a given f32 volume, stereo_cost.cu:399-445).  Each oracle function is
checked against an independent statement of the reference's own arithmetic (double sums in the
reference's orders, libm exp, the volume rules), written here in numpy, or against the pinned
census path."""
import math

import numpy as np
import pytest

import oracle_ffi as O


def test_wm_weight_is_the_rounded_exp_for_every_possible_input():
    # |dB| + |dG| + |dR| of u8 images takes the integer values 0..765 only
    bad = []
    for a in range(766):
        t = np.float32(-np.float32(a) / np.float32(10.0))
        ref = np.float32(math.exp(float(t)))
        got = np.float32(O.lib.orc_wm_weight(float(a)))
        if got != ref:
            bad.append((a, got, ref))
    assert not bad, bad[:5]


def reference_wm(bgr, fail, plane):
    """post_process.cu:127-185 as written: weights in float, their sum in double in row-major
    order, a stable sort by z, the running double sum compared with half the total."""
    H, W = fail.shape
    I = bgr.astype(np.float32)
    out = plane.copy()
    near = 0
    for y in range(H):
        for x in range(W):
            if not fail[y, x]:
                continue
            y0, y1, x0, x1 = max(y - 20, 0), min(y + 20, H - 1), max(x - 20, 0), min(x + 20, W - 1)
            d = np.abs(I[y, x] - I[y0:y1 + 1, x0:x1 + 1]).astype(np.float32)
            absdiff = (d[..., 0] + d[..., 1]) + d[..., 2]
            t = (-absdiff / np.float32(10.0)).astype(np.float32)
            w = np.exp(t.astype(np.float64)).astype(np.float32).ravel()
            pl = plane[y0:y1 + 1, x0:x1 + 1].reshape(-1, 3)
            z = ((pl[:, 0] * np.float32(x) + pl[:, 1] * np.float32(y)) + pl[:, 2]).astype(np.float32)
            sumw = 0.0
            for v in w.tolist():
                sumw += v
            center = sumw / 2.0
            order = np.argsort(z, kind="stable")
            c = 0.0
            for j in order.tolist():
                c += float(w[j])
                if c > center:
                    if abs(c - center) < 1e-9 or abs(c - float(w[j]) - center) < 1e-9:
                        near += 1
                    out[y, x] = pl[j]
                    break
    return out, near


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_weighted_median_matches_the_reference_arithmetic(seed):
    rng = np.random.default_rng(seed)
    H, W = 36, 52
    # smooth colours with a few edges, planes from a handful of labels (many equal z)
    bgr = np.clip(rng.normal(128, 40, (H, W, 3)).cumsum(axis=1) / np.arange(1, W + 1)[None, :, None],
                  0, 255).astype(np.uint8)
    labels = rng.integers(0, 7, (H // 6 + 1, W // 6 + 1)).repeat(6, 0).repeat(6, 1)[:H, :W]
    planes = rng.normal(0, 1, (7, 3)).astype(np.float32) * np.float32([0.05, 0.05, 10]) + np.float32([0, 0, 30])
    plane = planes[labels].astype(np.float32)
    fail = (rng.random((H, W)) < 0.2).astype(np.uint8) * 255
    ref, near = reference_wm(bgr, fail, plane)
    got = plane.copy()
    O.lib.orc_weighted_median(bgr, fail, got, W, H)
    assert near == 0
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_weighted_median_uniform_colour_is_the_plain_median():
    rng = np.random.default_rng(5)
    H, W = 30, 30
    bgr = np.full((H, W, 3), 77, np.uint8)
    plane = rng.normal(0, 1, (H, W, 3)).astype(np.float32)
    fail = np.zeros((H, W), np.uint8)
    fail[15, 15] = fail[0, 0] = fail[29, 3] = 255
    got = plane.copy()
    O.lib.orc_weighted_median(bgr, fail, got, W, H)
    for (y, x) in [(15, 15), (0, 0), (29, 3)]:
        y0, y1, x0, x1 = max(y - 20, 0), min(y + 20, H - 1), max(x - 20, 0), min(x + 20, W - 1)
        pl = plane[y0:y1 + 1, x0:x1 + 1].reshape(-1, 3)
        z = (pl[:, 0] * np.float32(x) + pl[:, 1] * np.float32(y)) + pl[:, 2]
        n = len(z)
        # equal weights: the first element whose count c satisfies 2c > n, i.e. c = n//2 + 1
        j = np.argsort(z, kind="stable")[n // 2]
        assert np.array_equal(got[y, x], pl[j])
    assert np.array_equal(got[fail == 0], plane[fail == 0])


def literal_median7(D):
    """main_stereo_hbp_isp.cu:120-193 line by line (insertion sort, small images only)."""
    H, W = D.shape
    D = D.copy()
    T = np.zeros_like(D)

    def isort(ws):
        vals = [0.0] * 7
        for j, t in enumerate(ws):
            i = j - 1
            while i >= 0 and vals[i] > t:
                vals[i + 1] = vals[i]
                i -= 1
            vals[i + 1] = t
        return vals[3]

    for u in range(3, W - 3):
        for v in range(3, H - 3):
            T[v, u] = isort([D[v, u2] for u2 in range(u - 3, u + 4)]) if D[v, u] >= 0 else D[v, u]
    for u in range(3, W - 3):
        for v in range(3, H - 3):
            if D[v, u] >= 0:
                D[v, u] = isort([T[v2, u] for v2 in range(v - 3, v + 4)])
    return D


def test_median7_literal_with_nan_and_negatives():
    rng = np.random.default_rng(3)
    D = rng.normal(20, 10, (17, 19)).astype(np.float32)
    D[rng.random(D.shape) < 0.1] = np.nan
    D[rng.random(D.shape) < 0.1] *= -1
    got = D.copy()
    O.lib.orc_median7(got, 19, 17)
    ref = literal_median7(D)
    assert np.array_equal(np.isnan(got), np.isnan(ref))
    assert np.array_equal(got[~np.isnan(got)], ref[~np.isnan(ref)])


def test_median7_is_the_separable_median_for_finite_values():
    rng = np.random.default_rng(4)
    H, W = 40, 50
    D = rng.random((H, W)).astype(np.float32) * 50
    got = D.copy()
    O.lib.orc_median7(got, W, H)
    T = np.zeros_like(D)
    win = np.lib.stride_tricks.sliding_window_view(D, 7, axis=1)          # [H][W-6][7]
    T[3:H - 3, 3:W - 3] = np.sort(win, axis=-1)[3:H - 3, :, 3]
    win2 = np.lib.stride_tricks.sliding_window_view(T, 7, axis=0)         # [H-6][W][7]
    ref = D.copy()
    ref[3:H - 3, 3:W - 3] = np.sort(win2, axis=-1)[:, 3:W - 3, 3]
    assert np.array_equal(got, ref)


# ---- the "MC-CNN" cost: a given f32 volume (stereo_cost.cu:399-445, 471-516) ----

def test_f32_volume_fill_truncate_wta_from_definition():
    rng = np.random.default_rng(11)
    D, H, W = 12, 9, 20
    raw = rng.random((D, H, W)).astype(np.float32) * 2 - 0.5
    raw[rng.random(raw.shape) < 0.02] = np.float32(np.nan)
    vol, wta = O.cost_volume_f32(raw, 0.5)
    ref = raw.copy()
    for d in range(D):
        ref[d, :, :d + 1] = raw[d, :, d:d + 1]           # columns x <= d take column d
    ref = np.fmin(ref, np.float32(0.5))                  # NaN -> 0.5
    assert np.array_equal(vol, ref)
    assert np.array_equal(wta, np.argmin(ref, axis=0).astype(np.float32))
    volr, wtar = O.cost_volume_f32(raw, 0.5, right_view=True)
    xs = np.arange(W)
    for d in range(D):
        assert np.array_equal(volr[d], ref[d][:, np.minimum(xs + d, W - 1)])
    assert np.array_equal(wtar, np.argmin(volr, axis=0).astype(np.float32))


def test_bp_on_an_f32_copy_of_the_census_volume_is_the_census_bp():
    from hbp_amd import synth
    left, right, _ = synth.make_pair(9, 80, 60, 24, (3, 6))
    vol, wta = O.cost_volume(left, right, 24)
    pyr = O.Pyramid(left)
    a = O.BP(pyr, vol, wta, 24, min_level=2)
    b = O.BP(pyr, vol.astype(np.float32), wta, 24, min_level=2, tau=25.0)
    a.step()
    b.step()
    da, pa = a.map()
    db, pb = b.map()
    assert np.array_equal(da.view(np.uint32), db.view(np.uint32))
    assert np.array_equal(pa.view(np.uint32), pb.view(np.uint32))


# ---- the "AD" cost (stereo_cost.cu:9-43, tau 30) ----

def numpy_grey_f32(bgr):
    """img_BGR_to_grey (imutil.cu:116-130): the double sum, stored as float"""
    b = bgr.astype(np.float64)
    return ((b[..., 0] * 0.114 + b[..., 1] * 0.587) + b[..., 2] * 0.299).astype(np.float32)


@pytest.mark.parametrize("right_view", [False, True])
def test_ad_volume_from_definition(right_view):
    rng = np.random.default_rng(21)
    H, W, D = 11, 30, 9
    left = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    right = np.clip(left.astype(np.int32) + rng.integers(-40, 41, (H, W, 3)), 0, 255).astype(np.uint8)
    vol, wta = O.cost_volume_ad(left, right, D, right_view=right_view)
    gl, gr = numpy_grey_f32(left), numpy_grey_f32(right)
    ref = np.zeros((D, H, W), np.float32)
    for d in range(D):
        for x in range(D - 1, W):
            ref[d, :, x] = np.minimum(np.abs(gl[:, x] - gr[:, x - d]), np.float32(30.0))
    if right_view:
        xs = np.minimum(np.arange(W)[None, :] + np.arange(D)[:, None], W - 1)
        ref = np.stack([ref[d][:, xs[d]] for d in range(D)])
    assert np.array_equal(vol, ref)
    assert np.array_equal(wta, np.argmin(ref, axis=0).astype(np.float32))
    # the border columns cost 0 at every disparity, so their WTA is 0 (left view)
    if not right_view:
        assert not vol[:, :, :D - 1].any() and not wta[:, :D - 1].any()
    assert vol.max() <= 30.0


def test_ad_grey_is_not_truncated():
    # the AD greys keep their fractions (the census path truncates them to u8)
    bgr = np.array([[[1, 2, 3]] * 12] * 9, np.uint8)
    other = bgr.copy()
    other[..., 2] = 4
    vol, _ = O.cost_volume_ad(bgr, other, 3)
    # |grey(1,2,3) - grey(1,2,4)| = 0.299 in float
    g1, g2 = numpy_grey_f32(bgr[:1, :1])[0, 0], numpy_grey_f32(other[:1, :1])[0, 0]
    assert vol[0, 0, 5] == abs(g1 - g2) and 0.29 < vol[0, 0, 5] < 0.31


def test_bp_on_ad_volume_uses_tau_30():
    from hbp_amd import synth
    left, right, _ = synth.make_pair(13, 72, 54, 16, (3, 6))
    vol, wta = O.cost_volume_ad(left, right, 16)
    pyr = O.Pyramid(left)
    a = O.BP(pyr, vol, wta, 16, min_level=2, tau=30.0)
    a.step()
    disp, plane = a.map()
    assert np.isfinite(disp).mean() > 0.99
    b = O.BP(pyr, vol, wta, 16, min_level=2, tau=0.5)       # another tau changes the result
    b.step()
    assert not np.array_equal(disp, b.map()[0])
