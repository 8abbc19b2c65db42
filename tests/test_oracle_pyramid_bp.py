"""CPU tests of the oracle's pyramid (b), BP (c) and LR (d) stages: closed forms, invariants and
independent library routines (scipy) that a plausible transcription mistake would break."""
import numpy as np
import pytest
from scipy import ndimage

import oracle_ffi as O
from hbp_amd import synth

L = O.lib


# ------------------------------------------------------------------------- Lab and 1-NN grid
def lab_of(bgr):
    H, W = bgr.shape[:2]
    out = np.zeros(H * W * 4, np.float32)
    L.orc_lab(np.ascontiguousarray(bgr), W, H, out)
    return out.reshape(H, W, 4)


def test_lab_known_colours():
    # white -> L=100, a=b=0; black -> L=0; sRGB red -> about (53.24, 80.09, 67.20) for D65
    bgr = np.array([[[255, 255, 255], [0, 0, 0], [0, 0, 255], [0, 255, 0]]], np.uint8)
    lab = lab_of(bgr)[0]
    assert abs(lab[0, 0] - 100) < 1e-3 and abs(lab[0, 1]) < 1e-2 and abs(lab[0, 2]) < 1e-2
    assert abs(lab[1, 0]) < 1e-6
    assert np.allclose(lab[2, :3], [53.24, 80.09, 67.20], atol=0.1)
    assert np.allclose(lab[3, :3], [87.73, -86.18, 83.18], atol=0.1)
    assert np.all(lab[:, 3] == 1.0)


def test_lab_against_float64_textbook_conversion():
    rng = np.random.default_rng(2)
    bgr = rng.integers(0, 256, (1, 500, 3), dtype=np.uint8)
    lab = lab_of(bgr)[0]
    c = bgr[0].astype(np.float64) / 255.0
    lin = np.where(c > 0.04045, ((c + 0.055) / 1.055) ** 2.4, c / 12.92)
    b, g, r = lin[:, 0], lin[:, 1], lin[:, 2]
    x = (r * 0.412453 + g * 0.357580 + b * 0.180423) / 0.950456
    y = r * 0.212671 + g * 0.715160 + b * 0.072169
    z = (r * 0.019334 + g * 0.119193 + b * 0.950227) / 1.088754
    f = lambda t: np.where(t > 216 / 24389.0, np.cbrt(t), 841 / 108.0 * t + 16 / 116.0)
    ref = np.stack([116 * f(y) - 16, 500 * (f(x) - f(y)), 200 * (f(y) - f(z))], 1)
    assert np.max(np.abs(lab[:, :3] - ref)) < 1e-4


def test_nn_grid_brute_force_and_tie_rule():
    rng = np.random.default_rng(4)
    H, W = 9, 11
    bgr = rng.integers(0, 256, (H, W, 3), dtype=np.uint8)
    lab = lab_of(bgr)
    nn = np.zeros(H * W, np.int32)
    dm = np.zeros(H * W, np.float32)
    bd = np.zeros(H * W * 8, np.int32)
    L.orc_nn_grid(lab.reshape(-1), W, H, 4, nn, dm, bd)
    bd = bd[: H * W * 8].reshape(-1, 2)[: H * W * 4]
    for i in range(H * W):
        y, x = divmod(i, W)
        cands = []
        for ok, j in ((y > 0, i - W), (y < H - 1, i + W), (x > 0, i - 1), (x < W - 1, i + 1)):
            if ok:
                s, t = lab.reshape(-1, 4)[i].astype(np.float64), lab.reshape(-1, 4)[j].astype(np.float64)
                d = ((s[:3] - t[:3]) ** 2).sum() + max(0.0, s[0] - t[0])
                cands.append((d, j))
        best = min(c[0] for c in cands)
        assert abs(dm[i] - best) <= 1e-3 * max(1.0, best)
        # first candidate within float rounding of the minimum
        assert nn[i] in [j for d, j in cands if abs(d - best) <= 1e-3 * max(1.0, best)]
        assert bd[4 * i][0] == i
    # constant image: all distances 0, so the first valid neighbour (up, else down) wins
    flat = lab_of(np.full((5, 6, 3), 100, np.uint8)).reshape(-1)
    nn2 = np.zeros(30, np.int32)
    L.orc_nn_grid(flat, 6, 5, 4, nn2, np.zeros(30, np.float32), np.zeros(30 * 8, np.int32))
    assert np.array_equal(nn2.reshape(5, 6)[0], np.arange(6) + 6)
    assert np.array_equal(nn2.reshape(5, 6)[1:], (np.arange(30) - 6).reshape(5, 6)[1:])


# ------------------------------------------------------------------------- CRTrees
def literal(nn, order):
    out = np.zeros(len(nn), np.int32)
    L.orc_crtrees_literal(nn, len(nn), order, out)
    return out


def components(nn):
    out = np.zeros(len(nn), np.int32)
    L.orc_cc_max(nn, len(nn), out)
    return out


@pytest.mark.parametrize("seed", range(6))
def test_crtrees_labels_weak_components_for_any_interleaving(seed):
    rng = np.random.default_rng(seed)
    for V in (2, 3, 17, 200, 3000):
        nn = rng.integers(0, V, V).astype(np.int32)
        ref = literal(nn, 0)
        for order in (-1, 1, 2, 3):           # snapshot, descending, two random orders
            assert np.array_equal(ref, literal(nn, order))
        comp = components(nn)
        # same partition: label constant on each weak component, distinct across components
        pairs = set(zip(comp.tolist(), ref.tolist()))
        assert len(pairs) == len(set(comp.tolist())) == len(set(ref.tolist()))
        # the representative lies in its component
        assert np.array_equal(comp[ref], comp)


def test_crtrees_representative_is_not_the_component_maximum():
    # 0 -> 1 -> 2 <-> 3: the branch node 2 (in-degree 2) represents the component
    assert list(literal(np.array([1, 2, 3, 2], np.int32), 0)) == [2, 2, 2, 2]
    # a pure 3-cycle is labelled with its maximum
    assert list(literal(np.array([1, 2, 0], np.int32), 0)) == [2, 2, 2]


def test_compact_ranks_labels():
    lab = np.array([5, 5, 2, 9, 2, 9, 9, 2, 5, 5], np.int32)
    n = L.orc_compact(lab, len(lab))
    assert n == 3 and list(lab) == [1, 1, 0, 2, 0, 2, 2, 0, 1, 1]


# ------------------------------------------------------------------------- pyramid
@pytest.fixture(scope="module")
def pyr():
    left, _, _ = synth.make_pair(3, 80, 60, 16, (3, 6))
    return left, O.Pyramid(left)


def test_pyramid_level_invariants(pyr):
    left, p = pyr
    H, W = left.shape[:2]
    N = H * W
    lv = [p.level(l) for l in range(p.levels)]
    assert lv[0]["V"] == N and lv[0]["E"] == 4 * N
    assert np.all(lv[0]["bdl"] == 1)
    # level 0 means: the pixel values (L, a, b, x, y, 1)
    lab = lab_of(left).reshape(-1, 4)
    assert np.array_equal(lv[0]["mean"][0], lab[:, 0])
    assert np.array_equal(lv[0]["mean"][3], (np.arange(N) % W).astype(np.float32))
    for l in range(p.levels):
        a = lv[l]
        V = a["V"]
        Vn = lv[l + 1]["V"] if l + 1 < p.levels else 1
        assert a["clus"].min() == 0 and a["clus"].max() == Vn - 1
        assert len(np.unique(a["clus"])) == Vn
        assert abs(a["mean"][5].sum() - N) < 1e-3 * N           # pixel counts add up
        if l >= 1:
            bd = a["bd"]
            keys = bd[:, 0].astype(np.int64) * V + bd[:, 1]
            assert np.all(np.diff(keys) > 0)                    # sorted, unique
            assert np.all(bd[:, 0] != bd[:, 1])                 # no self loops
            rev = set(map(tuple, bd[:, ::-1].tolist()))
            assert rev == set(map(tuple, bd.tolist()))          # symmetric
            # the bdl of level 1 count the pixel edges merged into each edge
            prev = lv[l - 1]
            m = prev["clus"][prev["bd"]]
            cross = m[:, 0] != m[:, 1]
            if l == 1:
                assert a["bdl"].sum() == cross.sum()
            # sums of children: x sums exact in float for this size
            sx = np.zeros(V)
            np.add.at(sx, prev["clus"], prev["mean"][3].astype(np.float64))
            assert np.allclose(a["mean"][3], sx, rtol=1e-6)
    # the last level maps everything to one cluster
    assert np.all(lv[-1]["clus"] == 0)


def test_pyramid_edges_match_set_semantics(pyr):
    # edges: the sorted set of mapped (s, t), s != t; bdl: the sum of the previous level's bdl
    # over each run's positions in the sorted order (the reference sorts the edges without their
    # bdl, segment.cu:1058-1078)
    _, p = pyr
    for l in range(p.levels - 1):
        a, b = p.level(l), p.level(l + 1)
        m = a["clus"][a["bd"]].astype(np.int64)
        order = np.lexsort((m[:, 1], m[:, 0]))
        sm = m[order]
        expect = {}
        for i, (s, t) in enumerate(sm.tolist()):
            if s != t:
                expect[(s, t)] = expect.get((s, t), 0) + int(a["bdl"][i])
        got = {(s, t): w for (s, t), w in zip(b["bd"].tolist(), b["bdl"].tolist())}
        assert got == expect


# ------------------------------------------------------------------------- BP
@pytest.fixture(scope="module")
def bp_case():
    left, right, gt = synth.make_pair(9, 80, 60, 16, (3, 6))
    vol, wta = O.cost_volume(left, right, 16)
    p = O.Pyramid(left)
    return left, right, gt, vol, wta, p


def test_bp_initial_costs_match_float64_plane_costs(bp_case):
    left, right, gt, vol, wta, p = bp_case
    bp = O.BP(p, vol, wta, 16, min_level=2)
    bp.step(1)
    s = bp.state()
    l = s["level"]
    # pixel -> cluster at level l
    im = np.arange(80 * 60)
    for k in range(l):
        im = p.level(k)["clus"][im]
    xs, ys = im % 1, None
    lam, tau = np.float32(1000.1), np.float32(25.0)
    X = (np.arange(80 * 60) % 80).astype(np.float32)
    Y = (np.arange(80 * 60) // 80).astype(np.float32)
    for c in range(s["V"]):
        pix = np.nonzero(im == c)[0]
        for slot in range(8):            # best and prior slots after the first sweep
            a, b, cc, _ = s["planes"][slot, c]
            df = (np.float32(a) * X[pix] + np.float32(b) * Y[pix]) + np.float32(cc)
            d = np.floor(df).astype(np.int64)
            ok = (d >= 0) & (d + 1 < 16)
            v = np.full(len(pix), float(tau))
            dd = d[ok]
            c0 = vol[dd, (pix[ok] // 80), pix[ok] % 80].astype(np.float64)
            c1 = vol[dd + 1, (pix[ok] // 80), pix[ok] % 80].astype(np.float64)
            v[ok] = (dd + 1 - df[ok].astype(np.float64)) * c0 + (df[ok].astype(np.float64) - dd) * c1
            ref = float(lam) * v.sum()
            got = s["cost"][slot, c]
            if np.isnan(got):
                assert np.isnan(a) or np.isnan(b) or np.isnan(cc)
            else:
                assert abs(got - ref) <= 1e-5 * max(1.0, abs(ref)), (c, slot, got, ref)


def test_bp_select_keeps_the_best_beliefs(bp_case):
    left, right, gt, vol, wta, p = bp_case
    bp = O.BP(p, vol, wta, 16, min_level=2)
    for _ in range(7):
        bp.step(1)
        s = bp.state()
        b = s["belief"]
        for c in range(s["V"]):
            col = b[:, c]
            if np.isnan(col).any():
                continue
            # slots 0..3 hold the 4 smallest of this sweep's 12 beliefs; slots 4..11 still hold
            # their own values, so each of them was either selected or is no smaller than
            # every kept belief
            kept = col[:4]
            for v in col[4:]:
                assert (v in kept) or v >= kept.max()


def test_bp_full_run_beats_wta_at_low_alpha(bp_case):
    left, right, gt, vol, wta, p = bp_case
    bp = O.BP(p, vol, wta, 16, min_level=2, alpha=0.05)
    bp.step()
    d, pl = bp.map()
    m = np.zeros_like(gt, bool)
    m[:, 16:] = True
    assert (np.abs(d - gt) > 2)[m].mean() < (np.abs(wta - gt) > 2)[m].mean()
    # disparities are the clamped plane evaluation
    X, Y = np.meshgrid(np.arange(80, dtype=np.float32), np.arange(60, dtype=np.float32))
    z = (pl[..., 0] * X + pl[..., 1] * Y) + pl[..., 2]
    assert np.array_equal(np.nan_to_num(np.clip(z, 0, 15), nan=0.0), d)


def test_bp_state_api_and_bounds(bp_case):
    left, right, gt, vol, wta, p = bp_case
    assert L.orc_bp_new(p.h, vol, wta, 80, 60, 16, 25.0, 1000.1, 0.6, 0, 1) is None
    assert L.orc_bp_new(p.h, vol, wta, 80, 60, 16, 25.0, 1000.1, 0.6, p.levels, 1) is None
    bp = O.BP(p, vol, wta, 16, min_level=p.levels - 1)
    assert bp.step() == 0            # nothing above the top level: map reads zero planes
    d, pl = bp.map()
    assert np.all(d == 0) and np.all(pl == 0)


# ------------------------------------------------------------------------- LR
def test_dilate_matches_scipy():
    rng = np.random.default_rng(7)
    f = (rng.random((23, 31)) < 0.1).astype(np.uint8) * 255
    out = np.zeros_like(f)
    L.orc_dilate3(f, 31, 23, out)
    ref = ndimage.grey_dilation(f, size=(3, 3), mode="constant", cval=0)
    assert np.array_equal(out, ref)


def test_consistency_check_properties():
    H, W = 10, 40
    dl = np.full((H, W), 5.0, np.float32)
    dr = np.full((H, W), 5.0, np.float32)
    fl, fr = np.zeros((H, W), np.uint8), np.zeros((H, W), np.uint8)
    L.orc_lr_consistency(dl, dr, W, H, fl, fr)
    assert not fl.any() and not fr.any()               # consistent maps never fail
    dr2 = dr.copy()
    dr2[:, 10] = 9.0                                   # left pixel x=15 maps to right x=10
    L.orc_lr_consistency(dl, dr2, W, H, fl, fr)
    assert np.all(fl[:, 15] == 255) and fl.sum() == 255 * H
    assert np.all(fr[:, 10] == 255)
    # rx = int(x - d + 0.5) truncates toward zero: x=0, d=0.6 -> int(-0.1) = 0, inside
    dl3 = np.zeros((1, W), np.float32)
    dl3[0, 0] = 0.6
    dr3 = np.full((1, W), 5.0, np.float32)
    f1, f2 = np.zeros((1, W), np.uint8), np.zeros((1, W), np.uint8)
    L.orc_lr_consistency(dl3, dr3, W, 1, f1, f2)
    assert f1[0, 0] == 255


def test_fill_takes_nearest_valid_plane_with_smaller_depth():
    H, W = 1, 9
    fail = np.zeros((H, W), np.uint8)
    fail[0, 4] = 255
    f2 = np.zeros_like(fail)
    L.orc_dilate3(fail, W, H, f2)                      # columns 3..5
    plane = np.zeros((H, W, 3), np.float32)
    plane[0, 2] = [0, 0, 7]
    plane[0, 6] = [0, 0, 3]
    plane[0, 4] = [9, 9, 9]
    L.orc_lr_fill(fail, f2, plane, W, H)
    assert list(plane[0, 4]) == [0, 0, 3]              # right neighbour is nearer (smaller z)
    plane[0, 4] = [9, 9, 9]
    plane[0, 6] = [0, 0, 7]                            # tie -> right
    plane[0, 2] = [1, 0, -1]                           # z at x=4: 3
    L.orc_lr_fill(fail, f2, plane, W, H)
    assert list(plane[0, 4]) == [1, 0, -1]
    d = np.zeros((H, W), np.float32)
    L.orc_plane_to_disp(plane, W, H, d)
    assert d[0, 4] == 3.0 and d[0, 2] == 1.0
