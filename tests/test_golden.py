"""The committed golden fixtures (tests/golden): digests, the oracle reproducing them bit for
bit, and checks of the fixtures against what the definitions fix, written independently of the
oracle in numpy (census words, cost volume entries, WTA ties, the right view, the coarsened edge
lists, MAP disparity from planes, left-right consistency masks)."""
import hashlib
import json
import os
import sys

import numpy as np
import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
GOLD = os.path.join(HERE, "golden")
sys.path.insert(0, GOLD)


def load(name):
    return np.load(os.path.join(GOLD, name + ".npy"), allow_pickle=False)


META = json.load(open(os.path.join(GOLD, "meta.json")))
W, H = META["W"], META["H"]


def same_bits(a, b):
    a, b = np.ascontiguousarray(a), np.ascontiguousarray(b)
    if a.dtype != b.dtype or a.shape != b.shape:
        return False
    return np.array_equal(a.view(np.uint8), b.view(np.uint8))


def test_fixture_digests():
    for line in open(os.path.join(GOLD, "SHA256SUMS")):
        digest, name = line.split()
        assert hashlib.sha256(open(os.path.join(GOLD, name), "rb").read()).hexdigest() == digest, name


def test_oracle_reproduces_every_fixture():
    import make_golden
    out, meta = make_golden.compute()
    assert meta == META
    for name, arr in out.items():
        assert same_bits(arr, load(name)), name


# ---- checks written from the definitions, not from the oracle ----

def grey(bgr):
    # B*0.114 + G*0.587 + R*0.299 in FP64, truncated to u8 (saturating at 255)
    g = bgr[..., 0].astype(np.float64) * 0.114 + bgr[..., 1].astype(np.float64) * 0.587 + \
        bgr[..., 2].astype(np.float64) * 0.299
    g = g.astype(np.float32)
    return np.where(g < 255.0, np.trunc(g), 255.0).astype(np.uint8)


def census_words(g):
    # 9x7 centre-symmetric census, zero outside the image: bit k*9+m (k<3, m<9) compares
    # I(y-3+k, x-4+m) >= I(y+3-k, x+4-m); bits 27+m (m<4) compare I(y, x-4+m) >= I(y, x+4-m)
    p = np.zeros((H + 6, W + 8), np.int32)
    p[3:3 + H, 4:4 + W] = g
    def at(dy, dx):
        return p[3 + dy:3 + dy + H, 4 + dx:4 + dx + W]
    out = np.zeros((H, W), np.uint32)
    for k in range(3):
        for m in range(9):
            out |= (at(-3 + k, -4 + m) >= at(3 - k, 4 - m)).astype(np.uint32) << np.uint32(k * 9 + m)
    for m in range(4):
        out |= (at(0, -4 + m) >= at(0, 4 - m)).astype(np.uint32) << np.uint32(27 + m)
    return out


def popcount32(x):
    x = x.astype(np.uint64)
    c = np.zeros(x.shape, np.int64)
    for b in range(32):
        c += ((x >> np.uint64(b)) & np.uint64(1)).astype(np.int64)
    return c


def test_census_from_definition():
    assert np.array_equal(census_words(grey(load("left"))), load("census_l"))
    assert np.array_equal(census_words(grey(load("right"))), load("census_r"))


@pytest.mark.parametrize("D", [16, 64])
def test_volume_and_wta_from_census(D):
    cl, cr = load("census_l"), load("census_r")
    xs = np.arange(W)
    vol = np.zeros((D, H, W), np.uint8)
    for d in range(D):
        vol[d] = np.minimum(popcount32(cl ^ cr[:, np.maximum(xs - d, 0)]), 25)
    assert np.array_equal(vol, load(f"vol_d{D}"))
    # WTA: the first (smallest) d of the minimum
    assert np.array_equal(np.argmin(vol, axis=0).astype(np.float32), load(f"wta_d{D}"))


def test_right_view_from_left_volume():
    vl, vr = load("vol_d16"), load("vol_right_d16")
    xs = np.arange(W)
    for d in range(16):
        assert np.array_equal(vr[d], vl[d][:, np.minimum(xs + d, W - 1)]), d
    assert np.array_equal(np.argmin(vr, axis=0).astype(np.float32), load("wta_right_d16"))


def literal_reduce(clus, bd, bdl, Vn):
    """reduce_boundary_dist (segment.cu:1050-1091) as the reference executes it: the mapped edges
    are sorted (thrust::sort moves only the edges), each run of equal (s, t) with s != t becomes
    one edge, and reduce_bdl adds bdl[i] for every *position* i of the run -- the bdl array is
    never permuted, so the sum is over the previous level's bdl at the run's positions."""
    m = np.stack([clus[bd[:, 0]], clus[bd[:, 1]]], axis=1).astype(np.int64)
    keys = np.sort(m[:, 0] * Vn + m[:, 1])
    uniq, start = np.unique(keys, return_index=True)
    end = np.r_[start[1:], len(keys)]
    csum = np.r_[0, np.cumsum(bdl.astype(np.int64))]
    keep = (uniq // Vn) != (uniq % Vn)
    edges = np.stack([uniq // Vn, uniq % Vn], axis=1)[keep].astype(np.int32)
    return edges, (csum[end] - csum[start])[keep].astype(np.int32)


@pytest.mark.parametrize("l", [0, 1])
def test_coarse_edges_from_fine_level(l):
    # the edges of level l+1: the sorted set of mapped (clus[s], clus[t]) with s != t; bdl by the
    # literal position rule, which for l = 0 (all bdl 1) is the multiplicity of each edge
    clus, bd, bdl = load(f"pyr_l{l}_clus"), load(f"pyr_l{l}_bd"), load(f"pyr_l{l}_bdl")
    edges, lens = literal_reduce(clus, bd, bdl, META["V"][l + 1])
    assert np.array_equal(load(f"pyr_l{l + 1}_bd"), edges)
    assert np.array_equal(load(f"pyr_l{l + 1}_bdl"), lens)
    assert len(clus) == META["V"][l] and clus.max() + 1 == META["V"][l + 1]
    if l == 0:
        m = clus[bd]
        m = m[m[:, 0] != m[:, 1]]
        _, mult = np.unique(m[:, 0].astype(np.int64) * META["V"][1] + m[:, 1], return_counts=True)
        assert np.array_equal(lens, mult)


def test_level2_bdl_differs_from_boundary_length():
    # the position rule is not the boundary length once the bdl differ (level 1 -> 2): the
    # fixture holds the reference's values, not the physically summed ones
    clus, bd, bdl = load("pyr_l1_clus"), load("pyr_l1_bd"), load("pyr_l1_bdl")
    m = np.stack([clus[bd[:, 0]], clus[bd[:, 1]]], axis=1)
    agg = {}
    for (s, t), w in zip(m.tolist(), bdl.tolist()):
        if s != t:
            agg[(s, t)] = agg.get((s, t), 0) + w
    physical = np.array([agg[k] for k in sorted(agg)], np.int32)
    assert not np.array_equal(load("pyr_l2_bdl"), physical)


def test_level0_edges_are_the_4_neighbourhood():
    bd = load("pyr_l0_bd").reshape(H * W, 4, 2)
    ids = np.arange(H * W).reshape(H, W)
    up = np.vstack([ids[:1], ids[:-1]])
    down = np.vstack([ids[1:], ids[-1:]])
    lft = np.hstack([ids[:, :1], ids[:, :-1]])
    rgt = np.hstack([ids[:, 1:], ids[:, -1:]])
    assert np.array_equal(bd[:, :, 0], np.repeat(ids.reshape(-1, 1), 4, axis=1))
    assert np.array_equal(bd[:, :, 1], np.stack([up, down, lft, rgt], axis=-1).reshape(-1, 4))
    assert np.all(load("pyr_l0_bdl") == 1)


def test_map_disparity_from_planes():
    D = META["D_bp"]
    for disp, plane in (("disp", "plane3"), ("disp_right", "plane3_right")):
        p = load(plane)
        ys, xs = np.mgrid[0:H, 0:W].astype(np.float32)
        d = (p[..., 0] * xs + p[..., 1] * ys) + p[..., 2]
        d = np.where(np.isnan(d), np.float32(0), np.clip(d, np.float32(0), np.float32(D - 1)))
        assert same_bits(d.astype(np.float32), load(disp)), disp


def test_bp_state_shapes_and_prior_slots():
    V4 = META["V"][META["bp_first_level"]]
    for tag in ("first", "level"):
        pl, bel = load(f"bp_{tag}_planes"), load(f"bp_{tag}_belief")
        assert pl.shape == (12, V4, 4) and bel.shape == (12, V4)
    # the prior slots P..2P-1 (the parents' best planes) are never rewritten within a level
    assert same_bits(load("bp_first_planes")[4:8], load("bp_level_planes")[4:8])


def test_left_right_masks_from_definition():
    dl, dr = load("disp"), load("disp_right")
    fail = np.zeros((H, W), np.uint8)
    for y in range(H):
        for x in range(W):
            d = dl[y, x]
            rx = int(np.float32(np.float32(x) - d + np.float32(0.5)))   # truncation toward 0
            if 0 <= rx < W and abs(dr[y, rx] - d) > 1.5:
                fail[y, x] = 255
    assert np.array_equal(fail, load("lr_fail_l"))
    pad = np.zeros((H + 2, W + 2), np.uint8)
    pad[1:-1, 1:-1] = fail
    dil = np.zeros_like(fail)
    for dy in range(3):
        for dx in range(3):
            dil |= pad[dy:dy + H, dx:dx + W]
    assert np.array_equal(dil, load("lr_fail2_l"))
    # pixels that did not fail keep their planes
    keep = fail == 0
    assert same_bits(load("lr_plane3")[keep], load("plane3")[keep])
