"""Serving under memory pressure (the 3840x2160, max disparity 384 stress case): one 4K pair
served beside smaller pairs under a 16 GB HBM budget, its result bit-identical to the plain
path's; a budget below the pair's estimate rejects it with an error instead of admitting it.
The 4K images repeat each pixel of a 960x540 synthetic pair 4x4 (the generator takes a minute
for a native 4K pair); the paths compare on the same input, so the content only has to be a
valid image."""
import sys

import numpy as np
import pytest

from hbp_amd import lib as H
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


@pytest.fixture(scope="module")
def pair4k():
    left, right, _ = synth.make_pair(31, 960, 540, 96, (8, 16))
    up = lambda im: np.ascontiguousarray(im.repeat(4, axis=0).repeat(4, axis=1))
    return up(left), up(right)


def plain(ctx, left, right, D, min_level):
    vol = ctx.cost_volume(left, right, D)
    pyr = ctx.pyramid(left)
    bp = ctx.bp(pyr, vol, min_level=max(1, min(min_level, pyr.levels - 1)))
    return bp.run()


def test_4k_pair_served_beside_small_pairs_under_16gb(ctx, pair4k):
    left, right = pair4k
    say("4K D384 served with 16 GB budget")
    small = [synth.make_pair(40 + i, w, h, d, (6, 12)) for i, (w, h, d) in
             enumerate([(640, 480, 64), (1280, 720, 128), (640, 480, 64)])]
    srv = H.Server(ctx, max_active=8, hbm_budget=16 << 30)
    tickets = [srv.submit(left, right, 384, 2)]
    for (l, r, _), d in zip(small, (64, 128, 64)):
        tickets.append(srv.submit(l, r, d, 2))
    results = {}
    while len(results) < len(tickets):
        _, _, ndone = srv.step()
        if ndone:
            for t in tickets:
                if t not in results:
                    r = srv.poll(t)
                    if r is not None:
                        results[t] = r
    srv.close()
    say("plain 4K")
    d4, p4 = plain(ctx, left, right, 384, 2)
    assert same_bits(results[tickets[0]][0], d4) and same_bits(results[tickets[0]][1], p4)
    for t, (l, r, _), d in zip(tickets[1:], small, (64, 128, 64)):
        ds, ps = plain(ctx, l, r, d, 2)
        assert same_bits(results[t][0], ds) and same_bits(results[t][1], ps)


def test_pair_above_the_budget_is_rejected(ctx, pair4k):
    left, right = pair4k
    srv = H.Server(ctx, max_active=4, hbm_budget=2 << 30)
    t = srv.submit(left, right, 384, 2)
    small = synth.make_pair(50, 320, 240, 32, (3, 6))
    t2 = srv.submit(small[0], small[1], 32, 2)
    got = {}
    for _ in range(100000):
        srv.step()
        for k in (t, t2):
            if k not in got:
                try:
                    r = srv.poll(k)
                except H.HbpError as e:
                    got[k] = e
                    continue
                if r is not None:
                    got[k] = r
        if len(got) == 2:
            break
    srv.close()
    assert isinstance(got[t], H.HbpError) and "budget" in str(got[t])
    assert not isinstance(got[t2], H.HbpError)
