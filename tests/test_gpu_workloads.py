"""GPU parity at the benchmark workloads' own sizes:
  * the 1472x984, D=256 pair through all 1754 sweeps of its BP (levels 6..9), against the oracle;
  * the batch of 64 pairs of mixed sizes (W U[320,1920], H U[240,1080], D from {64,128,192,256})
    through hbp_bp_run_batch, with its rounds queued and replayed from one HIP graph, every pair
    bit-identical to its own plain run, the two smallest also against the oracle;
  * the request server under Poisson arrivals of the serving mix (40% 640x480/D64, 40%
    1280x720/D128, 20% 1920x1080/D256), 120 requests, every result checked against the plain
    path;
  * the 3840x2160, D=384 scene with BP on 5 levels served under a 16 GB budget beside pairs of
    the mixed batch, every result bit-identical to the plain path."""
import numpy as np
import pytest

import oracle_ffi as O
import pairgen
from hbp_amd import lib as H
from hbp_amd import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def same_bits(a, b):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _plain(ctx, left, right, D, min_level):
    """One pair through the plain path; min_level < 0 means BP on -min_level levels."""
    pyr = ctx.pyramid(left)
    L = pyr.levels
    ml = L - 1 + min_level if min_level < 0 else min_level
    ml = max(1, min(ml, L - 1))
    bp = ctx.bp(pyr, ctx.cost_volume(left, right, D), min_level=ml)
    out = bp.run()
    return out, ml, bp.total_sweeps


def test_large_pair_full_bp_against_the_oracle(ctx):
    """Every sweep of the 1472x984 pair's BP (min_level = levels - 5), then the map, against the
    oracle run to the end."""
    say("1472x984 D256: full BP vs oracle")
    left, right = pairgen.pair(11, 1472, 984, 256, (20, 40))
    (dg, pg), ml, sweeps = _plain(ctx, left, right, 256, -4)
    vo, wo = O.cost_volume(left, right, 256)
    po = O.Pyramid(left)
    assert ml == po.levels - 5
    bo = O.BP(po, vo, wo, 256, min_level=ml)
    assert bo.step() == sweeps
    do, plo = bo.map()
    assert same_bits(do, dg) and same_bits(plo, pg)


def test_batch_of_64_mixed_pairs(ctx):
    """hbp_bp_run_batch over the 64-pair mix: every pair's disparity and planes equal its plain
    run bit for bit; the two smallest pairs are also run by the oracle."""
    specs = synth.mixed_batch(1, 64)
    say(f"batch of 64 mixed pairs: generating {sum(w * h for _, w, h, _ in specs) / 1e6:.1f} Mpixel")
    pairs = pairgen.pairs(specs)
    ref = []
    for (sd, w, h, d), (left, right) in zip(specs, pairs):
        (dg, pg), ml, _ = _plain(ctx, left, right, d, 2)
        ref.append((dg, pg, ml))
    for graph in (0, 1):
        say(f"batch of 64 mixed pairs: batched run, rounds {'in one HIP graph' if graph else 'queued'}")
        ctx.set_tuning("batch_graph", graph)
        try:
            bps = []
            for (sd, w, h, d), (left, right), (_, _, ml) in zip(specs, pairs, ref):
                bps.append(ctx.bp(ctx.pyramid(left), ctx.cost_volume(left, right, d), min_level=ml))
            ctx.bp_run_batch(bps)
        finally:
            ctx.set_tuning("batch_graph", 0)
        for i, bp in enumerate(bps):
            dg, pg = bp.run()
            assert same_bits(dg, ref[i][0]) and same_bits(pg, ref[i][1]), f"pair {i} {specs[i]} graph {graph}"
            bp.free()
    small = sorted(range(64), key=lambda i: specs[i][1] * specs[i][2])[:2]
    for i in small:
        sd, w, h, d = specs[i]
        say(f"batch pair {i} {w}x{h} D{d} vs oracle")
        left, right = pairs[i]
        vo, wo = O.cost_volume(left, right, d)
        bo = O.BP(O.Pyramid(left), vo, wo, d, min_level=ref[i][2])
        bo.step()
        do, plo = bo.map()
        assert same_bits(do, ref[i][0]) and same_bits(plo, ref[i][1])


SERVE_MIX = ((0.4, 640, 480, 64), (0.4, 1280, 720, 128), (0.2, 1920, 1080, 256))


def test_server_under_poisson_arrivals_of_the_serving_mix(ctx):
    """120 requests, Poisson arrivals at 40 pairs/s (submitted when the serving loop's clock
    passes their arrival time), 4 distinct pairs per class: every result equals that pair's
    plain run."""
    import time
    per = 4
    specs = [(900 + 10 * c + i, w, h, d) for c, (_, w, h, d) in enumerate(SERVE_MIX) for i in range(per)]
    say("serving mix: generating 12 pairs")
    pairs = pairgen.pairs(specs)
    ref = [_plain(ctx, l, r, s[3], 2)[0] for s, (l, r) in zip(specs, pairs)]
    rng = np.random.default_rng(3)
    n = 120
    arrivals = np.cumsum(rng.exponential(1 / 40.0, n))
    cls = rng.choice(3, n, p=[m[0] for m in SERVE_MIX])
    which = [c * per + int(rng.integers(per)) for c in cls]
    say(f"serving mix: {n} requests")
    srv = H.Server(ctx, max_active=64, hbm_budget=96 << 30)
    tickets, out, nxt = {}, {}, 0
    t0 = time.perf_counter()
    while nxt < n or tickets:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            j = which[nxt]
            left, right = pairs[j]
            tickets[srv.submit(left, right, specs[j][3], 2)] = nxt
            nxt += 1
        srv.step()
        for tk in list(tickets):
            r = srv.poll(tk)
            if r is not None:
                out[tickets.pop(tk)] = r
        assert time.perf_counter() - t0 < 240
    srv.close()
    assert len(out) == n
    for i in range(n):
        d, p = ref[which[i]]
        assert same_bits(out[i][0], d) and same_bits(out[i][1], p), f"request {i}"


def test_4k_pair_served_under_16gb_beside_the_mixed_batch(ctx):
    """3840x2160, D=384, BP on 5 levels (a request's min_level -5), served under a 16 GB budget
    with 8 pairs of the mixed batch: each result bit-identical to its plain run."""
    say("4K D384: generating")
    l4, r4 = pairgen.pair(5, 3840, 2160, 384, (20, 40))
    specs = synth.mixed_batch(2, 8)
    pairs = pairgen.pairs(specs)
    say("4K D384: plain runs")
    jobs = [(l4, r4, 384, -5)] + [(l, r, s[3], 2) for s, (l, r) in zip(specs, pairs)]
    ref = [_plain(ctx, *j)[0] for j in jobs]
    say("4K D384: served under 16 GB")
    srv = H.Server(ctx, max_active=16, hbm_budget=16 << 30)
    tickets = {srv.submit(l, r, d, ml): i for i, (l, r, d, ml) in enumerate(jobs)}
    out = {}
    for _ in range(10_000_000):
        srv.step()
        for tk in list(tickets):
            r = srv.poll(tk)
            if r is not None:
                out[tickets.pop(tk)] = r
        if not tickets:
            break
    srv.close()
    for i in range(len(jobs)):
        assert same_bits(out[i][0], ref[i][0]) and same_bits(out[i][1], ref[i][1]), f"job {i}"
