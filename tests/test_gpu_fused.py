"""Fused sweeps (tuning "fuse"): on segment-form BP levels a run of RANSAC sweeps takes one launch
per sweep, the previous sweep's node update recomputed where it is read (fuse 1), and with fuse 2
(the default) every RANSAC sweep's plane costs are computed ahead on a second stream.  The
results must be the bits of the two-launch sweeps (fuse 0) for every cost form (census of either
view, the WTA indicator, a float volume), both segment step batches, any split of the schedule
into hbp_bp_step calls, any batch of sweeps per ahead launch, and the oracle's."""
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


@pytest.fixture
def tuned(ctx):
    before = ctx.get_tuning()
    yield ctx
    for k, v in before.items():
        ctx.set_tuning(k, v)
    for l in range(24):
        ctx.set_tuning(f"level_form.{l}", -1)


def make_bp(ctx, left, right, D, ml, right_view=False, cost="Census", **kw):
    pyr = ctx.pyramid(right if right_view else left)
    vol = ctx.cost_volume(left, right, D, right_view, cost=cost)
    return ctx.bp(pyr, vol, min_level=ml, **kw)


CASES = [
    # seed, W, H, D, min_level, right_view, cost, bp keywords
    (71, 320, 240, 48, 1, False, "Census", {}),
    (72, 256, 200, 32, 2, True, "Census", {}),
    (73, 240, 176, 40, 2, False, "Census", {"use_cost_vol": 0, "alpha": 1.3}),
    (74, 240, 160, 32, 1, False, "AD", {}),
    (75, 200, 152, 24, 2, False, "Census", {"alpha": -3.7, "tau_s": 2.5}),
    (77, 217, 151, 27, 1, True, "Census", {}),     # odd sizes and disparity count
]


@pytest.mark.parametrize("seed,W,H,D,ml,right_view,cost,kw", CASES)
@pytest.mark.parametrize("hb", [4, 2])
def test_fused_runs_give_the_split_sweeps_bits(tuned, seed, W, H, D, ml, right_view, cost, kw, hb):
    say(f"fused vs split seed={seed} {W}x{H} D={D} ml={ml} right={right_view} {cost} {kw} hb={hb}")
    left, right, _ = synth.make_pair(seed, W, H, D, (3, 8))
    # every level in the segment form, so that runs are fused on every BP level
    tuned.set_tuning("coarse_min_pix", 1)
    tuned.set_tuning("single_hb", hb)
    tuned.set_tuning("pipe_hb", 6 - hb)
    out = {}
    for fuse in (0, 1, 2):
        tuned.set_tuning("fuse", fuse)
        out[fuse] = make_bp(tuned, left, right, D, ml, right_view, cost, **kw).run()
    for fuse in (1, 2):
        assert same_bits(out[0][0], out[fuse][0]) and same_bits(out[0][1], out[fuse][1]), fuse


@pytest.mark.parametrize("seed,W,H,D,ml,right_view,cost,kw", CASES)
@pytest.mark.parametrize("max_e", [64, 4096])
def test_one_workgroup_runs_give_the_split_sweeps_bits(tuned, seed, W, H, D, ml, right_view, cost, kw, max_e):
    """Levels of at most 32 superpixels (and run_max_e edges) run a fused run's sweeps after its
    first in one workgroup; with run_max_e 4096 every such level does."""
    say(f"one-workgroup runs seed={seed} {W}x{H} D={D} ml={ml} right={right_view} {cost} {kw} max_e={max_e}")
    left, right, _ = synth.make_pair(seed, W, H, D, (3, 8))
    tuned.set_tuning("coarse_min_pix", 1)
    tuned.set_tuning("fuse", 0)
    d0, p0 = make_bp(tuned, left, right, D, ml, right_view, cost, **kw).run()
    tuned.set_tuning("fuse", 2)
    tuned.set_tuning("run_max_e", max_e)
    d1, p1 = make_bp(tuned, left, right, D, ml, right_view, cost, **kw).run()
    assert same_bits(d0, d1) and same_bits(p0, p1)


@pytest.mark.parametrize("batch,blocks,window", [(1, 8, 0), (3, 64, 1), (7, 512, 0), (16, 4096, 2), (4, 512, 1),
                                                 (1, 512, 3)])
def test_ahead_launch_shapes_give_the_same_bits(tuned, batch, blocks, window):
    say(f"ahead batch={batch} blocks={blocks} window={window}")
    left, right, _ = synth.make_pair(79, 320, 240, 48, (3, 8))
    tuned.set_tuning("coarse_min_pix", 1)
    tuned.set_tuning("fuse", 0)
    d0, p0 = make_bp(tuned, left, right, 48, 1).run()
    tuned.set_tuning("fuse", 2)
    tuned.set_tuning("pipe_batch", batch)
    tuned.set_tuning("pipe_blocks", blocks)
    tuned.set_tuning("pipe_window", window)
    d1, p1 = make_bp(tuned, left, right, 48, 1).run()
    assert same_bits(d0, d1) and same_bits(p0, p1)


@pytest.mark.parametrize("fuse", [1, 2, 3, 4])
@pytest.mark.parametrize("chunk", [2, 3, 5, 7, 13])
def test_fused_state_after_every_step_call(tuned, chunk, fuse):
    """hbp_bp_step(n) ends a fused run where the call ends: the state read back after each call
    is the split sweeps' state, planes, costs, beliefs and messages alike."""
    say(f"fused {fuse} state, steps of {chunk}")
    left, right, _ = synth.make_pair(76, 288, 216, 40, (3, 8))
    tuned.set_tuning("coarse_min_pix", 1)
    if fuse == 3:        # fuse 2 with one-workgroup runs on every level of at most 32 superpixels
        tuned.set_tuning("run_max_e", 4096)
        fuse = 2
    elif fuse == 4:      # fuse 2 with the plane-cost launches issued one ahead of the waits
        tuned.set_tuning("pipe_window", 1)
        fuse = 2
    modes = (0, fuse)   # the form is the context's at each step call
    bps = [make_bp(tuned, left, right, 40, 2) for _ in modes]

    def step(i, n):
        tuned.set_tuning("fuse", modes[i])
        return bps[i].step(n)

    total = bps[0].total_sweeps
    done = 0
    calls = 0
    while done < total and calls < 40:
        n = [step(i, chunk) for i in (0, 1)]
        assert n[0] == n[1]
        done += n[0]
        calls += 1
        s0, s1 = bps[0].state(), bps[1].state()
        assert s0["level"] == s1["level"]
        for key in ("planes", "cost", "belief", "msg"):
            assert same_bits(s0[key], s1[key]), f"{key} differs after {done} sweeps"
    step(0, 1 << 30)
    step(1, 1 << 30)
    d0, p0 = bps[0].run()
    d1, p1 = bps[1].run()
    assert same_bits(d0, d1) and same_bits(p0, p1)


# the 90x70 pair's segment-form levels have about 20 segments: fewer ahead blocks than XCDs
@pytest.mark.parametrize("seed,W,H,D,ml", [(77, 200, 150, 32, 1), (78, 320, 240, 64, 2), (74, 90, 70, 16, 2),
                                           (75, 96, 72, 24, 2)])
def test_fused_runs_match_the_oracle(tuned, seed, W, H, D, ml):
    say(f"fused vs oracle seed={seed} {W}x{H} D={D} ml={ml}")
    left, right, _ = synth.make_pair(seed, W, H, D, (3, 8))
    vo, wo = O.cost_volume(left, right, D)
    bo = O.BP(O.Pyramid(left), vo, wo, D, min_level=ml)
    bo.step()
    do, plo = bo.map()
    for cmp in (1, 96):
        for fuse in (1, 2):
            tuned.set_tuning("coarse_min_pix", cmp)
            tuned.set_tuning("fuse", fuse)
            d, p = make_bp(tuned, left, right, D, ml).run()
            assert same_bits(d, do) and same_bits(p, plo), (cmp, fuse)
