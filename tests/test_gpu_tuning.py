"""The kernel-form tuning of a context (hbp_ctx_set_tuning, hbp_amd.tune) changes speed only:
one-pair and batched BP give the same bits for every value, including the extremes that put
every level in one form, and the oracle agrees."""
import sys

import numpy as np
import pytest

import oracle_ffi as O
from hbp_amd import lib as H
from hbp_amd import synth, tune

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


SETTINGS = [dict(coarse_min_pix=96, single_hb=4, batch_hb=2, seg_len=512, batch_graph=0, fuse=0),
            dict(coarse_min_pix=1, single_hb=2, batch_hb=4, seg_len=64, batch_graph=1, fuse=2),
            dict(coarse_min_pix=1 << 24, single_hb=4, batch_hb=4, seg_len=256, batch_graph=0, fuse=1),
            dict(coarse_min_pix=40, single_hb=2, batch_hb=2, seg_len=320, batch_graph=1, fuse=0),
            dict(coarse_min_pix=96, single_hb=4, batch_hb=2, seg_len=448, batch_graph=1, fuse=1)]


def run_pair(ctx, left, right, D, ml):
    return ctx.bp(ctx.pyramid(left), ctx.cost_volume(left, right, D), min_level=ml).run()


def test_one_pair_results_do_not_depend_on_tuning(tuned):
    left, right, _ = synth.make_pair(61, 640, 480, 64, (6, 12))
    outs = []
    for st in SETTINGS:
        say("tuning", st)
        for k, v in st.items():
            tuned.set_tuning(k, v)
        outs.append(run_pair(tuned, left, right, 64, 1))
    for d, p in outs[1:]:
        assert same_bits(d, outs[0][0]) and same_bits(p, outs[0][1])


def test_extreme_forms_match_the_oracle(tuned):
    left, right, _ = synth.make_pair(62, 160, 120, 24, (3, 8))
    vo, wo = O.cost_volume(left, right, 24)
    bo = O.BP(O.Pyramid(left), vo, wo, 24, min_level=1)
    bo.step()
    do, plo = bo.map()
    for cmp in (1, 1 << 24):
        for hb in (2, 4):
            tuned.set_tuning("coarse_min_pix", cmp)
            tuned.set_tuning("single_hb", hb)
            tuned.set_tuning("seg_len", 64 if hb == 2 else 512)
            d, p = run_pair(tuned, left, right, 24, 1)
            assert same_bits(d, do) and same_bits(p, plo), (cmp, hb)


def test_batched_results_do_not_depend_on_tuning(tuned):
    specs = [(63, 320, 240, 32), (64, 480, 272, 48), (65, 640, 480, 64)]
    pairs = [synth.make_pair(s, w, h, d, (3, 8)) for s, w, h, d in specs]
    ref = [run_pair(tuned, l, r, d, 2) for (l, r, _), (_, _, _, d) in zip(pairs, specs)]
    for st in SETTINGS:
        say("batched tuning", st)
        for k, v in st.items():
            tuned.set_tuning(k, v)
        keep = []
        bps = []
        for (l, r, _), (_, _, _, d) in zip(pairs, specs):
            v, p = tuned.cost_volume(l, r, d), tuned.pyramid(l)
            keep += [v, p]
            bps.append(tuned.bp(p, v, min_level=2))
        tuned.bp_run_batch(bps)
        for bp, (dr, pr) in zip(bps, ref):
            d, p = bp.run()
            assert same_bits(d, dr) and same_bits(p, pr)


def test_tuning_parameters_are_checked(tuned):
    with pytest.raises(H.HbpError, match="unknown tuning"):
        tuned.set_tuning("tile", 3)
    with pytest.raises(H.HbpError, match="2 or 4"):
        tuned.set_tuning("batch_hb", 3)
    with pytest.raises(H.HbpError, match="coarse_min_pix"):
        tuned.set_tuning("coarse_min_pix", 0)
    with pytest.raises(H.HbpError, match="4 or 8"):
        tuned.set_tuning("hamming_cols", 6)
    with pytest.raises(H.HbpError, match="seg_len"):
        tuned.set_tuning("seg_len", 100)
    with pytest.raises(H.HbpError, match="seg_len"):
        tuned.set_tuning("seg_len", 576)
    with pytest.raises(H.HbpError, match="batch_graph"):
        tuned.set_tuning("batch_graph", 2)
    with pytest.raises(H.HbpError, match="fuse"):
        tuned.set_tuning("fuse", 3)
    with pytest.raises(H.HbpError, match="pipe_batch"):
        tuned.set_tuning("pipe_batch", 17)
    with pytest.raises(H.HbpError, match="pipe_blocks"):
        tuned.set_tuning("pipe_blocks", 4)
    with pytest.raises(H.HbpError, match="pipe_window"):
        tuned.set_tuning("pipe_window", -1)
    with pytest.raises(H.HbpError, match="crt_rounds"):
        tuned.set_tuning("crt_rounds", 0)
    with pytest.raises(H.HbpError, match="crt_jumps"):
        tuned.set_tuning("crt_jumps", 9)
    with pytest.raises(H.HbpError, match="level form"):
        tuned.set_tuning("level_form.3", 2)
    with pytest.raises(H.HbpError, match="out of range"):
        tuned.set_tuning("level_form.24", 0)
    assert set(tuned.get_tuning()) == {"coarse_min_pix", "single_hb", "batch_hb", "hamming_cols", "seg_len",
                                      "batch_graph", "fuse", "pipe_batch", "pipe_blocks",
                                      "pipe_hb", "pipe_window", "run_max_e", "crt_rounds", "crt_jumps",
                                      "crt_max_jumps"}


def test_tuner_sets_the_fastest_grid_point(tuned):
    left, right, _ = synth.make_pair(66, 320, 240, 32, (3, 8))
    r = tune.tune_single(tuned, left, right, 32, 2, grid={"coarse_min_pix": (32, 256), "single_hb": (4, 2)}, reps=1)
    assert len(r["ms"]) == 4 and r["best"]["coarse_min_pix"] in (32, 256) and r["best"]["single_hb"] in (2, 4)
    assert tuned.get_tuning("coarse_min_pix") == r["best"]["coarse_min_pix"]
    pairs = [synth.make_pair(67 + i, 320, 240, 32, (3, 8))[:2] + (32,) for i in range(2)]
    rb = tune.tune_batch(tuned, pairs, 2, grid={"batch_hb": (2, 4)}, reps=1)
    assert tuned.get_tuning("batch_hb") == rb["best"]["batch_hb"] and len(rb["ms"]) == 2


def test_per_level_form_schedules_give_the_same_bits(tuned):
    """Any schedule of plane-cost forms over the BP levels (one pair and batched) gives the bits
    of the default rule."""
    rng = np.random.default_rng(7)
    left, right, _ = synth.make_pair(68, 640, 480, 64, (6, 12))
    d0, p0 = run_pair(tuned, left, right, 64, 1)
    specs = [(69, 320, 240, 32), (70, 480, 272, 48)]
    pairs = [synth.make_pair(s, w, h, d, (3, 8)) for s, w, h, d in specs]
    ref = [run_pair(tuned, l, r, d, 1) for (l, r, _), (_, _, _, d) in zip(pairs, specs)]
    for trial in range(3):
        forms = rng.integers(0, 2, 24)
        say("level forms", forms[:10])
        for l in range(24):
            tuned.set_tuning(f"level_form.{l}", int(forms[l]))
        d, p = run_pair(tuned, left, right, 64, 1)
        assert same_bits(d, d0) and same_bits(p, p0)
        keep, bps = [], []
        for (l, r, _), (_, _, _, dd) in zip(pairs, specs):
            v, py = tuned.cost_volume(l, r, dd), tuned.pyramid(l)
            keep += [v, py]
            bps.append(tuned.bp(py, v, min_level=1))
        tuned.bp_run_batch(bps)
        for bp, (dr, pr) in zip(bps, ref):
            d, p = bp.run()
            assert same_bits(d, dr) and same_bits(p, pr)


@pytest.mark.parametrize("right_view", [False, True])
def test_hamming_widths_give_the_oracle_volume(tuned, right_view):
    for W, Hh, D in ((333, 201, 37), (1472, 64, 256)):
        left, right, _ = synth.make_pair(71, W, Hh, D, (3, 8))
        vo, wo = O.cost_volume(left, right, D, right_view=right_view)
        for cols in (4, 8):
            tuned.set_tuning("hamming_cols", cols)
            v = tuned.cost_volume(left, right, D, right_view=right_view)
            vg, wg = v.download()
            assert np.array_equal(vg, vo) and np.array_equal(wg, wo), (W, D, cols)
            v.free()


def test_level_and_hamming_tuners_leave_their_choice(tuned):
    left, right, _ = synth.make_pair(72, 640, 480, 64, (6, 12))
    r = tune.tune_levels(tuned, left, right, 64, 2, reps=1)
    for name, f in r["schedule"].items():
        assert tuned.get_tuning(f"level_form.{name[1:]}") == (1 if f == "segments" else 0)
    rh = tune.tune_hamming(tuned, left, right, 64, reps=1)
    assert tuned.get_tuning("hamming_cols") == rh["best"]["hamming_cols"]
