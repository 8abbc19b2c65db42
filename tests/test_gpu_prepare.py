"""hbp_pairs_prepare_dev: many pairs' cost volumes, pyramids and BP objects built by several
host threads on streams of their own.  The objects equal the one-pair calls' bit for bit (BP
results through hbp_bp_run and hbp_bp_run_batch, with 1 and 4 workers, min_level given and
counted from the top); a pair whose CRTrees needs the cooperative form is redone on the
caller's stream; a bad pair or an injected allocation failure returns an error and leaves no
object behind."""
import ctypes as C

import numpy as np
import pytest

from hbp_amd import lib as H
from hbp_amd import synth

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        _hip = C.CDLL("libamdhip64.so")
    return _hip


def say(msg):
    print(f"[case] {msg}", flush=True)


class Dev:
    """A device copy of a host array (hipMalloc'ed by the test itself)."""

    def __init__(self, a):
        a = np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), C.c_size_t(max(a.nbytes, 1))) == 0
        assert hip().hipMemcpy(self.p, a.ctypes.data_as(C.c_void_p), C.c_size_t(a.nbytes), 1) == 0

    def free(self):
        hip().hipFree(self.p)


def same_bits(a, b):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


SPECS = [(301, 320, 240, 32, 2), (302, 480, 272, 64, 1), (303, 640, 480, 64, -3), (304, 200, 120, 16, 2),
         (305, 512, 384, 48, 3), (306, 256, 200, 24, -2)]


def _plain(ctx, left, right, D, ml):
    pyr = ctx.pyramid(left)
    L = pyr.levels
    m = L - 1 + ml if ml < 0 else ml
    m = max(1, min(m, L - 1))
    return ctx.bp(pyr, ctx.cost_volume(left, right, D), min_level=m).run()


@pytest.fixture(scope="module")
def pairs():
    return [synth.make_pair(s, w, h, d, (4, 10))[:2] for s, w, h, d, _ in SPECS]


@pytest.mark.parametrize("workers", [1, 4])
def test_prepared_pairs_equal_the_one_pair_calls(ctx, pairs, workers):
    ref = [_plain(ctx, l, r, spec[3], spec[4]) for (l, r), spec in zip(pairs, SPECS)]
    dev = [(Dev(l), Dev(r)) for l, r in pairs]
    try:
        for batched in (False, True):
            say(f"prepare {len(SPECS)} pairs, {workers} workers, batched={batched}")
            objs = ctx.prepare_pairs_dev([(dl.p.value, dr.p.value, s[1], s[2], s[3], s[4])
                                          for (dl, dr), s in zip(dev, SPECS)], workers=workers)
            if batched:
                ctx.bp_run_batch([b for _, _, b in objs])
            for i, (v, p, b) in enumerate(objs):
                d, pl = b.run()
                assert same_bits(d, ref[i][0]) and same_bits(pl, ref[i][1]), (i, SPECS[i], batched)
            for v, p, b in objs:
                b.free()
                p.free()
                v.free()
    finally:
        for dl, dr in dev:
            dl.free()
            dr.free()


def test_pair_needing_the_cooperative_crtrees_is_redone_on_the_callers_stream(ctx, pairs):
    """A flat 20x1000 image: every column is one 1-NN chain of 1000 nodes, longer than the
    scheduled pointer jumping covers, so a worker may not launch the cooperative kernel; the
    pair is rebuilt after the workers and equals the one-pair calls."""
    flat = np.full((1000, 20, 3), 90, np.uint8)
    say("prepare with a cooperative CRTrees pair")
    specs = [(flat, flat, 20, 1000, 8, 1)] + [(l, r, s[1], s[2], s[3], s[4]) for (l, r), s in zip(pairs[:3], SPECS)]
    ref = [_plain(ctx, l, r, D, ml) for l, r, _, _, D, ml in specs]
    dev = [(Dev(l), Dev(r)) for l, r, *_ in specs]
    try:
        ctx.profile(True)
        ctx.profile_reset()
        try:
            objs = ctx.prepare_pairs_dev([(dl.p.value, dr.p.value, W, Hh, D, ml)
                                          for (dl, dr), (_, _, W, Hh, D, ml) in zip(dev, specs)], workers=3)
            st = ctx.profile_stats()
        finally:
            ctx.profile(False)
        # the redo ran on the caller's context (the workers' contexts have no profiling on)
        assert st.get("crtrees_cooperative", {}).get("launches", 0) >= 1
        po = ctx.pyramid(flat)
        assert objs[0][1].levels == po.levels
        for l in range(po.levels):
            a, b = po.level(l), objs[0][1].level(l)
            assert np.array_equal(a["clus"], b["clus"]) and np.array_equal(a["bd"], b["bd"]), l
        for i, (v, p, b) in enumerate(objs):
            d, pl = b.run()
            assert same_bits(d, ref[i][0]) and same_bits(pl, ref[i][1]), i
        for v, p, b in objs:
            b.free()
            p.free()
            v.free()
    finally:
        for dl, dr in dev:
            dl.free()
            dr.free()


def test_errors_leave_no_object_behind(pairs):
    c = H.Ctx(0)
    dev = [(Dev(l), Dev(r)) for l, r in pairs[:3]]
    try:
        good = [(dl.p.value, dr.p.value, s[1], s[2], s[3], s[4]) for (dl, dr), s in zip(dev, SPECS)]
        say("prepare with a bad pair")
        with pytest.raises(H.HbpError, match="bad pair size"):
            c.prepare_pairs_dev(good + [(good[0][0], good[0][1], 64, 48, 0, 2)], workers=2)
        say("prepare under an injected allocation failure")
        H.inject_fault("alloc", 40)
        try:
            with pytest.raises(H.HbpError):
                c.prepare_pairs_dev(good, workers=3)
        finally:
            H.inject_fault("alloc", 0)
        say("prepare after the failures")
        objs = c.prepare_pairs_dev(good, workers=3)
        assert len(objs) == 3
        for v, p, b in objs:
            b.free()
            p.free()
            v.free()
        # nothing left alive: the context can be destroyed
        c.close()
        assert not c.h
    finally:
        for dl, dr in dev:
            dl.free()
            dr.free()
