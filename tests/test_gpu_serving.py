"""Request server and context lifetimes on the GPU: failures forced from the host (an
allocation that fails during one pair's admission, a batched round that fails before it
launches), cancellation of queued and running pairs, and contexts that refuse to be destroyed
under live handles.  Every kernel exits normally in these tests: a forced failure throws on the
host before the work it stands for is queued."""
import ctypes as C
import time

import numpy as np
import pytest

from hbp_amd import lib as H
from hbp_amd import synth

pytestmark = pytest.mark.gpu


def say(msg):
    print(f"[case] {msg}", flush=True)


def same_bits(a, b):
    a, b = np.ascontiguousarray(a, np.float32), np.ascontiguousarray(b, np.float32)
    return a.shape == b.shape and np.array_equal(a.view(np.uint32), b.view(np.uint32))


def _pair(seed, W, H_, D):
    left, right, _ = synth.make_pair(seed, W, H_, D, (3, 8))
    return left, right


def _single(ctx, left, right, D, ml):
    pyr = ctx.pyramid(left)
    ml = max(1, min(ml, pyr.levels - 1))
    return ctx.bp(pyr, ctx.cost_volume(left, right, D), min_level=ml).run()


def _drive(srv, tickets, limit=400000):
    """Steps until every ticket has a result or an error; returns {ticket: result or error}."""
    out = {}
    for _ in range(limit):
        srv.step()
        for t in tickets:
            if t in out:
                continue
            try:
                r = srv.poll(t)
            except H.HbpError as e:
                out[t] = e
                continue
            if r is not None:
                out[t] = r
        if len(out) == len(tickets):
            return out
    raise AssertionError("server did not finish")


@pytest.fixture(autouse=True)
def _disarm():
    yield
    H.inject_fault("alloc", 0)
    H.inject_fault("round", 0)


def _step_until_active(srv, n_active, limit=20000, tickets=()):
    """Steps until n_active pairs are on the device.  Between steps the main stream drains: the
    admission stream may share a hardware queue with it (a process holds more streams than the
    device's hardware queues), and then admission work runs only behind the queued rounds."""
    a = p = 0
    for _ in range(limit):
        a, p, _ = srv.step()
        if a >= n_active and p == 0:
            return
        time.sleep(0.0005)
    errors = {}
    for t in tickets:
        try:
            errors[t] = "done" if srv.poll(t) is not None else "pending"
        except H.HbpError as e:
            errors[t] = str(e)
    raise AssertionError(f"pairs were not admitted: active {a}, queued {p}, requests {errors}")


def test_allocation_failure_fails_one_admission_only(ctx):
    """A device allocation fails while one pair is being admitted and another runs: that pair
    polls as an error, the running one and a later one finish with their own results."""
    say("server: injected allocation failure during an admission")
    z = (_pair(101, 320, 240, 48), 48, 1)
    x = (_pair(102, 96, 72, 24), 24, 2)
    y = (_pair(103, 120, 90, 32), 32, 2)
    srv = H.Server(ctx, max_active=4)
    tz = srv.submit(z[0][0], z[0][1], z[1], z[2])
    _step_until_active(srv, 1)
    srv.step()
    H.inject_fault("alloc", 1)       # the next device allocation: x's admission
    tx = srv.submit(x[0][0], x[0][1], x[1], x[2])
    out = _drive(srv, [tx])
    assert isinstance(out[tx], H.HbpError) and "injected" in str(out[tx])
    ty = srv.submit(y[0][0], y[0][1], y[1], y[2])
    out = _drive(srv, [tz, ty])
    for t, ((left, right), D, ml) in ((tz, z), (ty, y)):
        d1, p1 = _single(ctx, left, right, D, ml)
        assert same_bits(d1, out[t][0]) and same_bits(p1, out[t][1])
    srv.close()


def test_round_failure_fails_its_pairs_and_the_server_keeps_serving(ctx):
    say("server: injected round failure")
    # a's rounds (min_level 1 on 320x240: about 1500 sweeps) outlast b's admission, so both are
    # on the device together before the failing round
    a = (_pair(111, 320, 240, 48), 48, 1)
    b = (_pair(112, 64, 48, 16), 16, 1)
    c = (_pair(113, 120, 90, 32), 32, 2)
    srv = H.Server(ctx, max_active=4)
    ta = srv.submit(a[0][0], a[0][1], a[1], a[2])
    tb = srv.submit(b[0][0], b[0][1], b[1], b[2])
    _step_until_active(srv, 2, tickets=(ta, tb))
    H.inject_fault("round", 1)
    out = _drive(srv, [ta, tb])
    assert all(isinstance(out[t], H.HbpError) and "round" in str(out[t]) for t in (ta, tb))
    tc = srv.submit(c[0][0], c[0][1], c[1], c[2])
    out = _drive(srv, [tc])
    d1, p1 = _single(ctx, c[0][0], c[0][1], c[1], c[2])
    assert same_bits(d1, out[tc][0]) and same_bits(p1, out[tc][1])
    srv.close()


def test_cancel_queued_and_running_pairs(ctx):
    say("server: cancel")
    a = (_pair(121, 160, 120, 32), 32, 1)
    b = (_pair(122, 96, 72, 24), 24, 2)
    c = (_pair(123, 64, 48, 16), 16, 2)
    d = (_pair(124, 80, 64, 16), 16, 2)
    srv = H.Server(ctx, max_active=2)
    ta = srv.submit(a[0][0], a[0][1], a[1], a[2])
    tb = srv.submit(b[0][0], b[0][1], b[1], b[2])
    tc = srv.submit(c[0][0], c[0][1], c[1], c[2])   # queued behind the two slots
    for _ in range(50):
        srv.step()
    srv.cancel(tb)                                   # admitted or running
    srv.cancel(tc)                                   # still queued
    td = srv.submit(d[0][0], d[0][1], d[1], d[2])
    out = _drive(srv, [ta, tb, tc, td])
    for t in (tb, tc):
        assert isinstance(out[t], H.HbpError) and "cancelled" in str(out[t])
    for t, ((left, right), D, ml) in ((ta, a), (td, d)):
        d1, p1 = _single(ctx, left, right, D, ml)
        assert same_bits(d1, out[t][0]) and same_bits(p1, out[t][1])
    with pytest.raises(H.HbpError):
        srv.cancel(987654)
    srv.close()


def test_context_refuses_destroy_under_live_handles():
    say("context lifetimes")
    lib = H.load()
    c2 = H.Ctx(0)
    left, right = _pair(131, 64, 48, 16)
    vol = c2.cost_volume(left, right, 16)
    pyr = c2.pyramid(left)
    assert lib.hbp_ctx_destroy(c2.h) != 0
    assert b"still alive" in lib.hbp_last_error()
    # a handle freed through another context is an error and stays alive
    c3 = H.Ctx(0)
    assert lib.hbp_volume_free(c3.h, vol.h) != 0
    c3.close()
    # close() frees the children first; their own free() afterwards does nothing
    c2.close()
    vol.free()
    pyr.free()
    assert c2.h is None


def test_raw_handles_then_destroy():
    say("raw C handles")
    lib = H.load()
    h = C.c_void_p()
    assert lib.hbp_ctx_create(0, C.byref(h)) == 0
    left, right = _pair(141, 64, 48, 16)
    v = C.c_void_p()
    assert lib.hbp_cost_volume(h, np.ascontiguousarray(left), np.ascontiguousarray(right), 64, 48, 16, b"Census",
                               C.byref(v)) == 0
    assert lib.hbp_ctx_destroy(h) != 0
    assert lib.hbp_volume_free(None, v) == 0      # NULL context: the handle's own is used
    assert lib.hbp_ctx_destroy(h) == 0


def test_backpressure_refuses_beyond_the_queue_bound(ctx):
    say("server: backpressure")
    pairs = [(_pair(151 + i, 96, 72, 24), 24, 2) for i in range(5)]
    srv = H.Server(ctx, max_active=1)
    srv.set_max_pending(2)
    tickets = []
    refused = 0
    for (l, r), d, ml in pairs:
        try:
            tickets.append((srv.submit(l, r, d, ml), (l, r, d, ml)))
        except H.HbpError as e:
            assert "queue full" in str(e)
            refused += 1
    assert refused == 3 and len(tickets) == 2
    assert srv.stats()[2] == 3
    out = _drive(srv, [t for t, _ in tickets])
    for t, (l, r, d, ml) in tickets:
        d1, p1 = _single(ctx, l, r, d, ml)
        assert same_bits(d1, out[t][0]) and same_bits(p1, out[t][1])
    (l, r), d, ml = pairs[4]
    t = srv.submit(l, r, d, ml)          # room again
    assert _drive(srv, [t])[t] is not None
    used, peak, _ = srv.stats()
    assert used == 0 and peak > 0
    srv.close()
