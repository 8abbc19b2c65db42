"""The tuner's search logic on the CPU, with a stand-in context whose run time is a known
function of the settings: the fastest grid point is chosen and left set, every point is
timed, and a failing run restores the settings it started from."""
import time

import pytest

from hbp_amd import tune


class FakeCtx:
    def __init__(self):
        self.t = {"coarse_min_pix": 96, "single_hb": 4, "batch_hb": 2}

    def get_tuning(self, name=None):
        return dict(self.t) if name is None else self.t[name]

    def set_tuning(self, name, value):
        self.t[name] = value

    def sync(self):
        pass


def test_search_picks_the_fastest_and_times_every_point():
    ctx = FakeCtx()
    cost = {(32, 4): 3e-3, (32, 2): 1e-3, (64, 4): 2e-3, (64, 2): 4e-3}
    run = lambda: time.sleep(cost[(ctx.t["coarse_min_pix"], ctx.t["single_hb"])])
    r = tune._search(ctx, {"coarse_min_pix": (32, 64), "single_hb": (4, 2)}, run, reps=2)
    assert r["best"] == {"coarse_min_pix": 32, "single_hb": 2}
    assert ctx.t["coarse_min_pix"] == 32 and ctx.t["single_hb"] == 2 and ctx.t["batch_hb"] == 2
    assert len(r["ms"]) == 4 and r["ms"]["coarse_min_pix=32/single_hb=2"] < r["ms"]["coarse_min_pix=64/single_hb=2"]


def test_search_restores_the_settings_when_a_run_fails():
    ctx = FakeCtx()

    def run():
        if ctx.t["batch_hb"] == 4:
            raise RuntimeError("device error")

    with pytest.raises(RuntimeError):
        tune._search(ctx, {"batch_hb": (2, 4)}, run, reps=1)
    assert ctx.t == {"coarse_min_pix": 96, "single_hb": 4, "batch_hb": 2}
