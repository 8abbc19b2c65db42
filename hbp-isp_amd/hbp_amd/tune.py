"""Kernel-form tuner: picks a context's kernel forms for a workload by timing it on the device.

Each value is a choice between forms that give the same results bit for bit (the plane-cost
cluster sums are exact integers in any order), so tuning only trades speed:
  * coarse_min_pix: average pixels per superpixel from which a BP level costs planes in the
    per-superpixel segment form instead of chunks of sorted pixels;
  * level_form.<l>: that choice for one BP level, overriding coarse_min_pix (tune_levels sets a
    schedule of forms level by level);
  * single_hb / batch_hb: steps per batch of a segment wavefront (2 or 4) in one-pair and in
    batched launches: 4 keeps more gathers in flight, 2 fewer registers and more wavefronts;
  * seg_len: pixels per segment wavefront of the segment form (a multiple of 64 up to 512);
  * hamming_cols: columns per lane of the Hamming cost kernel (4 or 8);
  * batch_graph: batched rounds queued one by one or replayed from one captured HIP graph;
  * pipe_blocks / pipe_batch / pipe_hb: the persistent launches that compute RANSAC sweeps'
    plane costs ahead on a second stream (fuse 2): their blocks (fewer leave more of every CU to
    the chain of message and node launches), RANSAC sweeps per launch, steps per batch.
The reference's kernels have one form each; this module has no counterpart there.
"""
import itertools
import time

SINGLE_GRID = {"coarse_min_pix": (32, 64, 96, 128, 256), "single_hb": (4, 2)}
PIPE_GRID = {"pipe_blocks": (384, 512, 640, 768), "pipe_batch": (4, 8)}
BATCH_GRID = {"coarse_min_pix": (64, 96, 128), "batch_hb": (2, 4)}


def _timed(ctx, fn, reps):
    """best of reps timed calls after one untimed call (seconds)"""
    fn()
    ctx.sync()
    best = float("inf")
    for _ in range(reps):
        t = time.perf_counter()
        fn()
        ctx.sync()
        best = min(best, time.perf_counter() - t)
    return best


def _search(ctx, grid, run, reps):
    keys = list(grid)
    before = ctx.get_tuning()
    times = {}
    try:
        for values in itertools.product(*(grid[k] for k in keys)):
            for k, v in zip(keys, values):
                ctx.set_tuning(k, v)
            times[values] = _timed(ctx, run, reps)
    finally:
        for k, v in before.items():
            ctx.set_tuning(k, v)
    best = min(times, key=times.get)
    for k, v in zip(keys, best):
        ctx.set_tuning(k, v)
    return {"best": dict(zip(keys, best)),
            "ms": {"/".join(f"{k}={v}" for k, v in zip(keys, vals)): round(1e3 * t, 3) for vals, t in times.items()}}


def tune_single(ctx, left, right, D, min_level=2, grid=None, reps=2):
    """Times one pair's BP (plane costs, messages, node updates, MAP) for every combination of
    the grid and leaves the fastest set on ctx."""
    vol = ctx.cost_volume(left, right, D)
    pyr = ctx.pyramid(left)
    ml = max(1, min(min_level, pyr.levels - 1))

    def run():
        ctx.bp(pyr, vol, min_level=ml).run()

    return _search(ctx, grid or SINGLE_GRID, run, reps)


def tune_batch(ctx, pairs, min_level=2, grid=None, reps=1):
    """Times the batched BP of pairs [(left, right, D), ...] (one sweep of every pair per round)
    for every combination of the grid and leaves the fastest set on ctx."""
    prepared = []
    for left, right, D in pairs:
        vol = ctx.cost_volume(left, right, D)
        pyr = ctx.pyramid(left)
        prepared.append((vol, pyr, max(1, min(min_level, pyr.levels - 1))))

    def run():
        bps = [ctx.bp(pyr, vol, min_level=ml) for vol, pyr, ml in prepared]
        ctx.bp_run_batch(bps)

    return _search(ctx, grid or BATCH_GRID, run, reps)


def tune_levels(ctx, left, right, D, min_level=2, reps=3):
    """A schedule of plane-cost forms, level by level, for one pair's BP: starting from the
    current rule, each BP level in turn takes the other form if that makes the whole BP faster
    (best of reps runs).  Leaves the schedule on ctx; returns it with the times."""
    vol = ctx.cost_volume(left, right, D)
    pyr = ctx.pyramid(left)
    L = pyr.levels
    ml = max(1, min(min_level, L - 1))

    def run():
        ctx.bp(pyr, vol, min_level=ml).run()

    # the forms the current rule gives, read back from a BP of this pair: level l is coarse when
    # its average superpixel holds at least coarse_min_pix pixels
    N = pyr.W * pyr.H
    cmp_ = ctx.get_tuning("coarse_min_pix")
    levels = [l for l in range(ml, L - 1) if l < 24]
    form = {l: int(N >= cmp_ * pyr.level_size(l)[0]) for l in levels}
    for l in levels:
        ctx.set_tuning(f"level_form.{l}", form[l])
    base = _timed(ctx, run, reps)
    trace = {"rule_ms": round(1e3 * base, 3)}
    for l in levels:
        ctx.set_tuning(f"level_form.{l}", 1 - form[l])
        t = _timed(ctx, run, reps)
        trace[f"L{l}:{'segments' if 1 - form[l] else 'chunks'}_ms"] = round(1e3 * t, 3)
        if t < base:
            base = t
            form[l] = 1 - form[l]
        else:
            ctx.set_tuning(f"level_form.{l}", form[l])
    return {"schedule": {f"L{l}": "segments" if form[l] else "chunks" for l in levels},
            "ms": round(1e3 * base, 3), "trace": trace}


def tune_hamming(ctx, left, right, D, reps=3):
    """Columns per lane of the Hamming cost kernel: times the cost volume (census, WTA and the
    u8 volume) for 4 and 8; leaves the faster on ctx."""
    lib = ctx.lib

    def run():
        v = ctx.cost_volume(left, right, D)
        lib.hbp_volume_dev_u8(v.h)
        v.free()

    return _search(ctx, {"hamming_cols": (4, 8)}, run, reps)
