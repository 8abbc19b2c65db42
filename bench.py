"""Throughput of the HBP-ISP stereo hot path on one MI355X.

A step is one stereo pair through the plain path (census + Hamming cost volume, CRTrees
superpixel pyramid, hierarchical BP down to min_level, MAP to disparity and planes), with the
pair's images already resident in HBM and the outputs left there.  Default workload: the
1472x984 RGB pair with max disparity 256 and 20-40 slanted planes, BP on 4 pyramid levels
(min_level = levels - 5), reference parameters otherwise (alpha 1000.1, tau_smooth 0.6).

Prints one JSON line.  `roofline` is the dominant kernel's algorithmic bytes over its
HIP-event time inside the timed region; `cpu_baseline` is the oracle (host C++ with OpenMP)
on a bounded sample of the same workload.
"""
import argparse
import json
import os
import statistics
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "hbp-isp_amd"))

HBM_PEAK_GBS = 8000.0
# timed region: one launch in 29 of the dominant kernel carries start/stop events stamped by its
# dispatch (kernel execution time, as a kernel trace measures it); 29 is coprime with the sweep
# kinds' period (4 or 5 sweeps per BP iteration), so the sample covers every kind evenly, and
# with 1754 launches per pair ten steps still time about 600 launches while the events' own
# cost (about 8 us per sampled launch) stays near 1 % of the step
EVENT_EVERY = 29
# the RANSAC plane costs computed ahead (340 launches per pair, on the second stream): one in 7
AHEAD_EVENT_EVERY = 7
# wavefront VALU instructions per second of the whole chip: 256 CUs x 4 SIMDs, one wave64
# instruction per SIMD every 4 cycles at 2.4 GHz
VALU_PEAK_WAVE_INSTR = 256 * 4 * 2.4e9 / 4
# profile names of the kernels whose counter summaries profiles/ holds, and their rocprof names
PROFILE_KERNELS = {"plane_cost": "cost_edge_kernel<4, true, 0, 2>", "plane_cost_ahead": "pc_ahead_kernel<0, 2>"}
# host threads (each with a stream of its own) that build the batched leg's pairs
PREP_WORKERS = 8
METRIC = "stereo pairs/s + Mpix·disp/s; p50/p99 latency, batched/served vs plain"
WORKLOADS = {
    "large": dict(width=1472, height=984, ndisp=256, n_planes=(20, 40), bp_levels=4,
                  name="plain method, one large pair: 1472x984 RGB u8, max disparity 256, 20-40 slanted "
                       "planes, BP on 4 pyramid levels"),
    "small": dict(width=450, height=375, ndisp=64, n_planes=(6, 12), bp_levels=0,
                  name="plain method, one pair: 450x375 RGB u8, max disparity 64, 6-12 slanted planes, "
                       "min_level 2"),
}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--workload", choices=sorted(WORKLOADS), default="large")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--batch-pairs", type=int, default=64,
                    help="pairs of the batched-vs-plain comparison (mixed sizes; 0 skips it)")
    ap.add_argument("--serve-seconds", type=float, default=4.0,
                    help="length of the Poisson arrival window of the served-vs-plain comparison (0 skips it)")
    ap.add_argument("--serve-rate", type=float, default=30.0,
                    help="mean arrivals per second (above both paths' capacity, so that pairs/s measures "
                         "capacity rather than the offered load)")
    ap.add_argument("--no-post", action="store_true",
                    help="skip the refinements after BP (right view, LR check, weighted median, 7-tap median)")
    ap.add_argument("--cpu-sweeps", type=int, default=0,
                    help="sweeps of the one-thread CPU sample (0: about 12 s worth)")
    ap.add_argument("--traffic-calibration", action="store_true",
                    help="run only the library's 256 MiB float4 copy kernel three times (the known traffic that "
                         "rocprofv3's FETCH_SIZE / WRITE_SIZE counts are scaled by) and exit")
    ap.add_argument("--traffic-summary", nargs=4, metavar=("FETCH_CSV", "WRITE_CSV", "CAL_FETCH_CSV", "CAL_WRITE_CSV"),
                    help="no GPU: the dominant kernel's HBM bytes per launch from rocprofv3 --pmc FETCH_SIZE and "
                         "WRITE_SIZE counter CSVs of a bench run and of a --traffic-calibration run; prints JSON")
    ap.add_argument("--serve-spec", action="store_true",
                    help="only the serving sweep: open-loop Poisson arrivals of the serving mix at 10 and 15 pairs/s "
                         "(below capacity, >= 1000 requests each) and 50, 100, 200, 400 pairs/s for 60 s with "
                         "backpressure; prints one JSON line")
    ap.add_argument("--stress-spec", action="store_true",
                    help="only the memory-pressure run: 3840x2160 D384 pairs (BP on 5 levels) interleaved with the "
                         "64-pair mixed batch, served under a 16 GB and a 256 GB budget; prints one JSON line")
    ap.add_argument("--mix-levels", action="store_true",
                    help="only the per-level split of the 64-pair mixed batch's plain BP time (each level's time as "
                         "the difference of runs ending one level apart); prints one JSON line")
    ap.add_argument("--batch-kernels", action="store_true",
                    help="only the batched BP kernels over the 64-pair mixed batch: each kernel's time "
                         "(HIP events around every launch), launches and algorithmic GB/s against the HBM peak")
    ap.add_argument("--trace-summary", nargs=2, metavar=("KERNEL_STATS_CSV", "PAIRS"), default=None,
                    help="per-pair table of a rocprofv3 --kernel-trace --stats kernel_stats.csv over PAIRS pairs")
    ap.add_argument("--valu-summary", metavar="VALU_CSV", default=None,
                    help="summarise a rocprofv3 --pmc SQ_INSTS_VALU CSV: wavefront VALU instructions per launch "
                         "of the kernel --traffic-kernel names")
    ap.add_argument("--traffic-kernel", default=None,
                    help="with --traffic-summary: the kernel (a name prefix) to summarise instead of the plain path's "
                         "plane-cost kernel")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the default kernel forms instead of timing the candidates first (the forms "
                         "change speed only, never the results)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the stereo path runs one pair per GPU and this build has "
              "no multi-GPU runner; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    if args.traffic_summary:
        print(json.dumps(traffic_summary(*args.traffic_summary, kernel=args.traffic_kernel), indent=1))
        return
    if args.trace_summary:
        print(trace_summary(args.trace_summary[0], int(args.trace_summary[1])))
        return
    if args.valu_summary:
        print(json.dumps(valu_summary(args.valu_summary, kernel=args.traffic_kernel), indent=1))
        return

    import ctypes as C
    import numpy as np
    from hbp_amd import lib as H
    from hbp_amd import synth

    if args.serve_spec:
        pool = make_serve_pool(args.seed, per_class=4)
        print(json.dumps(serve_spec(pool, args.seed)), flush=True)
        return
    if args.mix_levels:
        pairs = make_batch_pairs(64, args.seed)
        print(json.dumps(mix_levels(pairs)), flush=True)
        return
    if args.stress_spec:
        print(json.dumps(stress_spec(args.seed)), flush=True)
        return
    if args.batch_kernels:
        pairs = make_batch_pairs(64, args.seed)
        print(json.dumps(batch_kernels(pairs)), flush=True)
        return

    if args.traffic_calibration:
        ctx = H.Ctx(0)
        H._check(ctx.lib.hbp_profile_copy(ctx.h, 256 << 20, 3))
        ctx.close()
        print(json.dumps({"calibration": "calib_copy_kernel: 3 launches, each reading and writing 262144 KiB"}))
        return

    wl = WORKLOADS[args.workload]
    W, Hh, D = wl["width"], wl["height"], wl["ndisp"]
    left, right, _ = synth.make_pair(args.seed, W, Hh, D, wl["n_planes"])
    # the mixed-size pairs of the batched comparison, generated before the GPU is touched
    batch_pairs = make_batch_pairs(args.batch_pairs, args.seed) if args.batch_pairs > 0 else None
    serve_pool = make_serve_pool(args.seed) if args.serve_seconds > 0 else None

    ctx = H.Ctx(0)
    lib = ctx.lib
    hip = C.CDLL("libamdhip64.so")
    nbytes = W * Hh * 3
    d_l, d_r = C.c_void_p(), C.c_void_p()
    assert hip.hipMalloc(C.byref(d_l), C.c_size_t(nbytes)) == 0
    assert hip.hipMalloc(C.byref(d_r), C.c_size_t(nbytes)) == 0
    assert hip.hipMemcpy(d_l, left.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes), 1) == 0
    assert hip.hipMemcpy(d_r, right.ctypes.data_as(C.c_void_p), C.c_size_t(nbytes), 1) == 0

    info = {}

    def one_pair():
        vol, pyr, bp = C.c_void_p(), C.c_void_p(), C.c_void_p()
        H._check(lib.hbp_cost_volume_dev(ctx.h, d_l, d_r, W, Hh, D, b"Census", C.byref(vol)))
        H._check(lib.hbp_pyramid_build_dev(ctx.h, d_l, W, Hh, 2, 4, 0.0, C.byref(pyr)))
        L = lib.hbp_pyramid_levels(pyr)
        ml = max(1, L - 1 - wl["bp_levels"]) if wl["bp_levels"] > 0 else 2
        H._check(lib.hbp_bp_create(ctx.h, pyr, vol, 1000.1, 0.6, ml, 1, C.byref(bp)))
        info.update(levels=L, min_level=ml, sweeps=lib.hbp_bp_total_sweeps(bp))
        H._check(lib.hbp_bp_run(ctx.h, bp, None, None))
        lib.hbp_bp_free(ctx.h, bp)
        lib.hbp_pyramid_free(ctx.h, pyr)
        lib.hbp_volume_free(ctx.h, vol)

    def timed(steps):
        lat = []
        ctx.sync()
        t0 = time.perf_counter()
        for _ in range(steps):
            ts = time.perf_counter()
            one_pair()
            ctx.sync()
            lat.append(time.perf_counter() - ts)
        ctx.sync()
        return time.perf_counter() - t0, lat

    one_pair()   # fills info (levels, min_level) and the context's memory pool
    tuning = {}
    if not args.no_autotune:
        # single_hb for this pair's shape; coarse_min_pix matters for the mixed sizes below (the
        # coarse levels of this pair all hold thousands of pixels per superpixel).  The segment
        # length (512) and the sweeps per ahead launch (4) stay at the shape the dominant
        # kernel's traffic and VALU profiles were measured at, so that `roofline` reads them for
        # the launches it times (their other values measured within the noise of these)
        from hbp_amd import tune
        tuning["single"] = tune.tune_single(ctx, left, right, D, info["min_level"], grid={"single_hb": (4, 2)})
        tuning["levels"] = tune.tune_levels(ctx, left, right, D, info["min_level"])
        tuning["pipe"] = tune.tune_single(ctx, left, right, D, info["min_level"],
                                          grid={"pipe_blocks": tune.PIPE_GRID["pipe_blocks"], "pipe_batch": (4,)})
        tuning["hamming"] = tune.tune_hamming(ctx, left, right, D)
    for _ in range(args.warmup):
        one_pair()
    ctx.sync()
    # the dominant kernel: the most time in one pair with every kernel timed (outside the timed
    # region); with the RANSAC plane costs computed ahead on the second stream it is that launch
    ctx.profile(True)
    ctx.profile_reset()
    one_pair()
    ctx.sync()
    first_stats = ctx.profile_stats()
    ctx.profile(False)
    dom = max(group(first_stats).items(), key=lambda kv: kv[1]["ms"])[0]
    # the ahead launches are few (one per 4 sweeps) and off the chain's critical path: sample more
    every = EVENT_EVERY if dom == "plane_cost" else AHEAD_EVENT_EVERY
    # timed region: HIP events around the dominant kernel's launches only
    ctx.profile(True)
    ctx.profile_only(dom, every)
    ctx.profile_reset()
    elapsed, lat = timed(args.steps)
    stats = ctx.profile_stats()
    ctx.profile(False)
    ctx.profile_only(None)
    # the same steps without any events (the events' own cost)
    elapsed_plain, _ = timed(args.steps)
    # the host-buffer boundary (images from host memory, disparity + planes back to host):
    # PCIe-inclusive time per pair, outside the timed region
    disp_h = np.zeros((Hh, W), np.float32)
    plane_h = np.zeros((Hh, W, 3), np.float32)

    def one_pair_host():
        vol, pyr, bp = C.c_void_p(), C.c_void_p(), C.c_void_p()
        H._check(lib.hbp_cost_volume(ctx.h, left, right, W, Hh, D, b"Census", C.byref(vol)))
        H._check(lib.hbp_pyramid_build(ctx.h, left, W, Hh, 2, 4, 0.0, C.byref(pyr)))
        H._check(lib.hbp_bp_create(ctx.h, pyr, vol, 1000.1, 0.6, info["min_level"], 1, C.byref(bp)))
        H._check(lib.hbp_bp_run(ctx.h, bp, disp_h.ctypes.data_as(C.c_void_p), plane_h.ctypes.data_as(C.c_void_p)))
        lib.hbp_bp_free(ctx.h, bp)
        lib.hbp_pyramid_free(ctx.h, pyr)
        lib.hbp_volume_free(ctx.h, vol)

    one_pair_host()
    ctx.sync()
    th = time.perf_counter()
    for _ in range(3):
        one_pair_host()
    ms_host = 1e3 * (time.perf_counter() - th) / 3

    # one more pair with every kernel timed, outside the timed region: per-kernel breakdown
    ctx.profile(True)
    ctx.profile_reset()
    one_pair()
    ctx.sync()
    pair_stats = ctx.profile_stats()
    ctx.profile(False)

    # the cost-volume kernels: census, the WTA pass of the plain path, and the u8 volume a caller
    # of the volume getters gets (written on first request; BP reads only census words and WTA)
    ctx.profile(True)
    ctx.profile_reset()
    for _ in range(3):
        vol = C.c_void_p()
        H._check(lib.hbp_cost_volume_dev(ctx.h, d_l, d_r, W, Hh, D, b"Census", C.byref(vol)))
        if not lib.hbp_volume_dev_u8(vol):
            raise H.HbpError(lib.hbp_last_error().decode())
        lib.hbp_volume_free(ctx.h, vol)
    ctx.sync()
    vol_stats = ctx.profile_stats()
    ctx.profile(False)
    cost_kernels = {}
    for name in ("census", "hamming_wta", "hamming_volume"):
        v = vol_stats.get(name)
        if v and v["ms"] > 0:
            gbs = v["bytes"] / (v["ms"] * 1e-3) / 1e9
            cost_kernels[name] = dict(us=round(1e3 * v["ms"] / v["launches"], 1), gbs=round(gbs, 1),
                                      frac=round(gbs / HBM_PEAK_GBS, 4),
                                      bytes_per_launch=round(v["bytes"] / v["launches"]))

    pairs_per_s = args.steps / elapsed
    mpix_disp = W * Hh * D / 1e6 * pairs_per_s
    lat_ms = sorted(1e3 * x for x in lat)
    p50 = statistics.median(lat_ms)
    p99 = lat_ms[min(len(lat_ms) - 1, int(round(0.99 * (len(lat_ms) - 1))))]

    gd = group(stats).get(dom, dict(ms=0.0, bytes=0.0, launches=0))
    avg_ms = gd["ms"] / max(gd["launches"], 1)
    avg_bytes = gd["bytes"] / max(gd["launches"], 1)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    kernels = {g: dict(ms_per_pair=round(v["ms"], 3), launches=v["launches"],
                       gbs=round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None)
               for g, v in sorted(group(pair_stats).items(), key=lambda kv: -kv[1]["ms"])}

    # HBM bytes per launch of the dominant kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE runs,
    # and its VALU instructions per launch (SQ_INSTS_VALU), from the committed summaries
    traffic, valu = None, None
    # the summaries were taken with the default shape (segments of 512 pixels, 4 RANSAC sweeps
    # per ahead launch); launches of another shape move other bytes and instructions
    profiled_shape = ctx.get_tuning("seg_len") == 512 and ctx.get_tuning("pipe_batch") == 4
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_{args.workload}.json")
    vfile = os.path.join(ROOT, "profiles", f"valu_{dom}_{args.workload}.json")
    if profiled_shape and os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None
    if profiled_shape and os.path.exists(vfile) and avg_ms > 0:
        try:
            vi = json.load(open(vfile)).get("valu_instr_per_launch")
            if vi:
                rate = vi / (avg_ms * 1e-3)
                valu = {"achieved": round(rate / 1e9, 1), "peak": round(VALU_PEAK_WAVE_INSTR / 1e9, 1),
                        "unit": "G wave-instr/s", "frac": round(rate / VALU_PEAK_WAVE_INSTR, 4),
                        "instr_per_launch": vi}
        except (OSError, ValueError):
            valu = None

    post, post_io = (None, None) if args.no_post else post_process(ctx, left, right, D, info["min_level"])
    if batch_pairs and not args.no_autotune:
        from hbp_amd import tune
        sub = [(left_b, right_b, spec[3]) for spec, left_b, right_b in batch_pairs[:8]]
        tuning["batch"] = tune.tune_batch(ctx, sub, 2)
    batched = batched_vs_plain(ctx, hip, batch_pairs) if batch_pairs else None
    served = served_vs_plain(ctx, serve_pool, args.serve_rate, args.serve_seconds, args.seed) if serve_pool else None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(left, right, wl, info, args.cpu_sweeps, disp_h, plane_h)
        if post is not None:
            post.update(cpu_weighted_median(*post_io))

    out = {
        "metric": METRIC,
        "value": round(pairs_per_s, 4),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * elapsed / args.steps, 3),
        "ms_per_step_without_events": round(1e3 * elapsed_plain / args.steps, 3),
        "ms_per_pair_host_buffers": round(ms_host, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (u8 costs, int64 cluster sums)",
        "data": "synthetic slanted-plane stereo pair (seeded generator)",
        "mpix_disp_per_s": round(mpix_disp, 2),
        "p50_ms": round(p50, 3),
        "p99_ms": round(p99, 3),
        "config": {"workload": wl["name"], "width": W, "height": Hh, "max_disp": D, "levels": info["levels"],
                   "min_level": info["min_level"], "sweeps": info["sweeps"], "alpha": 1000.1,
                   "tau_smooth": 0.6, "seed": args.seed},
        "roofline": {"bound": "hbm", "kernel": dom, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic, "algorithmic_bytes_per_launch": round(avg_bytes),
                     "avg_launch_us": round(1e3 * avg_ms, 2), "launches_timed": gd["launches"],
                     "event_sampling": f"1 in {every} launches in the timed region", "valu": valu},
        "kernels_one_pair": kernels,
        "cost_volume_kernels": cost_kernels,
        "launch_us_by_level": {n: round(1e3 * v["ms"] / max(v["launches"], 1), 2)
                               for n, v in sorted(pair_stats.items()) if "_L" in n},
        "tuning": dict(tuning, final=dict(ctx.get_tuning(), **{
            f"level_form.{l}": ctx.get_tuning(f"level_form.{l}") for l in range(info["min_level"], info["levels"] - 1)})),
        "batched_vs_plain": batched,
        "served_vs_plain": served,
        "post_process": post,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    hip.hipFree(d_l)
    hip.hipFree(d_r)
    ctx.close()


DOMINANT_KERNEL = "cost_edge_kernel<4, true, 0, 4>"   # plane cost, coarse form, left view, 4-step batches
CALIB_KIB = 262144.0                                  # the calibration copy reads and writes 256 MiB


def group(st):
    """Per-kernel profile entries summed over BP levels (names "<kernel>_L<level>")."""
    out = {}
    for name, v in st.items():
        g = name.split("_L")[0]
        a = out.setdefault(g, dict(ms=0.0, bytes=0.0, launches=0))
        a["ms"] += v["ms"]
        a["bytes"] += v["bytes"]
        a["launches"] += v["launches"]
    return out


def traffic_summary(fetch_csv, write_csv, cal_fetch_csv, cal_write_csv, kernel=None):
    """HBM bytes per launch of `kernel`: the mean FETCH_SIZE and WRITE_SIZE (KiB per dispatch)
    over its launches, each scaled by the ratio of the known bytes to the counter's reading on
    the calibration copy (calib_copy_kernel)."""
    import csv

    def per_launch(path, name):
        vals, grids = [], {}
        with open(path) as f:
            for row in csv.DictReader(f):
                if name in row["Kernel_Name"]:
                    v = float(row["Counter_Value"])
                    vals.append(v)
                    grids.setdefault(row["Grid_Size"], []).append(v)
        return vals, grids

    kernel = PROFILE_KERNELS.get(kernel, kernel) if kernel else DOMINANT_KERNEL
    cf, _ = per_launch(cal_fetch_csv, "calib_copy_kernel")
    cw, _ = per_launch(cal_write_csv, "calib_copy_kernel")
    sf = CALIB_KIB / (sum(cf) / len(cf))
    sw = CALIB_KIB / (sum(cw) / len(cw))
    f, fg = per_launch(fetch_csv, kernel)
    w, wg = per_launch(write_csv, kernel)
    fk, wk = sum(f) / len(f), sum(w) / len(w)
    return {"kernel": kernel, "bytes_per_launch": round((fk * sf + wk * sw) * 1024),
            "fetch_kib_per_launch_raw": round(fk, 1), "write_kib_per_launch_raw": round(wk, 1),
            "scale_fetch": round(sf, 4), "scale_write": round(sw, 4),
            "calibration": f"calib_copy_kernel ({CALIB_KIB:.0f} KiB read + written per launch): FETCH_SIZE reads "
                           f"{sum(cf) / len(cf):.2f} KiB, WRITE_SIZE {sum(cw) / len(cw):.2f} KiB per dispatch",
            "launches": len(f),
            "by_grid_size": {g: {"launches": len(v), "fetch_kib": round(sum(v) / len(v), 2),
                                 "write_kib": round(sum(wg.get(g, [0])) / max(len(wg.get(g, [])), 1), 2)}
                             for g, v in sorted(fg.items())}}


def trace_summary(stats_csv, pairs):
    """Launches per pair, mean duration and time per pair of every kernel in a rocprofv3
    kernel_stats.csv that covers `pairs` pairs."""
    import csv
    rows = []
    with open(stats_csv) as f:
        for row in csv.DictReader(f):
            rows.append((row["Name"], int(row["Calls"]), float(row["TotalDurationNs"])))
    rows.sort(key=lambda r: -r[2])
    total = sum(r[2] for r in rows)
    lines = [f"total kernel time {total / 1e6:.3f} ms over {pairs} pairs = {total / 1e6 / pairs:.3f} ms per pair "
             "(kernels of both streams, overlapping ones counted each)",
             "  per pair    avg_us  ms/pair  kernel"]
    for name, calls, ns in rows:
        lines.append(f"{calls / pairs:10.1f} {ns / calls / 1e3:9.2f} {ns / 1e6 / pairs:8.3f}  {name[:110]}")
    return "\n".join(lines)


def valu_summary(valu_csv, kernel=None):
    """Wavefront VALU instructions per launch of `kernel`: the mean SQ_INSTS_VALU per dispatch."""
    import csv
    kernel = PROFILE_KERNELS.get(kernel, kernel) if kernel else DOMINANT_KERNEL
    vals = []
    with open(valu_csv) as f:
        for row in csv.DictReader(f):
            if kernel in row["Kernel_Name"]:
                vals.append(float(row["Counter_Value"]))
    return {"kernel": kernel, "valu_instr_per_launch": round(sum(vals) / len(vals)), "launches": len(vals)}


def batch_spec(n, seed):
    """Mixed sizes: W ~ U[320, 1920], H ~ U[240, 1080] (multiples of 8), D from {64, 128, 192,
    256}, one generator seed per pair."""
    from hbp_amd import synth
    import numpy as np
    k = synth._key(seed, 77)
    r = synth._u01(k, np.arange(3 * n, dtype=np.uint64)).reshape(n, 3)
    out = []
    for i in range(n):
        w = int(round((320 + r[i, 0] * 1600) / 8)) * 8
        h = int(round((240 + r[i, 1] * 840) / 8)) * 8
        d = (64, 128, 192, 256)[min(3, int(r[i, 2] * 4))]
        out.append((seed * 1000 + i, w, h, d))
    return out


def _gen_pair(spec):
    from hbp_amd import synth
    sd, w, h, d = spec
    left, right, _ = synth.make_pair(sd, w, h, d, (6, 12))
    return spec, left, right


def make_batch_pairs(n, seed):
    import multiprocessing as mp
    specs = batch_spec(n, seed)
    workers = max(1, min(16, os.cpu_count() or 1))
    with mp.get_context("fork").Pool(workers) as pool:
        return pool.map(_gen_pair, specs)


def batched_vs_plain(ctx, hip, pairs, min_level=2):
    """The same mixed-size pairs through the plain loop (one pair after the other) and through
    the batched BP (every pair's cost volume and pyramid, then all BP sweeps together); images
    resident in HBM; outputs checked bit for bit."""
    import ctypes as C
    import numpy as np
    from hbp_amd import lib as H
    lib = ctx.lib
    dev = []
    for spec, left, right in pairs:
        nb = left.nbytes
        dl, dr = C.c_void_p(), C.c_void_p()
        assert hip.hipMalloc(C.byref(dl), C.c_size_t(nb)) == 0 and hip.hipMalloc(C.byref(dr), C.c_size_t(nb)) == 0
        assert hip.hipMemcpy(dl, left.ctypes.data_as(C.c_void_p), C.c_size_t(nb), 1) == 0
        assert hip.hipMemcpy(dr, right.ctypes.data_as(C.c_void_p), C.c_size_t(nb), 1) == 0
        dev.append((spec, dl, dr))

    def prepare(spec, dl, dr):
        _, w, h, d = spec
        vol, pyr, bp = C.c_void_p(), C.c_void_p(), C.c_void_p()
        H._check(lib.hbp_cost_volume_dev(ctx.h, dl, dr, w, h, d, b"Census", C.byref(vol)))
        H._check(lib.hbp_pyramid_build_dev(ctx.h, dl, w, h, 2, 4, 0.0, C.byref(pyr)))
        ml = min(min_level, lib.hbp_pyramid_levels(pyr) - 1)
        H._check(lib.hbp_bp_create(ctx.h, pyr, vol, 1000.1, 0.6, ml, 1, C.byref(bp)))
        return vol, pyr, bp

    def release(objs):
        for vol, pyr, bp in objs:
            lib.hbp_bp_free(ctx.h, bp)
            lib.hbp_pyramid_free(ctx.h, pyr)
            lib.hbp_volume_free(ctx.h, vol)

    def download(bp, spec):
        _, w, h, _ = spec
        disp = np.zeros((h, w), np.float32)
        plane = np.zeros((h, w, 3), np.float32)
        H._check(lib.hbp_bp_run(ctx.h, bp, disp.ctypes.data_as(C.c_void_p), plane.ctypes.data_as(C.c_void_p)))
        return disp, plane

    def plain(keep):
        outs = []
        for spec, dl, dr in dev:
            o = prepare(spec, dl, dr)
            H._check(lib.hbp_bp_run(ctx.h, o[2], None, None))
            if keep:
                outs.append(download(o[2], spec))
            release([o])
        return outs

    split = {}

    def batched(keep, workers):
        t0 = time.perf_counter()
        if workers > 1:   # the pairs built by several host threads on streams of their own
            objs = ctx.prepare_pairs_dev([(dl.value, dr.value, spec[1], spec[2], spec[3], min_level)
                                          for spec, dl, dr in dev], workers=workers)
            objs = [(v, p, b) for v, p, b in objs]
        else:
            objs = [prepare(spec, dl, dr) for spec, dl, dr in dev]
        ctx.sync()
        t1 = time.perf_counter()
        hs = [o[2].h if workers > 1 else o[2] for o in objs]
        arr = (C.c_void_p * len(objs))(*hs)
        H._check(lib.hbp_bp_run_batch(ctx.h, len(objs), arr))
        ctx.sync()
        split["prepare_ms"] = round((t1 - t0) * 1e3, 1)
        split["bp_rounds_ms"] = round((time.perf_counter() - t1) * 1e3, 1)
        outs = [download(h, spec) for h, (spec, _, _) in zip(hs, dev)] if keep else []
        if workers > 1:
            for v, p, b in objs:
                b.free()
                p.free()
                v.free()
        else:
            release(objs)
        return outs

    # each form once untimed first (the context's memory pool then holds its buffers)
    plain(False)
    ctx.sync()
    t0 = time.perf_counter()
    plain(False)
    ctx.sync()
    t_plain = time.perf_counter() - t0
    # the pairs prepared one after the other on the context's stream or by PREP_WORKERS host
    # threads; the batched rounds queued from the host or captured in one HIP graph
    graph0 = ctx.get_tuning("batch_graph")
    forms = [(1, 0), (1, 1), (PREP_WORKERS, graph0)]
    t_form, split_form = {}, {}
    for wk, g in forms:
        ctx.set_tuning("batch_graph", g)
        batched(False, wk)
        ctx.sync()
        t0 = time.perf_counter()
        batched(False, wk)
        ctx.sync()
        t_form[(wk, g)] = time.perf_counter() - t0
        split_form[(wk, g)] = dict(split)
    # bit-identical outputs (outside the timed runs)
    ref = plain(True)
    same = {}
    for wk, g in forms:
        ctx.set_tuning("batch_graph", g)
        got = batched(True, wk)
        same[(wk, g)] = all(np.array_equal(a[0].view(np.uint32), b[0].view(np.uint32)) and
                            np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)) for a, b in zip(ref, got))
    ctx.set_tuning("batch_graph", graph0)
    t_batch = t_form[(PREP_WORKERS, graph0)]
    for _, dl, dr in dev:
        hip.hipFree(dl)
        hip.hipFree(dr)
    n = len(dev)
    mpx = sum(w * h * d for _, w, h, d in (p for p, _, _ in dev)) / 1e6
    return {"workload": f"{n} pairs, W U[320,1920] x H U[240,1080] (multiples of 8), D from {{64,128,192,256}}, "
                        f"6-12 planes, min_level 2, reference parameters; images resident in HBM",
            "pairs": n, "plain_pairs_per_s": round(n / t_plain, 3), "batched_pairs_per_s": round(n / t_batch, 3),
            "speedup": round(t_plain / t_batch, 3), "plain_mpix_disp_per_s": round(mpx / t_plain, 1),
            "batched_mpix_disp_per_s": round(mpx / t_batch, 1), "bit_identical": bool(all(same.values())),
            "prepare": f"hbp_pairs_prepare_dev, {PREP_WORKERS} host threads on streams of their own",
            "rounds": "in one HIP graph" if graph0 else "queued from the host",
            "split_ms": split_form[(PREP_WORKERS, graph0)],
            "ab": {"one_stream_prepare_queued_rounds": {"pairs_per_s": round(n / t_form[(1, 0)], 3),
                                                        "split_ms": split_form[(1, 0)],
                                                        "bit_identical": bool(same[(1, 0)])},
                   "one_stream_prepare_graph_rounds": {"pairs_per_s": round(n / t_form[(1, 1)], 3),
                                                       "split_ms": split_form[(1, 1)],
                                                       "bit_identical": bool(same[(1, 1)])},
                   "graph_timed": "capture of every round, instantiation, one replay, synchronisation"},
            "timed": "plain: cost volume + pyramid + BP + map per pair in turn; batched: every pair's cost volume, "
                     "pyramid and BP set-up, then hbp_bp_run_batch over all pairs"}


SERVE_MIX = ((0.4, 640, 480, 64), (0.4, 1280, 720, 128), (0.2, 1920, 1080, 256))


def make_serve_pool(seed, per_class=3):
    """A few generated pairs per class of the serving mix (requests reuse them in turn)."""
    import multiprocessing as mp
    specs = [(seed * 7000 + 100 * c + i, w, h, d) for c, (_, w, h, d) in enumerate(SERVE_MIX) for i in range(per_class)]
    workers = max(1, min(16, os.cpu_count() or 1))
    with mp.get_context("fork").Pool(workers) as pool:
        out = pool.map(_gen_pair, specs)
    return [[p for p in out if p[0][1] == w] for _, w, _, _ in SERVE_MIX]


def served_vs_plain(ctx, pool, rate, seconds, seed, min_level=2):
    """Open-loop Poisson arrivals of the serving mix, through the request server (continuous
    batching) and, as the baseline, through the plain path one pair after the other in arrival
    order.  Host images in, disparity and planes back to host.  Latency = completion time -
    scheduled arrival time."""
    import ctypes as C
    import numpy as np
    from hbp_amd import lib as H
    from hbp_amd import synth
    lib = ctx.lib
    k = synth._key(seed, 91)
    n_max = int(rate * seconds * 3 + 16)
    u = synth._u01(k, np.arange(2 * n_max, dtype=np.uint64)).reshape(n_max, 2)
    arrivals, kinds, t = [], [], 0.0
    counts = [0, 0, 0]
    for i in range(n_max):
        t += -np.log(1.0 - u[i, 0]) / rate
        if t > seconds:
            break
        c = 0 if u[i, 1] < SERVE_MIX[0][0] else (1 if u[i, 1] < SERVE_MIX[0][0] + SERVE_MIX[1][0] else 2)
        arrivals.append(t)
        kinds.append((c, counts[c] % len(pool[c])))
        counts[c] += 1
    n = len(arrivals)

    def pair(i):
        c, j = kinds[i]
        return pool[c][j]

    # served: submit each request at its arrival time, one round per step
    srv = H.Server(ctx, max_active=64, hbm_budget=96 << 30)
    done_at = [None] * n
    tickets = {}
    outputs = {}
    nxt = 0
    t0 = time.perf_counter()
    while nxt < n or tickets:
        now = time.perf_counter() - t0
        while nxt < n and arrivals[nxt] <= now:
            (_, w, h, d), left, right = pair(nxt)
            tickets[srv.submit(left, right, d, min_level)] = nxt
            nxt += 1
        if not tickets:
            time.sleep(max(0.0, min(0.002, arrivals[nxt] - now)))
            continue
        _, _, ndone = srv.step()
        if ndone:
            now = time.perf_counter() - t0
            for tk in list(tickets):
                r = srv.poll(tk)
                if r is not None:
                    i = tickets.pop(tk)
                    done_at[i] = now
                    outputs[i] = _digest(r)
    srv.close()
    lat_s = sorted(d - a for d, a in zip(done_at, arrivals))
    span_s = max(done_at) - arrivals[0]

    # plain baseline: one pair at a time in arrival order
    done_p = []
    ref = {}
    t0 = time.perf_counter()
    for i in range(n):
        now = time.perf_counter() - t0
        if now < arrivals[i]:
            time.sleep(arrivals[i] - now)
        (_, w, h, d), left, right = pair(i)
        vol, pyr, bp = C.c_void_p(), C.c_void_p(), C.c_void_p()
        H._check(lib.hbp_cost_volume(ctx.h, left, right, w, h, d, b"Census", C.byref(vol)))
        H._check(lib.hbp_pyramid_build(ctx.h, left, w, h, 2, 4, 0.0, C.byref(pyr)))
        ml = max(1, min(min_level, lib.hbp_pyramid_levels(pyr) - 1))
        H._check(lib.hbp_bp_create(ctx.h, pyr, vol, 1000.1, 0.6, ml, 1, C.byref(bp)))
        disp = np.zeros((h, w), np.float32)
        plane = np.zeros((h, w, 3), np.float32)
        H._check(lib.hbp_bp_run(ctx.h, bp, disp.ctypes.data_as(C.c_void_p), plane.ctypes.data_as(C.c_void_p)))
        lib.hbp_bp_free(ctx.h, bp)
        lib.hbp_pyramid_free(ctx.h, pyr)
        lib.hbp_volume_free(ctx.h, vol)
        done_p.append(time.perf_counter() - t0)
        ref[i] = _digest((disp, plane))
    lat_p = sorted(d - a for d, a in zip(done_p, arrivals))
    span_p = max(done_p) - arrivals[0]
    differ = sum(1 for i in range(n) if outputs.get(i) != ref[i])

    def pct(v, q):
        return round(1e3 * v[min(len(v) - 1, int(round(q * (len(v) - 1))))], 1)

    return {"workload": f"Poisson arrivals at {rate:g} pairs/s for {seconds:g} s: 40% 640x480/D64, 40% 1280x720/D128, "
                        f"20% 1920x1080/D256, min_level 2, host images in, disparity + planes out",
            "requests": n,
            "served": {"pairs_per_s": round(n / span_s, 3), "p50_ms": pct(lat_s, 0.5), "p99_ms": pct(lat_s, 0.99)},
            "plain": {"pairs_per_s": round(n / span_p, 3), "p50_ms": pct(lat_p, 0.5), "p99_ms": pct(lat_p, 0.99)},
            "results_not_identical": differ,
            "checked": "every request: a 64-bit xxh3 digest of its disparity and planes, served vs plain"}


def _digest(out):
    """64-bit digest of a (disparity, planes) result's bytes"""
    import xxhash
    h = xxhash.xxh3_64()
    for a in out:
        h.update(memoryview(a).cast("B"))
    return h.intdigest()


def _progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def _same(a, b):
    import numpy as np
    return bool(np.array_equal(np.ascontiguousarray(a).view(np.uint32), np.ascontiguousarray(b).view(np.uint32)))


def serve_spec(pool, seed, min_level=2, max_pending=64,
               plan=((10.0, 100.0), (15.0, 70.0), (50.0, 60.0), (100.0, 60.0), (200.0, 60.0), (400.0, 60.0))):
    """The serving mix (40% 640x480/D64, 40% 1280x720/D128, 20% 1920x1080/D256) as open-loop
    Poisson arrivals at each (rate, seconds) of plan, through the request server with a queue
    bound of max_pending (an arrival that finds it full is refused: backpressure).  Per rate:
    offered, accepted, refused, completed, sustained pairs/s (completions inside the window per
    second), p50/p99 of completion - arrival over the completed requests; every result checked
    against the plain path.  Also the plain per-pair loop's capacity on the same mix."""
    import ctypes as C
    import numpy as np
    from hbp_amd import lib as H
    from hbp_amd import synth
    ctx = H.Ctx(0)
    flat = [(c, j) for c in range(len(pool)) for j in range(len(pool[c]))]
    ref = {}
    for c, j in flat:
        (_, w, h, d), left, right = pool[c][j]
        ref[(c, j)] = ctx.bp(ctx.pyramid(left), ctx.cost_volume(left, right, d), min_level=min_level).run()
    # plain loop capacity: 60 requests of the mix back to back, host images in, results out
    k = synth._key(seed, 92)
    u = synth._u01(k, np.arange(120, dtype=np.uint64)).reshape(60, 2)
    seq = []
    for i in range(60):
        c = 0 if u[i, 0] < 0.4 else (1 if u[i, 0] < 0.8 else 2)
        seq.append((c, int(u[i, 1] * len(pool[c]))))
    ctx.sync()
    t0 = time.perf_counter()
    for c, j in seq:
        (_, w, h, d), left, right = pool[c][j]
        ctx.bp(ctx.pyramid(left), ctx.cost_volume(left, right, d), min_level=min_level).run()
    plain_rate = len(seq) / (time.perf_counter() - t0)
    _progress(f"plain loop capacity {plain_rate:.2f} pairs/s")
    runs = []
    for rate, seconds in plan:
        k = synth._key(seed, 93, int(rate))
        n_max = int(rate * seconds * 1.5 + 64)
        u = synth._u01(k, np.arange(3 * n_max, dtype=np.uint64)).reshape(n_max, 3)
        arrivals, kinds, t = [], [], 0.0
        for i in range(n_max):
            t += -np.log(1.0 - u[i, 0]) / rate
            if t > seconds:
                break
            c = 0 if u[i, 1] < 0.4 else (1 if u[i, 1] < 0.8 else 2)
            arrivals.append(t)
            kinds.append((c, int(u[i, 2] * len(pool[c]))))
        n = len(arrivals)
        srv = H.Server(ctx, max_active=64, hbm_budget=96 << 30)
        srv.set_max_pending(max_pending)
        tickets, lat, refused, done_in_window, mismatched = {}, [], 0, 0, 0
        nxt, last_note = 0, 0.0
        t0 = time.perf_counter()
        while nxt < n or tickets:
            now = time.perf_counter() - t0
            while nxt < n and arrivals[nxt] <= now:
                c, j = kinds[nxt]
                (_, w, h, d), left, right = pool[c][j]
                try:
                    tickets[srv.submit(left, right, d, min_level)] = nxt
                except H.HbpError:
                    refused += 1
                nxt += 1
            if not tickets:
                time.sleep(max(0.0, min(0.002, arrivals[nxt] - now)) if nxt < n else 0.0)
                continue
            _, _, ndone = srv.step()
            if not ndone:
                continue
            now = time.perf_counter() - t0
            for tk in list(tickets):
                r = srv.poll(tk)
                if r is None:
                    continue
                i = tickets.pop(tk)
                lat.append(now - arrivals[i])
                done_in_window += now <= seconds
                rd, rp = ref[kinds[i]]
                mismatched += not (_same(r[0], rd) and _same(r[1], rp))
            if now - last_note >= 10.0:
                last_note = now
                _progress(f"rate {rate:g}/s: t={now:.0f} s, {len(lat)} completed, {refused} refused")
        span = time.perf_counter() - t0
        srv.close()
        lat.sort()

        def pct(q):
            return round(1e3 * lat[min(len(lat) - 1, int(round(q * (len(lat) - 1))))], 1) if lat else None

        runs.append({"offered_per_s": rate, "window_s": seconds, "offered": n, "accepted": n - refused,
                     "refused": refused, "completed": len(lat),
                     "sustained_pairs_per_s": round(done_in_window / seconds, 3),
                     "completed_per_s_incl_drain": round(len(lat) / span, 3),
                     "p50_ms": pct(0.5), "p99_ms": pct(0.99), "results_not_identical": mismatched})
        _progress(f"rate {rate:g}/s done: {runs[-1]}")
    ctx.close()
    return {"serve_spec": {
        "workload": "open-loop Poisson arrivals: 40% 640x480/D64, 40% 1280x720/D128, 20% 1920x1080/D256 "
                    f"(4 generated pairs per class), min_level {min_level}, host images in, disparity + planes out; "
                    f"server: 64 slots, 96 GB budget, queue bound {max_pending} (backpressure)",
        "plain_loop_pairs_per_s": round(plain_rate, 3), "runs": runs}}


def mix_levels(pairs, min_level=2, latency_us=25.0):
    """Per-level split of the plain path's time over the mixed batch: for each pair, the cost
    volume, the pyramid and the BP ending at each level from the top down to min_level (best of
    two runs each); a level's time is the difference of the runs ending at it and one level
    above.  A level whose time per sweep is below latency_us is counted as latency-bound (its
    sweeps are chains of small launches that batching amortises); the others as throughput-bound
    (their launches already fill the GPU with one pair)."""
    from hbp_amd import lib as H
    ctx = H.Ctx(0)
    lib = ctx.lib

    def best(fn, reps=2):
        fn()
        ctx.sync()
        b = float("inf")
        for _ in range(reps):
            t = time.perf_counter()
            fn()
            ctx.sync()
            b = min(b, time.perf_counter() - t)
        return b

    rows = []
    t_cost = t_pyr = 0.0
    for spec, left, right in pairs:
        _, w, h, d = spec
        t_cost += best(lambda: ctx.cost_volume(left, right, d).free())
        t_pyr += best(lambda: ctx.pyramid(left).free())
        vol, pyr = ctx.cost_volume(left, right, d), ctx.pyramid(left)
        L = pyr.levels
        prev = best(lambda: ctx.bp(pyr, vol, min_level=L - 1).run())   # no BP level: the map alone
        for l in range(L - 2, min_level - 1, -1):
            t = best(lambda: ctx.bp(pyr, vol, min_level=l).run())
            sweeps = lib.hbp_schedule_sweeps(L, l, d) - lib.hbp_schedule_sweeps(L, l + 1, d) if l + 1 < L - 1 \
                else lib.hbp_schedule_sweeps(L, l, d)
            V, E = pyr.level_size(l)
            rows.append(dict(w=w, h=h, d=d, level=l, nodes=V, edges=E, sweeps=sweeps, ms=1e3 * max(t - prev, 0.0)))
            prev = t
        vol.free()
        pyr.free()
    ctx.close()
    lat = [r for r in rows if r["ms"] * 1e3 / max(r["sweeps"], 1) < latency_us]
    thr = [r for r in rows if r not in lat]
    t_lat = sum(r["ms"] for r in lat)
    t_thr = sum(r["ms"] for r in thr)
    total = 1e3 * (t_cost + t_pyr) + t_lat + t_thr
    by_pixels = {}
    for r in rows:
        k = "superpixels >= 1000" if r["nodes"] >= 1000 else "superpixels < 1000"
        a = by_pixels.setdefault(k, dict(ms=0.0, sweeps=0))
        a["ms"] += r["ms"]
        a["sweeps"] += r["sweeps"]
    return {"mix_levels": {
        "workload": f"{len(pairs)} pairs of the mixed batch, plain path, min_level {min_level}",
        "cost_volume_ms": round(1e3 * t_cost, 1), "pyramid_ms": round(1e3 * t_pyr, 1),
        "bp_latency_bound_ms": round(t_lat, 1), "bp_throughput_bound_ms": round(t_thr, 1),
        "latency_rule_us_per_sweep": latency_us,
        "levels_latency_bound": len(lat), "levels_throughput_bound": len(thr),
        "sweeps_latency_bound": sum(r["sweeps"] for r in lat), "sweeps_throughput_bound": sum(r["sweeps"] for r in thr),
        "us_per_sweep_latency_bound": round(1e3 * t_lat / max(sum(r["sweeps"] for r in lat), 1), 2),
        "by_level_size": {k: dict(ms=round(v["ms"], 1), sweeps=v["sweeps"],
                                  us_per_sweep=round(1e3 * v["ms"] / max(v["sweeps"], 1), 2))
                          for k, v in by_pixels.items()},
        "total_ms": round(total, 1),
        "ceiling_if_latency_bound_work_were_free": round(total / max(total - t_lat, 1e-9), 3),
        "ceiling_if_it_and_the_pyramid_were_free": round(total / max(total - t_lat - 1e3 * t_pyr, 1e-9), 3)}}


def batch_kernels(pairs, min_level=2):
    """The batched BP rounds of the mixed batch with HIP events around every launch: per kernel
    the time, the launches and the algorithmic bytes (per launch the sum of the one-pair
    launches' figures: 12 B per pixel and 24 B per plane and superpixel for the plane costs, 96 B
    per edge and 16 K + 48 B per superpixel for the messages, 84 B per edge and 48 K + 16 B per
    superpixel for the node update), hence GB/s against the HBM peak.  The rounds' wall time
    with and without the events is reported beside them."""
    from hbp_amd import lib as H
    ctx = H.Ctx(0)

    def prepare():
        keep, bps = [], []
        for (sd, w, h, d), left, right in pairs:
            v, p = ctx.cost_volume(left, right, d), ctx.pyramid(left)
            keep += [v, p]
            bps.append(ctx.bp(p, v, min_level=min(min_level, p.levels - 1)))
        ctx.sync()
        return keep, bps

    def rounds(profiled):
        keep, bps = prepare()
        if profiled:
            ctx.profile(True)
            ctx.profile_reset()
            ctx.profile_only(None, 1)
        t0 = time.perf_counter()
        ctx.bp_run_batch(bps)
        ctx.sync()
        el = time.perf_counter() - t0
        st = ctx.profile_stats() if profiled else None
        ctx.profile(False)
        sweeps = sum(b.total_sweeps for b in bps)
        for b in bps:
            b.free()
        for o in keep:
            o.free()
        return el, st, sweeps

    rounds(False)
    t_plain_ev, _, _ = rounds(False)
    t_ev, st, sweeps = rounds(True)
    kern = {}
    for name, v in sorted(st.items(), key=lambda kv: -kv[1]["ms"]):
        if not name.startswith("batch_") or v["launches"] == 0:
            continue
        gbs = v["bytes"] / (v["ms"] * 1e-3) / 1e9 if v["ms"] > 0 else 0.0
        kern[name] = {"ms": round(v["ms"], 1), "launches": v["launches"],
                      "us_per_launch": round(1e3 * v["ms"] / v["launches"], 2),
                      "mb_per_launch": round(v["bytes"] / v["launches"] / 1e6, 2),
                      "gbs": round(gbs, 1), "frac_of_hbm_peak": round(gbs / HBM_PEAK_GBS, 4)}
    return {"batch_kernels": {
        "workload": f"{len(pairs)} pairs of the mixed batch, min_level {min_level}, hbp_bp_run_batch (rounds queued)",
        "sweeps": sweeps, "rounds_s": round(t_plain_ev, 3), "rounds_s_with_events": round(t_ev, 3),
        "kernels": kern,
        "bound": "the messages (batch_edge and the coarse launches' message waves) are VALU-bound; the HBM "
                 "fractions are of the algorithmic bytes, not of measured traffic"}}


def stress_spec(seed, n4k=8, budgets=(16 << 30, 256 << 30)):
    """Memory pressure: the 64-pair mixed batch with a 3840x2160 D384 pair (BP on 5 levels)
    after every 8 of them, all submitted at once, served under each HBM budget; pairs/s, the
    accounted device bytes at most, how long the 4K pairs waited for admission; every result
    checked against the plain path; the 4K pair's BP message bytes per level."""
    import numpy as np
    from hbp_amd import lib as H
    from hbp_amd import synth
    specs = synth.mixed_batch(seed, 64)
    mix = synth.make_pairs_parallel(specs)
    big = [synth.make_pair_parallel(seed * 100 + i, 3840, 2160, 384, (20, 40))[:2] for i in range(2)]
    ctx = H.Ctx(0)
    jobs = []
    for i in range(64):
        jobs.append((mix[i][0], mix[i][1], specs[i][3], 2, "mix"))
        if i % 8 == 7:
            b = big[(i // 8) % 2]
            jobs.append((b[0], b[1], 384, -5, "4k"))
    # plain references (and the 4K pair's levels)
    ref = []
    for l, r, d, ml, kind in jobs:
        pyr = ctx.pyramid(l)
        L = pyr.levels
        m = max(1, min(L - 1 + ml if ml < 0 else ml, L - 1))
        ref.append(ctx.bp(pyr, ctx.cost_volume(l, r, d), min_level=m).run())
    pyr4 = ctx.pyramid(big[0][0])
    L4 = pyr4.levels
    ml4 = L4 - 6
    levels = []
    max_e = max_slots = 0
    for lv in range(ml4, L4 - 1):
        info = pyr4.level(lv)
        deg = np.bincount(info["bd"][:, 0], minlength=info["V"]).max() if info["E"] else 1
        max_e = max(max_e, info["E"])
        max_slots = max(max_slots, info["V"] * int(deg))
        levels.append({"level": lv, "nodes": info["V"], "edges": info["E"], "max_degree": int(deg),
                       "msg_bytes": 32 * info["E"], "msg_t_bytes": 48 * info["V"] * int(deg)})
    E0 = pyr4.level(0)["E"]
    runs = []
    for budget in budgets:
        srv = H.Server(ctx, max_active=128, hbm_budget=budget)
        t0 = time.perf_counter()
        tickets = {srv.submit(l, r, d, ml): i for i, (l, r, d, ml, _) in enumerate(jobs)}
        first_active, done_at, out_ok = {}, {}, 0
        while tickets:
            a, p, nd = srv.step()
            if not nd:
                continue
            now = time.perf_counter() - t0
            for tk in list(tickets):
                r = srv.poll(tk)
                if r is not None:
                    i = tickets.pop(tk)
                    done_at[i] = now
                    out_ok += _same(r[0], ref[i][0]) and _same(r[1], ref[i][1])
        span = time.perf_counter() - t0
        _, peak, _ = srv.stats()
        srv.close()
        big_done = sorted(done_at[i] for i, j in enumerate(jobs) if j[4] == "4k")
        runs.append({"budget_gb": budget / (1 << 30), "pairs": len(jobs), "seconds": round(span, 3),
                     "pairs_per_s": round(len(jobs) / span, 3), "peak_accounted_gb": round(peak / (1 << 30), 2),
                     "results_identical": out_ok, "first_4k_done_s": round(big_done[0], 3),
                     "last_4k_done_s": round(big_done[-1], 3)})
        _progress(f"budget {budget >> 30} GB: {runs[-1]}")
    ctx.close()
    return {"stress_spec": {
        "workload": "64 pairs of the mixed batch (W U[320,1920], H U[240,1080], D 64-256, min_level 2) with a "
                    "3840x2160 D384 pair (20-40 planes, BP on 5 levels) after every 8, all submitted at once",
        "runs": runs,
        "bp_messages_4k": {"levels": levels, "allocated_bytes": 32 * max_e + 48 * (max_slots + 16),
                           "reference_bytes": 2 * 4 * 12 * E0,
                           "note": "messages [E][8] f32 and node-major sweep messages [V][maxdeg][12] f32 per BP "
                                   "level, allocated at the largest BP level; the reference sizes msg and msg_t "
                                   "by level 0 (12 E_0 floats each, hbp_isp.hpp:367-368)"}}}


def post_process(ctx, left, right, D, min_level):
    """The refinements after BP on the same pair (the sequence of main_stereo_hbp_isp.cu:84-113):
    the right view's BP, the left-right check + fill, the weighted-median refinement of the
    failed pixels' planes, plane_label_to_disp, and the CLI's 7-tap median of the left
    disparity.  Device time of each stage (HIP events), host buffers in between."""
    Hh, W = left.shape[:2]
    dl, pl = ctx.bp(ctx.pyramid(left), ctx.cost_volume(left, right, D), min_level=min_level).run()
    dr, _ = ctx.bp(ctx.pyramid(right), ctx.cost_volume(left, right, D, right_view=True), min_level=min_level).run()
    runs = []
    for _ in range(3):
        ctx.profile(True)
        ctx.profile_reset()
        filled, fl, _, _ = ctx.lr_check_fill(dl, dr, pl)
        wm = ctx.weighted_median(left, fl, filled)
        ctx.plane_to_disp(wm)
        ctx.median7(dl)
        ctx.sync()
        runs.append(ctx.profile_stats())
        ctx.profile(False)
    st = runs[-1]

    def ms(name):
        return round(min(r[name]["ms"] for r in runs), 3) if name in st else None

    nfail = int((fl != 0).sum())
    out = {"workload": f"{W}x{Hh} D{D}: right-view BP, LR check + fill, weighted median (41x41, failed pixels), "
                       f"7-tap median",
           "failed_pixels": nfail, "lr_check_ms": ms("lr_check"), "weighted_median_ms": ms("weighted_median"),
           "median7_ms": ms("median7")}
    wmm = out["weighted_median_ms"]
    if wmm:
        out["weighted_median_mpix_window_per_s"] = round(nfail * 41 * 41 / (wmm * 1e-3) / 1e6, 1)
    return out, (left, fl, filled, wm)


def cpu_weighted_median(left, fl, filled, wm_gpu):
    """The oracle's weighted median on the same failed pixels (host cores), and the GPU result
    checked against it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_ffi as O
    Hh, W = fl.shape
    ref = np.ascontiguousarray(filled, np.float32).copy()
    t0 = time.perf_counter()
    O.lib.orc_weighted_median(np.ascontiguousarray(left), np.ascontiguousarray(fl), ref, W, Hh)
    t = time.perf_counter() - t0
    same = bool(np.array_equal(np.isnan(ref), np.isnan(wm_gpu)) and
                np.array_equal(ref.view(np.uint32)[~np.isnan(ref)], wm_gpu.view(np.uint32)[~np.isnan(wm_gpu)]))
    return {"cpu_weighted_median_ms": round(1e3 * t, 1), "cpu_threads": int(os.environ.get("OMP_NUM_THREADS", "1")),
            "weighted_median_bit_identical": same}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(left, right, wl, info, sweeps, gpu_disp, gpu_plane):
    """The oracle (host C++17, -O3, OpenMP over pixels, clusters and edges) on the GPU host's
    cores.  All cores: the whole pair (cost volume, pyramid, every BP sweep, map), its disparity
    and planes compared with the GPU's.  One thread: cost volume, pyramid and a bounded number of
    sweeps, scaled to the whole pair by the sweep count."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_ffi as O

    D = wl["ndisp"]
    ml, total = info["min_level"], info["sweeps"]
    cores = os.cpu_count() or 1

    def prefix(threads):
        O.lib.orc_set_threads(threads)
        t0 = time.perf_counter()
        vol, wta = O.cost_volume(left, right, D)
        pyr = O.Pyramid(left)
        return time.perf_counter() - t0, O.BP(pyr, vol, wta, D, min_level=ml)

    # all cores, the whole pair
    t_pre, bp = prefix(cores)
    t0 = time.perf_counter()
    ran = bp.step()
    disp, plane = bp.map()
    t_all = t_pre + time.perf_counter() - t0
    same_d = disp.view(np.uint32) == gpu_disp.view(np.uint32)
    same_p = (plane.view(np.uint32) == gpu_plane.view(np.uint32)).all(axis=2)
    differ = float(1.0 - (same_d & same_p).mean())
    del bp
    # one thread, a sample: the prefix and `n1` sweeps
    t_pre1, bp1 = prefix(1)
    n1 = sweeps if sweeps > 0 else 0
    tc = time.perf_counter()
    done = bp1.step(4)
    per = (time.perf_counter() - tc) / max(done, 1)
    if n1 <= 0:
        n1 = int(max(8, min(total, 12.0 / max(per, 1e-6))))
    rest = max(0, n1 - done)
    tc = time.perf_counter()
    done2 = bp1.step(rest) if rest else 0
    per1 = (time.perf_counter() - tc) / done2 if done2 else per
    t_one = t_pre1 + per1 * total
    O.lib.orc_set_threads(cores)
    return {"value": round(1.0 / t_all, 5), "unit": "pairs/s", "cores": cores, "kind": "port",
            "cpu": cpu_model(), "nproc": cores,
            "sample": f"oracle (host C++17 -O3 -march=x86-64-v3, OpenMP) on {cores} threads, the whole pair: cost "
                      f"volume + pyramid {t_pre:.2f} s, all {ran} BP sweeps and the map, {t_all:.2f} s in all",
            "matches_gpu": {"bit_identical": differ == 0.0, "pixels_differing_fraction": differ},
            "one_thread": {"value": round(1.0 / t_one, 6), "unit": "pairs/s", "cores": 1,
                           "sample": f"cost volume + pyramid {t_pre1:.2f} s and {done + done2} of {total} BP sweeps "
                                     f"({per1 * 1e3:.1f} ms/sweep), scaled to all sweeps"}}


if __name__ == "__main__":
    main()
