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
EVENT_EVERY = 8   # timed region: HIP events around one in 8 launches of the dominant kernel
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
    ap.add_argument("--cpu-sweeps", type=int, default=0,
                    help="sweeps the CPU baseline runs (0: scale to about 20 s)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: the stereo path runs one pair per GPU and this build has "
              "no multi-GPU runner; use --gpus 1", file=sys.stderr)
        sys.exit(2)

    import ctypes as C
    import numpy as np
    from hbp_amd import lib as H
    from hbp_amd import synth

    wl = WORKLOADS[args.workload]
    W, Hh, D = wl["width"], wl["height"], wl["ndisp"]
    left, right, _ = synth.make_pair(args.seed, W, Hh, D, wl["n_planes"])

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

    for _ in range(args.warmup):
        one_pair()
    ctx.sync()
    # timed region: HIP events around the dominant kernel's launches only
    dom = "plane_cost"
    ctx.profile(True)
    ctx.profile_only(dom, EVENT_EVERY)
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

    pairs_per_s = args.steps / elapsed
    mpix_disp = W * Hh * D / 1e6 * pairs_per_s
    lat_ms = sorted(1e3 * x for x in lat)
    p50 = statistics.median(lat_ms)
    p99 = lat_ms[min(len(lat_ms) - 1, int(round(0.99 * (len(lat_ms) - 1))))]

    def group(st):
        out = {}
        for name, v in st.items():
            g = name.split("_L")[0]
            a = out.setdefault(g, dict(ms=0.0, bytes=0.0, launches=0))
            a["ms"] += v["ms"]
            a["bytes"] += v["bytes"]
            a["launches"] += v["launches"]
        return out

    gd = group(stats).get(dom, dict(ms=0.0, bytes=0.0, launches=0))
    avg_ms = gd["ms"] / max(gd["launches"], 1)
    avg_bytes = gd["bytes"] / max(gd["launches"], 1)
    achieved = avg_bytes / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0
    kernels = {g: dict(ms_per_pair=round(v["ms"], 3), launches=v["launches"],
                       gbs=round(v["bytes"] / (v["ms"] * 1e-3) / 1e9, 1) if v["ms"] > 0 else None)
               for g, v in sorted(group(pair_stats).items(), key=lambda kv: -kv[1]["ms"])}

    # HBM bytes per launch of the dominant kernel from rocprofv3 FETCH_SIZE / WRITE_SIZE runs
    traffic = None
    tfile = os.path.join(ROOT, "profiles", f"traffic_{dom}_{args.workload}.json")
    if os.path.exists(tfile):
        try:
            traffic = json.load(open(tfile)).get("bytes_per_launch")
        except (OSError, ValueError):
            traffic = None

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(left, right, wl, info, args.cpu_sweeps)

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
                     "event_sampling": f"1 in {EVENT_EVERY} launches in the timed region"},
        "kernels_one_pair": kernels,
        "launch_us_by_level": {n: round(1e3 * v["ms"] / max(v["launches"], 1), 2)
                               for n, v in sorted(pair_stats.items()) if "_L" in n},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)
    hip.hipFree(d_l)
    hip.hipFree(d_r)
    ctx.close()


def cpu_baseline(left, right, wl, info, sweeps):
    """The oracle on the box's host cores: cost volume + pyramid + the first `sweeps` BP sweeps
    of the same pair, scaled to a whole pair by the sweep count."""
    import ctypes as C
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import numpy as np
    import oracle_ffi as O

    cores = min(16, os.cpu_count() or 1)
    os.environ.setdefault("OMP_NUM_THREADS", str(cores))
    D = wl["ndisp"]
    t0 = time.perf_counter()
    vol, wta = O.cost_volume(left, right, D)
    pyr = O.Pyramid(left)
    t1 = time.perf_counter()
    ml = info["min_level"]
    bp = O.BP(pyr, vol, wta, D, min_level=ml)
    total = info["sweeps"]
    if sweeps <= 0:
        # calibrate on 8 sweeps, then run about 20 s worth
        tc = time.perf_counter()
        done = bp.step(8)
        per = (time.perf_counter() - tc) / max(done, 1)
        sweeps = int(max(8, min(total, 20.0 / max(per, 1e-6))))
        rest = sweeps - done
    else:
        done, rest = 0, sweeps
    t2 = time.perf_counter()
    done2 = bp.step(rest) if rest > 0 else 0
    t3 = time.perf_counter()
    ran = done + done2
    per_sweep = (t3 - t2) / max(done2, 1) if done2 else (t3 - t2)
    pair_s = (t1 - t0) + per_sweep * total
    return {"value": round(1.0 / pair_s, 5), "unit": "pairs/s", "cores": int(os.environ["OMP_NUM_THREADS"]),
            "kind": "port",
            "sample": f"oracle (host C++, OpenMP) on the same pair: cost volume + pyramid ({t1 - t0:.2f} s) and "
                      f"{ran} of {total} BP sweeps ({per_sweep * 1e3:.1f} ms/sweep at the top levels), "
                      f"scaled to all sweeps"}


if __name__ == "__main__":
    main()
