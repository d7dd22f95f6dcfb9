#!/usr/bin/env python3
"""Bandwidth benchmark of the fused D2Q9 stream-collide path on one MI355X.

Workload: one periodic 16384 x 16384 block, rho = 1, (u, v) ~ N(0, 0.01^2)
from a fixed seed, tau = 0.6; one step = one fused pull-stream + BGK update of
all nodes, F planes in HBM the whole time. The library takes two steps per
launch on this block with FP32 planes (k_periodic3: step n+1 of a tile and a
two-node ring, then step n+2 of the tile and a one-node ring, in LDS, then
step n+3; the same bits as three single steps), so K steps are K/3 launches
(16-bit planes: two steps per launch, k_periodic2). MLUPS = nodes * steps /
wall seconds. The roofline figure uses the algorithmic bytes of one launch,
72 B per node for FP32 planes (9 loads of step n + 9 stores of step n+3, 4 B
each; three node updates), against the 8 TB/s HBM peak, with the per-launch
time taken from HIP events on the library's stream. The single-step kernel
(k_stream_collide, 72 B per node update) is timed beside it (`single_step`).

FP32 planes are updated in FP64 arithmetic (same per-launch time as FP32
arithmetic on this HBM-bound kernel; `dtype` names the arithmetic, `storage`
the planes).

The metric's second half, the time to a drag coefficient converged within
+-0.5 %, is measured on the two-level (Re = 20) and three-level (Re = 100)
cylinder channels for plain FP32 planes against 16-bit 1-3-12 planes
(`time_to_converged_cd`; about 40 s, --no-converge skips it).

The CPU baseline times the FP64 oracle (three passes per step, OpenMP) on a
bounded sample of the same workload (a smaller periodic block, a few steps).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(REPO, "multi-block-lbm-gpu_amd"))

METRIC = "MLUPS & %HBM roofline; time-to-converged Cd (±0.5%) vs plain FP32"
HBM_PEAK_GBS = 8000.0
BYTES_PER_NODE = {0: 72, 1: 72, 2: 36, 3: 36, 4: 36, 5: 36}
STORAGE_NAMES = {0: "f32", 1: "f32s", 2: "f16s", 3: "bf16s", 4: "f16c", 5: "f16d"}


def traffic_from_profile(storage, kernel="k_stream_collide"):
    """HBM bytes per launch of a kernel measured with rocprofv3 PMC counters, if recorded."""
    path = os.path.join(REPO, "profiles", f"traffic_periodic16384_{STORAGE_NAMES[storage]}.json")
    if not os.path.exists(path):
        return None
    with open(path) as fh:
        d = json.load(fh)
    if kernel == "k_stream_collide":
        return d.get("bytes_per_launch")
    return d.get("kernels", {}).get(kernel, {}).get("bytes_per_launch")


LAUNCH_MODES = {3: 3, 2: 1, 1: 0}  # --launch-steps -> mbl_set_step_pairs mode


def gpu_run(lib, n, steps, warmup, storage, tuning=None, launch_steps=3):
    """launch_steps 3: the library's default (three steps per launch for 4-byte planes, two for
    16-bit planes); 2: two steps per launch; 1: one step per launch."""
    import mbl

    blk = mbl.PeriodicBlock(lib, n, n, tau=0.6, storage=storage, tuning=tuning)
    mbl.check(lib, lib.mbl_set_step_pairs(blk.h, LAUNCH_MODES[launch_steps]), "steps per launch")
    m0 = blk.moments()  # mass, momentum and non-finite count of the initial field (device sums)
    # bring the GPU clocks up first: untimed steps for at least half a second, then the W warmup steps
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < 0.5:
        blk.step(10)
        blk.sync()
    blk.step(warmup)
    blk.sync()
    t0 = time.perf_counter()
    ev_ms = blk.timed_steps(steps)  # events bracket the launches of K steps; ends with a synchronize
    wall = time.perf_counter() - t0
    launches = blk.launches(steps)
    m1 = blk.moments()
    blk.close()
    nodes = n * n
    # a periodic BGK block conserves mass and momentum: their change over the run (relative to
    # the mass) is rounding, and every entry stays finite
    inv = {"mass_rel_change": float((m1[0] - m0[0]) / m0[0]),
           "momentum_change_per_node": [float((m1[1] - m0[1]) / nodes), float((m1[2] - m0[2]) / nodes)],
           "nonfinite_entries": int(m1[3])}
    return {"wall_s": wall, "event_ms": ev_ms, "launches": launches, "per_launch_ms": ev_ms / launches,
            "steps_per_launch": steps / launches, "mlups": nodes * steps / wall / 1e6, "nodes": nodes,
            "invariants": inv}


def stream_copy_gbs(lib, mib=4096, reps=10):
    """Streaming-copy rates (bytes read + written per second) of the library's copy kernels over
    MIB MiB: one stream (the HBM ceiling the fused kernel is compared with besides the 8 TB/s
    datasheet peak) and 9 planes read + 9 written (the fused kernel's own access pattern)."""
    import ctypes

    import mbl

    cfg = mbl.default_config(lib, LEVEL=1, write_files=0)
    cfg.mi[0] = 1
    h = ctypes.c_void_p()
    if lib.mbl_handle_create(ctypes.byref(h), ctypes.byref(cfg)):
        return None, None
    g, g9 = ctypes.c_double(), ctypes.c_double()
    rc = lib.mbl_bandwidth_probe(h, mib << 20, reps, ctypes.byref(g))
    rc9 = lib.mbl_bandwidth_probe_planes(h, (mib << 20) // 9, reps, ctypes.byref(g9))
    lib.mbl_handle_destroy(h)
    return (g.value if rc == 0 else None), (g9.value if rc9 == 0 else None)


def copy_calibration(mib):
    """Five device-to-device copies of a known size: what FETCH_SIZE / WRITE_SIZE
    read for them calibrates the counters against bytes actually moved."""
    import ctypes
    import mbl

    lib = mbl.load_library(False)
    n = mib << 20
    a, b = lib.mbl_malloc(n), lib.mbl_malloc(n)
    lib.mbl_memset_d(a, 1, n)
    for _ in range(5):
        lib.mbl_memcpy_d2d(b, a, n)
    lib.mbl_free(a)
    lib.mbl_free(b)
    print(json.dumps({"copy_calibration_bytes": n, "copies": 5}))


def cpu_baseline(target_s):
    """FP64 oracle on a 4096 x 4096 periodic sample, steps sized to ~target_s seconds."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle

    n = 4096
    o = oracle.OraclePeriodic(n, n, tau=0.6)
    t1 = o.run(1)
    steps = max(1, min(50, int(target_s / max(t1, 1e-3))))
    secs = o.run(steps)
    threads = o.lib.orc_threads()
    o.close()
    return {"value": round(n * n * steps / secs / 1e6, 3), "unit": "MLUPS", "cores": threads, "kind": "port",
            "sample": f"periodic {n}x{n}, tau=0.6, FP64 oracle (stb1/stb2/stb3 passes, OpenMP), {steps} steps, "
                      f"{secs:.1f} s"}


def converged_cd():
    """The metric's second half: coarse steps and wall seconds until the drag coefficient stays
    within +-0.5 % of the FP64 run's converged value, plain FP32 planes against 16-bit 1-3-12
    planes, on the two-level channel at Re = 20 (steady Cd, 20,000 coarse steps) and the
    three-level channel at Re = 100 (Cd averaged over one shedding period, 50,000 coarse
    steps), with the per-result error of each against FP64."""
    from mbl import precision

    print("# bench: time to converged Cd (two-level Re=20, three-level Re=100)", file=sys.stderr, flush=True)
    res = {}
    for case in ("c1", "c2b"):
        r = precision.case_c1(["f32", "f16d"]) if case == "c1" else precision.case_c2b(["f32", "f16d"])
        rows = {x["storage"]: x for x in r["rows"]}
        f32 = rows["f32"]
        entry = {"node_updates_per_step": r["node_updates_per_step"], "steps_run": r["steps"]}
        for st, x in rows.items():
            e = {"steps": x["t_conv_steps"], "wall_s": None if x["t_conv_s"] is None else round(x["t_conv_s"], 4),
                 "mlups": round(x["mlups"], 1)}
            if st != "f64":
                errs = {k: round(v, 8) for k, v in x.items() if k.endswith("_err")}
                e["err_vs_f64"] = errs
                if x["t_conv_s"] is not None and f32["t_conv_s"]:
                    e["wall_vs_f32"] = round(x["t_conv_s"] / f32["t_conv_s"], 4)
            entry[st] = e
        res["two_level_re20" if case == "c1" else "three_level_re100"] = entry
    ws = os.path.join(REPO, "profiles", "warmstart_c5.json")
    if os.path.exists(ws):
        with open(ws) as fh:
            res["learned_warm_start_recorded"] = json.load(fh)
    return res


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=1200)  # a multiple of 3 and 2: whole launches
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--n", type=int, default=16384)
    ap.add_argument("--storage", type=int, default=0, help="0 f32, 1 f32 shifted, 2 f16 shifted, 3 bf16 shifted, 4 1-4-11 half shifted, 5 1-3-12 half shifted")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-converge", action="store_true", help="skip the time-to-converged-Cd runs (about 40 s)")
    ap.add_argument("--variants", action="store_true", help="also time the 16-bit storage modes")
    ap.add_argument("--launch-steps", type=int, default=3, choices=[1, 2, 3],
                    help="steps per launch on the periodic block: 3 = the library default (three for 4-byte "
                         "planes, two for 16-bit planes), 2 = two, 1 = one")
    ap.add_argument("--vx", type=int, default=0, help="nodes per thread (0: library default)")
    ap.add_argument("--tpb", type=int, default=0, help="threads per workgroup (0: library default)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run 5 device-to-device copies of MIB MiB (HBM counter calibration) and exit")
    args = ap.parse_args()
    if args.copy_calibration:
        copy_calibration(args.copy_calibration)
        return
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is not supported "
              f"(the multi-GPU sharding is a design, not a run)", file=sys.stderr)
        sys.exit(2)
    import mbl

    lib = mbl.load_library(False)
    print(f"# bench: periodic {args.n}^2 storage={STORAGE_NAMES[args.storage]} steps={args.steps}", file=sys.stderr,
          flush=True)
    tuning = (args.vx, args.tpb) if (args.vx or args.tpb) else None
    r = gpu_run(lib, args.n, args.steps, args.warmup, args.storage, tuning, args.launch_steps)
    bpn = BYTES_PER_NODE[args.storage]  # algorithmic bytes per node of one launch (either kernel)
    achieved = bpn * r["nodes"] / (r["per_launch_ms"] * 1e-3) / 1e9
    spl = r["steps_per_launch"]
    kernel = ("k_periodic3" if spl > 2.5 else "k_periodic2") if spl > 1.5 else "k_stream_collide"
    kname = {"k_periodic3": "k_periodic3 (three steps per launch)", "k_periodic2": "k_periodic2 (two steps per launch)",
             "k_stream_collide": "k_stream_collide (periodic)"}[kernel]
    single = gpu_run(lib, args.n, max(100, args.steps // 5), args.warmup, args.storage, tuning, launch_steps=1)
    single_achieved = bpn * single["nodes"] / (single["per_launch_ms"] * 1e-3) / 1e9
    out = {
        "metric": METRIC,
        "value": round(r["mlups"], 1),
        "unit": "MLUPS",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(r["wall_s"] / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64" if args.storage in (0, 1) else "f32",
        "data": "synthetic (seeded random velocity field)",
        "config": {"workload": f"periodic {args.n}x{args.n} D2Q9 BGK block, tau=0.6, "
                               f"{STORAGE_NAMES[args.storage]} planes, "
                               + (f"{round(spl)} steps per launch, {bpn} B/node/launch" if spl > 1.5
                                  else f"{bpn} B/node/step"),
                   "nodes": r["nodes"], "storage": STORAGE_NAMES[args.storage],
                   "arithmetic": "f64" if args.storage in (0, 1) else "f32",
                   "steps_per_launch": round(r["steps_per_launch"], 3)},
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": traffic_from_profile(args.storage, kernel),
                     "kernel": kname,
                     "per_launch_ms": round(r["per_launch_ms"], 5), "bytes_per_node": bpn,
                     "node_updates_per_launch": round(r["steps_per_launch"] * r["nodes"])},
        "single_step": {"kernel": "k_stream_collide (periodic)", "mlups": round(single["mlups"], 1),
                        "per_launch_ms": round(single["per_launch_ms"], 5),
                        "achieved": round(single_achieved, 1), "frac": round(single_achieved / HBM_PEAK_GBS, 4),
                        "traffic": traffic_from_profile(args.storage, "k_stream_collide")},
        "invariants": r["invariants"],
    }
    copy, copy9 = stream_copy_gbs(lib)
    if copy:
        out["roofline"]["copy_gbs"] = round(copy, 1)
        out["roofline"]["frac_of_copy"] = round(achieved / copy, 4)
    if copy9:
        out["roofline"]["copy_9planes_gbs"] = round(copy9, 1)
        out["roofline"]["frac_of_copy_9planes"] = round(achieved / copy9, 4)
    if args.variants:
        var = {}
        for st in (1, 2, 3, 4, 5):
            v = gpu_run(lib, args.n, max(200, args.steps // 4), args.warmup, st)
            b = BYTES_PER_NODE[st]
            var[STORAGE_NAMES[st]] = {"mlups": round(v["mlups"], 1), "per_launch_ms": round(v["per_launch_ms"], 5),
                                      "achieved_gbs": round(b * v["nodes"] / (v["per_launch_ms"] * 1e-3) / 1e9, 1),
                                      "bytes_per_node": b}
        out["storage_variants"] = var
    if not args.no_converge:
        out["time_to_converged_cd"] = converged_cd()
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
