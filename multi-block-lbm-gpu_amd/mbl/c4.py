"""Batched sweep of independent two-level cylinder cases.

    python -m mbl.c4 [--cases 256] [--seed 2024] [--steps 2000] [--storage f32|f16s] [--no-graph] [--no-tune]
                     [--persistent] [--vx N --tpb N]

256 cases from one seed: Re ~ U[40, 200], D in {10, 16, 24, 32, 40} finest
cells, blockage D/H ~ U[0.1, 0.3], channel length 20 D, refined block
4 D x 2 D. All cases advance through shared launches (mbl_batch_*), or with
--persistent through one launch per timed run, one workgroup per case; the
autotuner picks the launch shape first. Prints one JSON line with the
aggregate MLUPS (node updates of all levels and cases per wall second).
"""
import argparse
import json
import time

import numpy as np

from . import Batch, abi, c4_specs, load_library


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", type=int, default=256)
    ap.add_argument("--seed", type=int, default=2024)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16s"])
    ap.add_argument("--no-graph", action="store_true", help="launch the cycle's kernels one by one")
    ap.add_argument("--persistent", action="store_true",
                    help="one launch per step() call, one workgroup per case")
    ap.add_argument("--no-tune", action="store_true")
    ap.add_argument("--vx", type=int, default=0, help="nodes per thread without tuning (0: library default)")
    ap.add_argument("--tpb", type=int, default=0, help="threads per workgroup without tuning (0: library default)")
    args = ap.parse_args(argv)
    lib = load_library(False)
    storage = abi.STORE_NATIVE if args.storage == "f32" else abi.STORE_F16S
    t0 = time.perf_counter()
    b = Batch(lib, c4_specs(args.cases, seed=args.seed), storage=storage, use_graph=0 if args.no_graph else 1,
              persistent=args.persistent)
    out = {"cases": args.cases, "storage": args.storage, "graph": not args.no_graph, "persistent": args.persistent,
           "create_s": round(time.perf_counter() - t0, 3), "node_updates_per_step": b.node_updates()}
    if args.vx or args.tpb:
        from . import check
        check(lib, lib.mbl_set_tuning(b.h, args.vx, args.tpb), "tuning")
        out["shape"] = {"vx": args.vx, "tpb": args.tpb}
    elif not args.no_tune:
        out["tune"] = b.autotune(trial_steps=5)
    b.step(args.warmup)
    b.sync()
    t0 = time.perf_counter()
    b.step(args.steps)
    b.sync()
    wall = time.perf_counter() - t0
    cd, _ = b.forces(1)
    out.update({"steps": args.steps, "wall_s": round(wall, 4), "ms_per_step": round(wall / args.steps * 1e3, 4),
                "mlups": round(b.node_updates() * args.steps / wall / 1e6, 1),
                "finite_cases": int(np.isfinite(cd[:, -1]).sum())})
    b.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
