"""Cylinder cases of the multi-block solver, run through the C ABI.

    python -m mbl.cases c1  [--steps 20000] [--storage f32|f32s|f16s|f16c|f16d|bf16s] [--fp64]
    python -m mbl.cases c2b [--steps 50000] ...

c1:  two-level channel of 300 x 80 coarse cells, one ratio-2 block of 80 x 40
     coarse cells around a cylinder of D = 10 coarse at (60, 40); parabolic
     inlet Umax = 0.1, pressure outlet, bounce-back walls, Re = 20, rest start.
c2b: three-level channel of 2048 x 512 coarse cells, two nested ratio-2
     blocks, cylinder D = 16 coarse = 64 finest cells at (512, 256), Re = 100.

Prints one JSON line: wall time, MLUPS (node updates of all levels per wall
second), final residual, Cd/Cl statistics and the time to a converged Cd.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import MultiBlockSim, c1_geometry, c2b_geometry, check, load_library, postproc
from . import abi

STORAGE = {"f32": abi.STORE_NATIVE, "f32s": abi.STORE_F32S, "f16s": abi.STORE_F16S, "f16c": abi.STORE_F16C, "f16d": abi.STORE_F16D,
           "bf16s": abi.STORE_BF16S}


def node_updates_per_step(sim):
    return sum(sim.areas[a].nx * sim.areas[a].ny * sim.areas[a].m for a in range(sim.na))


def run(lib, geom, steps, storage=abi.STORE_NATIVE, chunk=1000, D=None, U=0.1, progress=False, tuning=None, fork=False,
        fused=True, **kw):
    sim = MultiBlockSim(lib, geom, storage=storage, flush_interval=min(chunk, 4096), fused=fused, **kw)
    if tuning:
        check(lib, lib.mbl_set_tuning(sim.h, tuning[0], tuning[1]), "tuning")
    if fork:
        check(lib, lib.mbl_set_cycle_fork(sim.h, 1), "fork")
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        n = min(chunk, steps - done)
        sim.step(n)
        done += n
        if progress:
            sim.sync()
            print(f"# step {done} / {steps}  {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    sim.sync()
    wall = time.perf_counter() - t0
    l2, drag = sim.series()
    nu = node_updates_per_step(sim)
    out = {"steps": steps, "wall_s": round(wall, 3), "mlups": round(nu * steps / wall / 1e6, 1),
           "node_updates_per_step": nu, "residual_final": float(l2["residual"][-1])}
    cd, cl = drag["Cd"], drag["Cl"]
    out["Cd_final"] = float(cd[-1])
    out["Cl_final"] = float(cl[-1])
    if D is not None:
        out.update(postproc.force_statistics(cd, cl, D=D, U=U, transient=0.5))
    # steady flow: the instantaneous Cd; shedding: the Cd mean over one shedding period
    window = None
    if D is not None and out.get("St", 0.0) > 0.02 and out.get("Cl_amp", 0.0) > 1e-3:
        window = int(round(D / (out["St"] * U)))
        out["shedding_period_steps"] = window
    out["t_converged_Cd"] = postproc.time_to_converged(cd, rel=0.005, window=window)
    sim.close()
    return out, {"t": drag["t"], "Cd": cd, "Cl": cl, "residual": l2["residual"]}


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("case", choices=["c1", "c2b"])
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--storage", default="f32", choices=list(STORAGE))
    ap.add_argument("--fp64", action="store_true")
    ap.add_argument("--save", default=None, help="write the Cd/Cl/residual series to this .npz")
    ap.add_argument("--progress", action="store_true")
    ap.add_argument("--vx", type=int, default=0, help="nodes per thread of the fused kernel (0: library default)")
    ap.add_argument("--tpb", type=int, default=0, help="threads per workgroup (0: library default)")
    ap.add_argument("--fork", action="store_true",
                    help="launches at one position of the scheduled cycle on forked side streams")
    ap.add_argument("--three-calls", action="store_true",
                    help="each step as evolution, error_L2 and getDrag_block (default: one mbl_evolution_step "
                         "call per chunk)")
    args = ap.parse_args(argv)
    lib = load_library(args.fp64)
    if args.case == "c1":
        geom, steps, D = c1_geometry(lib), args.steps or 20000, 10.0
    else:
        geom, steps, D = c2b_geometry(lib), args.steps or 50000, 16.0
    out, series = run(lib, geom, steps, storage=STORAGE[args.storage] if not args.fp64 else abi.STORE_NATIVE,
                      D=D, progress=args.progress, tuning=(args.vx, args.tpb) if (args.vx or args.tpb) else None,
                      fork=args.fork, fused=not args.three_calls)
    out.update({"case": args.case, "precision": "f64" if args.fp64 else args.storage})
    if args.save:
        np.savez(args.save, **series)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
