"""Per-result error of each plane storage against the FP64 build, and the time to a converged
drag coefficient, on the two-level (C1) and three-level (C2b) cylinder channels.

    python -m mbl.precision [--cases c1,c2b,poiseuille] [--storages f32,f32s,f16d,f16c,f16s,bf16s]
                            [--markdown FILE]

For each case the FP64 build runs first (the reference of every comparison; it equals the
FP64 C restatement of the reference to 1e-12 on these layouts), then each storage of the FP32
build. Reported per run:

* C1 (Re = 20, steady, 20,000 coarse steps): the final Cd and its relative error; the
  analytic-profile velocity L2 is reported separately (`poiseuille`).
* C2b (Re = 100, shedding, 50,000 coarse steps): Cd mean, Cl rms, Cl amplitude and St over the
  same whole shedding periods of the second half (phase-locked windows) and their relative
  errors.
* Time to a converged Cd: the first coarse step after which Cd (C1) or its running mean over
  one shedding period (C2b) stays within +-0.5 % of the FP64 run's converged value, in steps
  and in wall seconds of that run (steps x its measured seconds per step).

Prints one JSON line per case and, with --markdown, writes the tables.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import MultiBlockSim, c1_geometry, c2b_geometry, load_library, postproc
from . import abi
from .cases import STORAGE, node_updates_per_step

NAMES = {"f32": "FP32 planes (f)", "f32s": "FP32 planes (f - w)", "f16s": "IEEE FP16 (f - w)",
         "bf16s": "BF16 (f - w)", "f16c": "1-4-11 (f - w)", "f16d": "1-3-12 (f - w)", "f64": "FP64"}
D_C1, D_C2B, U = 10.0, 16.0, 0.1


def run_series(lib, geom, steps, storage, chunk=5000):
    sim = MultiBlockSim(lib, geom, storage=storage, flush_interval=4096)
    nu = node_updates_per_step(sim)
    sim.step(1)  # graph capture and first launches outside the clock
    sim.sync()
    t0 = time.perf_counter()
    done = 1
    while done < steps:
        n = min(chunk, steps - done)
        sim.step(n)
        done += n
    sim.sync()
    wall = time.perf_counter() - t0
    _, drag = sim.series()
    sim.close()
    return np.asarray(drag["Cd"]), np.asarray(drag["Cl"]), wall / (steps - 1), nu


def converged_step(cd, target, window=None, rel=0.005):
    """First step after which Cd (or its running mean over `window` steps) stays within rel of
    target; -1 if it never does."""
    x = np.asarray(cd, dtype=np.float64)
    off = 0
    if window and window > 1:
        c = np.cumsum(np.insert(x, 0, 0.0))
        x = (c[window:] - c[:-window]) / window
        off = window - 1
    bad = np.nonzero(np.abs(x - target) > rel * abs(target))[0]
    if bad.size == 0:
        return off
    if bad[-1] == x.size - 1:
        return -1
    return int(bad[-1] + 1 + off)


def case_c1(storages, steps=20000):
    lib64, lib32 = load_library(True), load_library(False)
    cd, cl, sps, nu = run_series(lib64, c1_geometry(lib64), steps, abi.STORE_NATIVE)
    target = float(cd[-1])
    rows = [{"storage": "f64", "Cd": target, "Cd_err": 0.0, "s_per_step": sps,
             "mlups": nu / sps / 1e6, "t_conv_steps": converged_step(cd, target)}]
    for st in storages:
        c, _, s, _ = run_series(lib32, c1_geometry(lib32), steps, STORAGE[st])
        rows.append({"storage": st, "Cd": float(c[-1]), "Cd_err": abs(c[-1] - target) / abs(target),
                     "s_per_step": s, "mlups": nu / s / 1e6, "t_conv_steps": converged_step(c, target)})
    for r in rows:
        r["t_conv_s"] = r["t_conv_steps"] * r["s_per_step"] if r["t_conv_steps"] >= 0 else None
    return {"case": "c1", "steps": steps, "node_updates_per_step": nu, "rows": rows}


def case_c2b(storages, steps=50000):
    lib64, lib32 = load_library(True), load_library(False)
    start = steps // 2
    cd, cl, sps, nu = run_series(lib64, c2b_geometry(lib64), steps, abi.STORE_NATIVE)
    s = postproc.phase_locked_statistics(cd, cl, D_C2B, U, start)
    n = s["periods"] - 1
    ref = postproc.phase_locked_statistics(cd, cl, D_C2B, U, start, periods=n)
    window = int(round(ref["period"]))
    target = ref["Cd_mean"]
    keys = ("Cd_mean", "Cl_rms", "Cl_amp", "St")
    rows = [dict({k: ref[k] for k in keys}, storage="f64", s_per_step=sps, mlups=nu / sps / 1e6,
                 t_conv_steps=converged_step(cd, target, window))]
    for st in storages:
        c, l, sp, _ = run_series(lib32, c2b_geometry(lib32), steps, STORAGE[st])
        r = postproc.phase_locked_statistics(c, l, D_C2B, U, start, periods=n, near=ref["t_first"])
        row = {"storage": st, "s_per_step": sp, "mlups": nu / sp / 1e6,
               "t_conv_steps": converged_step(c, target, window)}
        if r is None:
            row["error"] = "fewer shedding periods than the FP64 run"
        else:
            for k in keys:
                row[k] = r[k]
                row[k + "_err"] = abs(r[k] - ref[k]) / abs(ref[k])
        rows.append(row)
    for r in rows:
        r["t_conv_s"] = r["t_conv_steps"] * r["s_per_step"] if r["t_conv_steps"] >= 0 else None
    drift = postproc.estimator_spread(cd, cl, D_C2B, U, start)
    return {"case": "c2b", "steps": steps, "node_updates_per_step": nu, "periods": n,
            "limit_cycle_drift_between_halves": drift, "rows": rows}


def case_poiseuille(storages, checkpoints=(300, 1000, 3000)):
    """Velocity L2 distance from the analytic plane Poiseuille profile on an empty two-level
    channel (80 x 20 cells, refined block 30 x 8 cells) started at rest, per storage against
    the FP64 build, at several points of the flow's development."""
    from . import Geometry
    rows = []
    for st in ["f64"] + list(storages):
        lib = load_library(st == "f64")
        geom = Geometry.channel(lib, 81, 21, [(20, 50, 6, 14)], R=2.5, CX=-1000.0, CY=-1000.0, Re=20.0)
        sim = MultiBlockSim(lib, geom, storage=abi.STORE_NATIVE if st == "f64" else STORAGE[st])
        done, vals = 0, []
        for n in checkpoints:
            sim.step(n - done)
            done = n
            vals.append(sim.poiseuille_l2())
        sim.close()
        rows.append({"storage": st, "L2": vals})
    ref = rows[0]["L2"]
    for r in rows:
        r["L2_err"] = [abs(v - q) / q for v, q in zip(r["L2"], ref)]
    return {"case": "poiseuille", "checkpoints": list(checkpoints), "rows": rows}


def markdown(res):
    out = []
    for r in res:
        if r["case"] == "poiseuille":
            cps = r["checkpoints"]
            out.append(f"\nVelocity L2 against the analytic Poiseuille profile (empty two-level channel from rest; "
                       f"FP64 L2 at {', '.join(map(str, cps))} coarse steps: "
                       f"{', '.join(f'{v:.6f}' for v in r['rows'][0]['L2'])}), relative error per storage:\n")
            out.append("| Planes | " + " | ".join(f"{c} steps" for c in cps) + " |")
            out.append("|---|" + "---:|" * len(cps))
            for x in r["rows"][1:]:
                out.append(f"| {NAMES[x['storage']]} | " + " | ".join(f"{e:.1e}" for e in x["L2_err"]) + " |")
            continue
        if r["case"] == "c1":
            out.append(f"C1 (two-level channel, Re = 20, {r['steps']} coarse steps; FP64 final Cd "
                       f"{r['rows'][0]['Cd']:.10f}):\n")
            out.append("| Planes | Cd rel. error | MLUPS | steps to Cd within 0.5 % | wall s to Cd within 0.5 % |")
            out.append("|---|---:|---:|---:|---:|")
            for x in r["rows"]:
                out.append(f"| {NAMES[x['storage']]} | {x['Cd_err']:.1e} | {x['mlups']:.0f} | {x['t_conv_steps']} | "
                           f"{x['t_conv_s']:.3f} |" if x["t_conv_s"] is not None else
                           f"| {NAMES[x['storage']]} | {x['Cd_err']:.1e} | {x['mlups']:.0f} | never | — |")
        else:
            f = r["rows"][0]
            out.append(f"\nC2b (three-level channel, Re = 100, {r['steps']} coarse steps, {r['periods']} shedding "
                       f"periods of the second half; FP64: Cd mean {f['Cd_mean']:.6f}, Cl rms {f['Cl_rms']:.6f}, "
                       f"Cl amplitude {f['Cl_amp']:.6f}, St {f['St']:.6f}):\n")
            out.append("| Planes | Cd mean | Cl rms | Cl amplitude | St | MLUPS | steps to Cd mean within 0.5 % "
                       "| wall s |")
            out.append("|---|---:|---:|---:|---:|---:|---:|---:|")
            for x in r["rows"]:
                if x["storage"] == "f64":
                    errs = "| — | — | — | — "
                elif "error" in x:
                    errs = "| " + x["error"] + " | | | "
                else:
                    errs = "".join(f"| {x[k + '_err']:.1e} " for k in ("Cd_mean", "Cl_rms", "Cl_amp", "St"))
                ts = f"{x['t_conv_s']:.2f}" if x["t_conv_s"] is not None else "—"
                out.append(f"| {NAMES[x['storage']]} {errs}| {x['mlups']:.0f} | {x['t_conv_steps']} | {ts} |")
    return "\n".join(out) + "\n"


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--cases", default="c1,c2b,poiseuille")
    ap.add_argument("--storages", default="f32,f32s,f16d,f16c,f16s,bf16s")
    ap.add_argument("--markdown", default=None)
    args = ap.parse_args(argv)
    sts = [s for s in args.storages.split(",") if s]
    res = []
    for c in args.cases.split(","):
        print(f"# precision: {c}", file=sys.stderr, flush=True)
        r = {"c1": case_c1, "c2b": case_c2b, "poiseuille": case_poiseuille}[c](sts)
        res.append(r)
        print(json.dumps(r), flush=True)
    if args.markdown:
        with open(args.markdown, "w") as fh:
            fh.write(markdown(res))


if __name__ == "__main__":
    main()
