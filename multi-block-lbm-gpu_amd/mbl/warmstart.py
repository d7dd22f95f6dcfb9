"""Learned warm start of batched cylinder cases.

    python -m mbl.warmstart [--train 128] [--test 64] [--spin 40000] [--steps 60000] [--epochs 30] ...

1. Training cases (two-level channels, D in {16, 24} finest cells, blockage
   D/H ~ U[0.1, 0.3], Re ~ U[80, 200], the cylinder 0.25 coarse cells above
   the channel centre line) start at rest with the parabolic profile and run
   `spin` steps, then `nsnap` snapshots of (rho, ux, uy) are rasterised onto an
   H x W grid of each channel, every `every` steps.
2. Inputs per pixel: solid mask, Re/200, D/H, sin and cos of the shedding
   phase (from the analytic signal of Cl), x and y; targets:
   ((rho - 1)/U^2, ux/U, uy/U). The encoder-decoder trains on the GPU.
3. Held-out cases run twice from step 0: cold (profile at rest) and warm
   (the network's field at phase 0, with the Chapman-Enskog non-equilibrium
   part). For each, the steps until the Cl amplitude (half the peak-to-peak
   over one shedding period) stays within 1% of the converged amplitude (the
   cold run's over its last two periods). A warm case whose forces turn
   non-finite or exceed |Cd| > 50 falls back to a cold start, and is charged
   the steps it used plus the cold steps.

Prints one JSON line with the medians, the ratio and the wall times.
"""
import argparse
import json
import sys
import time

import numpy as np
from scipy.ndimage import maximum_filter1d, minimum_filter1d
from scipy.signal import hilbert

from . import Batch, abi, case_spec, load_library
from .net import Net

IN_CH, OUT_CH = 7, 3


def c5_specs(n, seed, offset=0.25):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        Dc = int(rng.choice([8, 12]))
        sp = case_spec(Dc, float(rng.uniform(0.1, 0.3)), float(rng.uniform(80.0, 200.0)))
        sp.CY += offset
        out.append(sp)
    return out


def raster_inputs(specs, H, W, phase):
    """[n, H, W, 7]: mask, Re/200, D/H, sin phase, cos phase, x, y."""
    n = len(specs)
    x = np.zeros((n, H, W, IN_CH), np.float32)
    phase = np.broadcast_to(np.asarray(phase, np.float64), (n,))
    for c, sp in enumerate(specs):
        Y = np.arange(H)[:, None] * (sp.LY - 1.0) / (H - 1)
        X = np.arange(W)[None, :] * (sp.LX - 1.0) / (W - 1)
        x[c, :, :, 0] = ((X - sp.CX) ** 2 + (Y - sp.CY) ** 2 < sp.R ** 2)
        x[c, :, :, 1] = sp.Re / 200.0
        x[c, :, :, 2] = 2 * sp.R / (sp.LY - 1.0)
        x[c, :, :, 3] = np.sin(phase[c])
        x[c, :, :, 4] = np.cos(phase[c])
        x[c, :, :, 5] = X / (sp.LX - 1.0)
        x[c, :, :, 6] = Y / (sp.LY - 1.0)
    return x


def to_target(fields, specs):
    U = np.array([sp.U for sp in specs])[:, None, None]
    t = np.empty_like(fields)
    t[..., 0] = (fields[..., 0] - 1.0) / U ** 2
    t[..., 1] = fields[..., 1] / U
    t[..., 2] = fields[..., 2] / U
    return t


def from_target(t, specs):
    U = np.array([sp.U for sp in specs])[:, None, None]
    f = np.empty_like(t)
    f[..., 0] = 1.0 + t[..., 0] * U ** 2
    f[..., 1] = t[..., 1] * U
    f[..., 2] = t[..., 2] * U
    return f


def shedding_period(cl):
    """Mean period (steps) of the upward zero crossings of cl - mean over its second half."""
    x = np.asarray(cl[len(cl) // 2:], np.float64)
    x = x - x.mean()
    s = np.nonzero((x[:-1] < 0) & (x[1:] >= 0))[0]
    if s.size < 3:
        return 0
    return int(round((s[-1] - s[0]) / (s.size - 1)))


def amplitude(cl, T):
    """Half the peak-to-peak of cl over the last T steps, at every step (valid from T-1)."""
    hi = maximum_filter1d(cl, T, origin=(T - 1) // 2)
    lo = minimum_filter1d(cl, T, origin=(T - 1) // 2)
    return 0.5 * (hi - lo)


def time_to_amplitude(cl, T, a_ref, tol=0.01):
    """First step from which the amplitude stays within tol * a_ref of a_ref; -1 if never."""
    a = amplitude(np.asarray(cl, np.float64), T)[T - 1:]
    bad = np.nonzero(np.abs(a - a_ref) > tol * a_ref)[0]
    if bad.size == 0:
        return T
    if bad[-1] == a.size - 1:
        return -1
    return int(bad[-1] + 1 + T)


def run(lib, specs, steps, storage, init=None, snap=None, chunk=4096, log=None):
    """Runs a batch for `steps` steps; returns (cd, cl, wall seconds, snapshots).
    snap = (start, every, count, H, W) records rasters at those steps."""
    b = Batch(lib, specs, storage=storage)
    if init is not None:
        b.init_from_raster(init)
    b.sync()
    n = len(specs)
    cd = np.zeros((n, steps))
    cl = np.zeros((n, steps))
    shots, shot_t = [], []
    marks = set()
    if snap:
        s0, every, count, H, W = snap
        marks = {s0 + k * every for k in range(count)}
    t0 = time.perf_counter()
    done = 0
    while done < steps:
        m = min(chunk, steps - done)
        nxt = [t for t in marks if done < t <= done + m]
        if nxt:
            m = min(nxt) - done
        b.step(m)
        d, l = b.forces(m)
        cd[:, done:done + m] = d
        cl[:, done:done + m] = l
        done += m
        if done in marks:
            shots.append(b.raster(snap[3], snap[4]))
            shot_t.append(done)
        if log and done % (8 * chunk) < m:
            print(f"# {log}: step {done}/{steps} {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    b.sync()
    wall = time.perf_counter() - t0
    nu = b.node_updates()
    b.close()
    return cd, cl, wall, (shots, shot_t), nu


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=128)
    ap.add_argument("--test", type=int, default=64)
    ap.add_argument("--spin", type=int, default=40000)
    ap.add_argument("--nsnap", type=int, default=32)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--steps", type=int, default=60000)
    ap.add_argument("--H", type=int, default=32)
    ap.add_argument("--W", type=int, default=128)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=30)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--storage", default="f32", choices=["f32", "f16s"])
    ap.add_argument("--save", default=None, help="write the series and the trained parameters to this .npz")
    args = ap.parse_args(argv)
    lib = load_library(False)
    storage = abi.STORE_NATIVE if args.storage == "f32" else abi.STORE_F16S
    H, W = args.H, args.W
    out = {"train_cases": args.train, "test_cases": args.test, "raster": [H, W], "storage": args.storage}

    # 1. snapshots of the training cases
    tr = c5_specs(args.train, seed=101)
    steps_tr = args.spin + args.nsnap * args.every
    _, cl_tr, wall_tr, (shots, shot_t), nu_tr = run(lib, tr, steps_tr, storage,
                                                    snap=(args.spin + args.every, args.every, args.nsnap, H, W),
                                                    log="train cases")
    out["snapshot_run_s"] = round(wall_tr, 2)
    out["snapshot_run_mlups"] = round(nu_tr * steps_tr / wall_tr / 1e6, 1)
    xs, ts = [], []
    for c, sp in enumerate(tr):
        x = cl_tr[c, args.spin // 2:]
        ph = np.angle(hilbert(x - x.mean()))
        for k, t in enumerate(shot_t):
            phase = ph[t - 1 - args.spin // 2]
            xs.append(raster_inputs([sp], H, W, phase)[0])
            ts.append(to_target(shots[k][c:c + 1], [sp])[0])
    X = np.stack(xs)
    T = np.stack(ts).astype(np.float32)
    ok = np.isfinite(T).all(axis=(1, 2, 3))
    X, T = X[ok], T[ok]
    out["samples"] = int(X.shape[0])

    # 2. training
    net = Net(lib, H, W, IN_CH, OUT_CH, base=args.base, max_batch=args.batch, seed=2024)
    t0 = time.perf_counter()
    hist = net.train(X, T, batch=args.batch, epochs=args.epochs, lr=args.lr, seed=99)
    out["train_s"] = round(time.perf_counter() - t0, 2)
    out["epoch_loss_first_last"] = [float(hist[0]), float(hist[-1])]
    print(f"# trained: loss {hist[0]:.4e} -> {hist[-1]:.4e} in {out['train_s']} s", file=sys.stderr, flush=True)

    # 3. held-out cases, cold and warm
    te = c5_specs(args.test, seed=202)
    cd_c, cl_c, wall_c, _, nu = run(lib, te, args.steps, storage, log="cold")
    t0 = time.perf_counter()
    pred = net.forward(raster_inputs(te, H, W, 0.0))
    init = from_target(pred, te).astype(np.float32)
    t_inf = time.perf_counter() - t0
    cd_w, cl_w, wall_w, _, _ = run(lib, te, args.steps, storage, init=init, log="warm")
    ms_step = 0.5 * (wall_c + wall_w) / args.steps * 1e3
    tc, tw, fallback = [], [], 0
    for c in range(len(te)):
        T_c = shedding_period(cl_c[c])
        if T_c < 8:
            tc.append(-1)
            tw.append(-1)
            continue
        a_ref = float(np.mean(amplitude(cl_c[c], T_c)[-2 * T_c:]))
        t_cold = time_to_amplitude(cl_c[c], T_c, a_ref)
        bad = ~np.isfinite(cd_w[c]) | (np.abs(cd_w[c]) > 50)
        if bad.any():
            fallback += 1
            t_warm = int(np.argmax(bad)) + 1 + t_cold if t_cold >= 0 else -1
        else:
            t_warm = time_to_amplitude(cl_w[c], T_c, a_ref)
        tc.append(t_cold)
        tw.append(t_warm)
    tc, tw = np.array(tc), np.array(tw)
    both = (tc > 0) & (tw > 0)
    out.update({
        "steps_run": args.steps, "ms_per_step": round(ms_step, 4), "mlups": round(nu * args.steps / wall_c / 1e6, 1),
        "converged_cold": int((tc > 0).sum()), "converged_warm": int((tw > 0).sum()), "fallbacks": fallback,
        "median_steps_cold": float(np.median(tc[both])) if both.any() else None,
        "median_steps_warm": float(np.median(tw[both])) if both.any() else None,
        "inference_s": round(t_inf, 3),
    })
    if both.any():
        out["speedup_median_steps"] = round(float(np.median(tc[both]) / np.median(tw[both])), 3)
        out["speedup_total_steps"] = round(float(tc[both].sum() / tw[both].sum()), 3)
        wall_cold = tc[both].sum() * ms_step / 1e3
        wall_warm = tw[both].sum() * ms_step / 1e3 + t_inf
        out["time_to_shedding_cold_s"] = round(wall_cold, 3)
        out["time_to_shedding_warm_s"] = round(wall_warm, 3)
        out["speedup_wall"] = round(wall_cold / wall_warm, 3)
    if args.save:
        np.savez(args.save, tc=tc, tw=tw, cl_cold=cl_c.astype(np.float32), cl_warm=cl_w.astype(np.float32),
                 loss=hist, params=net.get_params())
    net.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
