"""Learned warm start of batched cylinder cases.

    python -m mbl.warmstart [--train 4096] [--test 64] [--spin 40000] [--steps 40000] [--epochs 40]
                            [--settled 0.01] [--snap-storage f32] ...

1. Training cases (two-level channels, D in {16, 24} finest cells, blockage
   D/H ~ U[0.1, 0.3], Re ~ U[80, 200], the cylinder 0.25 coarse cells above
   the channel centre line) start at rest with the parabolic profile and run
   `spin` steps (FP32 planes by default), then `nsnap` snapshots of
   (rho, ux, uy) are rasterised onto an H x W grid of each channel, every
   `every` steps. Cases whose Cl amplitude still drifts by more than
   `settled` (relative) over the snapshot window are not yet on their limit
   cycle and are left out of training.
2. Phase model (default): inputs per pixel are the solid mask, Re/200, D/H,
   sin and cos of the shedding phase (linear in time between the upward zero
   crossings of Cl that bracket the snapshot), x and y; targets
   (0.25 (rho - 1)/U^2, ux/U, uy/U). Harmonic model: the inputs without the
   phase, and as targets each case's least-squares mean and two shedding
   harmonics per pixel (15 channels).
   The encoder-decoder trains on the GPU; 1/8 of the training cases are held
   out for a validation loss.
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
from scipy.sparse import coo_matrix
from scipy.sparse.linalg import splu

from . import Batch, abi, case_spec, load_library
from .net import Net

IN_CH, OUT_CH = 7, 3


def c5_specs(n, seed, offset=0.25, re_min=80.0, re_max=200.0):
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(n):
        Dc = int(rng.choice([8, 12]))
        sp = case_spec(Dc, float(rng.uniform(0.1, 0.3)), float(rng.uniform(re_min, re_max)))
        sp.CY += offset
        out.append(sp)
    return out


def raster_inputs(specs, H, W, phase):
    """[n, H, W, 7]: mask, Re/200, D/H, sin and cos of the phase, x, y."""
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


RHO_SCALE = 0.25  # scale of the density target: its weight in the loss (measured best of 0.25 and 1)


STATIC_CH = 5  # harmonic model inputs: mask, Re/200, D/H, x, y
NHARM = 5      # mean, cos and sin of the phase and of twice the phase


def static_inputs(specs, H, W):
    """[n, H, W, 5]: the phase-free inputs (mask, Re/200, D/H, x, y)."""
    x = raster_inputs(specs, H, W, 0.0)
    return np.ascontiguousarray(x[..., [0, 1, 2, 5, 6]])


def harmonic_basis(phase):
    phase = np.asarray(phase, np.float64)
    return np.stack([np.ones_like(phase), np.cos(phase), np.sin(phase), np.cos(2 * phase), np.sin(2 * phase)], -1)


def harmonic_fit(targets, phases):
    """Least-squares mean + two shedding harmonics of a case's snapshots:
    targets [k, H, W, 3], phases [k] -> coefficients [H, W, 5 * 3] (harmonic-major)."""
    A = harmonic_basis(phases)
    k, H, W, C = targets.shape
    coef, *_ = np.linalg.lstsq(A, targets.reshape(k, -1).astype(np.float64), rcond=None)
    return coef.reshape(NHARM, H, W, C).transpose(1, 2, 0, 3).reshape(H, W, NHARM * C).astype(np.float32)


def harmonic_eval(coef, phase):
    """Targets from coefficients [..., H, W, 15]: at a scalar phase -> [..., H, W, 3]; one field
    [H, W, 15] at phases [k] -> [k, H, W, 3]; fields [n, H, W, 15] at phases [n] -> [n, H, W, 3]."""
    c = coef.reshape(coef.shape[:-1] + (NHARM, 3))
    b = harmonic_basis(phase)
    if np.ndim(phase) == 0:
        return np.einsum("...hc,h->...c", c, b)
    if c.ndim == 4:
        return np.einsum("yxhc,kh->kyxc", c, b)
    return np.einsum("nyxhc,nh->nyxc", c, b)


def to_target(fields, specs):
    U = np.array([sp.U for sp in specs])[:, None, None]
    t = np.empty_like(fields)
    t[..., 0] = RHO_SCALE * (fields[..., 0] - 1.0) / U ** 2
    t[..., 1] = fields[..., 1] / U
    t[..., 2] = fields[..., 2] / U
    return t


def from_target(t, specs):
    U = np.array([sp.U for sp in specs])[:, None, None]
    f = np.empty_like(t)
    f[..., 0] = 1.0 + t[..., 0] / RHO_SCALE * U ** 2
    f[..., 1] = t[..., 1] * U
    f[..., 2] = t[..., 2] * U
    return f


def poisson_density(fields, specs):
    """Density consistent with the rasters' velocity: rho = 1 + 3 p, with the
    kinematic pressure p of the incompressible momentum balance,
        lap p = -d_i d_j (u_i u_j)
    on the raster (central differences, pixel spacing in coarse cells, solid
    pixels taken as fluid at rest), p = 0 at the outlet column (the outlet
    holds rho = 1), dp/dx = nu u''(y) = -8 nu Umax / (LY - 2)^2 at the inlet
    (the parabolic profile) and dp/dy = 0 on the walls. A density that does
    not balance the velocity launches pressure waves through the channel that
    viscosity damps only slowly. Returns a copy of fields with channel 0
    replaced."""
    out = np.array(fields, np.float64)
    n, H, W, _ = out.shape
    for c, sp in enumerate(specs):
        dx = (sp.LX - 1.0) / (W - 1)
        dy = (sp.LY - 1.0) / (H - 1)
        Y = np.arange(H)[:, None] * dy
        X = np.arange(W)[None, :] * dx
        solid = (X - sp.CX) ** 2 + (Y - sp.CY) ** 2 < sp.R ** 2
        u = np.where(solid, 0.0, out[c, :, :, 1])
        v = np.where(solid, 0.0, out[c, :, :, 2])
        uu, uv, vv = u * u, u * v, v * v
        src = -(np.gradient(np.gradient(uu, dx, axis=1), dx, axis=1)
                + 2 * np.gradient(np.gradient(uv, dx, axis=1), dy, axis=0)
                + np.gradient(np.gradient(vv, dy, axis=0), dy, axis=0))
        nu = sp.U * 2 * sp.R / sp.Re
        g_in = -8.0 * nu * sp.U / (sp.LY - 2.0) ** 2
        # unknowns: every pixel left of the outlet column
        Wn = W - 1
        idx = np.arange(H * Wn).reshape(H, Wn)
        rows, cols, vals = [], [], []
        rhs = src[:, :Wn].copy()
        ax, ay = 1.0 / dx ** 2, 1.0 / dy ** 2
        for j in range(H):
            for i in range(Wn):
                r = idx[j, i]
                rows.append(r)
                cols.append(r)
                vals.append(-2 * ax - 2 * ay)
                # x neighbours: ghost p[-1] = p[1] - 2 dx g_in at the inlet; p = 0 at the outlet column
                for di in (-1, 1):
                    ii = i + di
                    if ii < 0:
                        rows.append(r)
                        cols.append(idx[j, 1])
                        vals.append(ax)
                        rhs[j, i] += 2 * dx * g_in * ax
                    elif ii < Wn:
                        rows.append(r)
                        cols.append(idx[j, ii])
                        vals.append(ax)
                # y neighbours: mirror at the walls
                for dj in (-1, 1):
                    jj = j + dj
                    if jj < 0 or jj >= H:
                        jj = j - dj
                    rows.append(r)
                    cols.append(idx[jj, i])
                    vals.append(ay)
        A = coo_matrix((vals, (rows, cols)), shape=(H * Wn, H * Wn)).tocsc()
        pr = splu(A).solve(rhs.ravel()).reshape(H, Wn)
        out[c, :, :Wn, 0] = 1.0 + 3.0 * pr
        out[c, :, Wn, 0] = 1.0
    return out.astype(np.float32)


def crossing_phase(cl, idx):
    """Shedding phase in [0, 2 pi) at sample idx: linear in time between the
    upward zero crossings of cl - mean(cl) that bracket idx; nan without both."""
    x = np.asarray(cl, np.float64) - np.mean(cl)
    s = np.nonzero((x[:-1] < 0) & (x[1:] >= 0))[0]
    tc = s + (-x[s]) / (x[s + 1] - x[s])
    k = np.searchsorted(tc, idx, side="right") - 1
    if k < 0 or k + 1 >= tc.size:
        return float("nan")
    return float(2 * np.pi * (idx - tc[k]) / (tc[k + 1] - tc[k]))


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


def run(lib, specs, steps, storage, init=None, snap=None, chunk=4096, log=None, probe=0, res_out=None):
    """Runs a batch for `steps` steps; returns (cd, cl, wall seconds, snapshots).
    snap = (start, every, count, H, W) records rasters at those steps. With probe > 0 the
    per-case residual (error_L2 of each case, mbl_batch_residual) after the first `probe`
    steps is appended to the list res_out: the relative change of the velocity field since the
    initial state, which decides the warm start's fallback."""
    b = Batch(lib, specs, storage=storage)
    if init is not None:
        b.init_from_raster(init)
    b.sync()
    n = len(specs)
    cd = np.zeros((n, steps))
    cl = np.zeros((n, steps))
    shots, shot_t = [], []
    snap_marks = set()
    if snap:
        s0, every, count, H, W = snap
        snap_marks = {s0 + k * every for k in range(count)}
    marks = snap_marks | ({probe} if probe > 0 else set())
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
        if probe > 0 and done == probe and res_out is not None:
            res_out.append(b.residual())
        if done in snap_marks:
            shots.append(b.raster(snap[3], snap[4]))
            shot_t.append(done)
        if log and done % (4 * chunk) < m:
            print(f"# {log}: step {done}/{steps} {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    b.sync()
    wall = time.perf_counter() - t0
    nu = b.node_updates()
    b.close()
    return cd, cl, wall, (shots, shot_t), nu


def predict(net, model, specs, H, W, phase):
    """Targets [n, H, W, 3] of the cases at phase (scalar or per case)."""
    if model == "phase":
        return net.forward(raster_inputs(specs, H, W, phase))
    coef = net.forward(static_inputs(specs, H, W))
    ph = np.broadcast_to(np.asarray(phase, np.float64), (len(specs),))
    return np.stack([harmonic_eval(coef[i], ph[i]) for i in range(len(specs))])


def convergence(cl_c, cl_w, cd_w, cd_c=None, res0=None, threshold=None, probe=0):
    """Per case: steps to convergence, cold and warm, and the number of fallbacks.

    A case that sheds in its cold run (a period of 8 steps or more and a Cl amplitude above
    1e-3) has converged when its Cl amplitude stays within 1 % of the cold run's converged
    amplitude; a steady case when its Cd stays within 0.5 % of the cold run's final Cd. A warm
    case falls back to a cold start, and is charged the steps it used plus the cold steps, when
    its residual after the first `probe` steps exceeds `threshold` (res0: per-case error_L2 of
    the warm run), or when its forces turn non-finite or exceed |Cd| > 50."""
    from .postproc import time_to_converged
    tc, tw, fallback = [], [], 0
    for c in range(cl_c.shape[0]):
        T_c = shedding_period(cl_c[c])
        sheds = T_c >= 8 and float(np.mean(amplitude(cl_c[c], T_c)[-2 * T_c:])) > 1e-3
        if sheds:
            a_ref = float(np.mean(amplitude(cl_c[c], T_c)[-2 * T_c:]))
            t_cold = time_to_amplitude(cl_c[c], T_c, a_ref)
        elif cd_c is not None:
            t_cold = time_to_converged(cd_c[c], rel=0.005)
        else:
            tc.append(-1)
            tw.append(-1)
            continue
        bad = ~np.isfinite(cd_w[c]) | (np.abs(cd_w[c]) > 50)
        resid_bad = res0 is not None and threshold is not None and not (res0[c] <= threshold)
        if resid_bad:
            fallback += 1
            t_warm = probe + t_cold if t_cold >= 0 else -1
        elif bad.any():
            fallback += 1
            t_warm = int(np.argmax(bad)) + 1 + t_cold if t_cold >= 0 else -1
        elif sheds:
            t_warm = time_to_amplitude(cl_w[c], T_c, a_ref)
        else:
            # the warm run converges to the same steady state; judged against the cold run's
            t_warm = time_to_converged_to(cd_w[c], float(cd_c[c][-1]), 0.005)
        tc.append(t_cold)
        tw.append(t_warm)
    return np.array(tc), np.array(tw), fallback


def time_to_converged_to(x, target, rel):
    """First step after which x stays within rel of target; -1 if never."""
    x = np.asarray(x, np.float64)
    bad = np.nonzero(np.abs(x - target) > rel * abs(target))[0]
    if bad.size == 0:
        return 0
    if bad[-1] == x.size - 1:
        return -1
    return int(bad[-1] + 1)


def fallback_cases(res0, threshold):
    """Cases whose initial residual exceeds the threshold (or is not finite): restarted cold."""
    r = np.asarray(res0, np.float64)
    return ~(r <= threshold)


def early_overshoot(cl_c, cl_w, periods=(1, 2)):
    """Median over cases of the warm run's Cl amplitude after the given numbers of
    shedding periods, relative to the cold run's converged amplitude."""
    res = {p: [] for p in periods}
    for c in range(cl_c.shape[0]):
        T_c = shedding_period(cl_c[c])
        if T_c < 8:
            continue
        a_ref = float(np.mean(amplitude(cl_c[c], T_c)[-2 * T_c:]))
        a_w = amplitude(cl_w[c], T_c)
        for p in periods:
            if p * T_c < a_w.size and a_ref > 0:
                res[p].append(a_w[p * T_c] / a_ref)
    return {f"amp_ratio_after_{p}T": round(float(np.median(v)), 3) for p, v in res.items() if v}


def summary(out, tc, tw, fallback, ms_step, t_inf):
    both = (tc > 0) & (tw > 0)
    out.update({
        "ms_per_step": round(ms_step, 4), "converged_cold": int((tc > 0).sum()), "converged_warm": int((tw > 0).sum()),
        "fallbacks": fallback,
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
    return out


def self_init(lib, args, storage, out):
    te = c5_specs(args.test, seed=202)
    cd_c, cl_c, wall_c, (shots, _), _ = run(lib, te, args.steps, storage,
                                             snap=(args.steps, 1, 1, args.H, args.W), log="cold")
    if args.dump:
        np.savez(args.dump, fields=shots[0], specs=np.array([[sp.LX, sp.LY, sp.R, sp.CX, sp.CY, sp.Re, sp.U]
                                                             for sp in te]),
                 cl=cl_c[:, -4000:].astype(np.float32), cd=cd_c[:, -4000:].astype(np.float32))
    init = shots[0].copy()
    if args.self_init_noise > 0:
        # uncorrelated pixel noise of the given size (relative to U) on the velocity
        rng = np.random.default_rng(11)
        U = np.array([sp.U for sp in te])[:, None, None]
        init[..., 1] += (args.self_init_noise * U * rng.standard_normal(init.shape[:3])).astype(init.dtype)
        init[..., 2] += (args.self_init_noise * U * rng.standard_normal(init.shape[:3])).astype(init.dtype)
        out["self_init_noise"] = args.self_init_noise
    if args.self_init_rho == "one":
        init[..., 0] = 1.0
    elif args.self_init_rho == "poisson":
        init = poisson_density(init, te)
    out["self_init_rho"] = args.self_init_rho
    cd_w, cl_w, wall_w, _, _ = run(lib, te, args.steps, storage, init=init, log="self-init")
    tc, tw, fb = convergence(cl_c, cl_w, cd_w)
    out["mode"] = "self-init"
    out["steps_run"] = args.steps
    summary(out, tc, tw, fb, 0.5 * (wall_c + wall_w) / args.steps * 1e3, 0.0)
    if args.save:
        np.savez(args.save, tc=tc, tw=tw, cl_cold=cl_c[:, ::10].astype(np.float32),
                 cl_warm=cl_w[:, ::10].astype(np.float32))
    print(json.dumps(out), flush=True)
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--train", type=int, default=4096)
    ap.add_argument("--test", type=int, default=64)
    ap.add_argument("--spin", type=int, default=40000)
    ap.add_argument("--nsnap", type=int, default=32)
    ap.add_argument("--every", type=int, default=50)
    ap.add_argument("--tail", type=int, default=1500)
    ap.add_argument("--steps", type=int, default=40000)
    ap.add_argument("--H", type=int, default=48)
    ap.add_argument("--W", type=int, default=192)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--epochs", type=int, default=None, help="default 40 (phase model), 400 (harmonic model)")
    ap.add_argument("--model", default="phase", choices=["harmonic", "phase"],
                    help="harmonic: the network predicts each pixel's mean and two shedding harmonics from the "
                         "geometry and Re; phase: it predicts the field at a given phase")
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--net-seed", type=int, default=2024,
                    help="seed of the network's initial weights; the batch order's seeds move with it")
    ap.add_argument("--storage", default="f32", choices=["f32", "f16s"], help="planes of the held-out runs")
    ap.add_argument("--snap-storage", default="f32", choices=["f32", "f16s", "f16d"],
                    help="planes of the snapshot runs (f16d: the 1-3-12 16-bit format)")
    ap.add_argument("--save", default=None, help="write the series and the trained parameters to this .npz")
    ap.add_argument("--self-init", action="store_true",
                    help="no network: each held-out case restarts from the raster of its own final cold state "
                         "(the best any network can do at this raster)")
    ap.add_argument("--self-init-rho", default="raster", choices=["raster", "one", "poisson"],
                    help="with --self-init: keep the raster's density, start from rho = 1 or solve for it")
    ap.add_argument("--self-init-noise", type=float, default=0.0,
                    help="with --self-init: add uncorrelated noise of this size (relative to U) to u and v")
    ap.add_argument("--dump", default=None, help="with --self-init: save the final rasters to this .npz")
    ap.add_argument("--diagnose", action="store_true",
                    help="also restart from hybrids of the network's field and each case's own final cold state "
                         "at the same phase (exact density + network velocity, and the reverse)")
    ap.add_argument("--settled", type=float, default=0.01,
                    help="keep only training cases whose Cl amplitude over the snapshot window stays within this "
                         "relative band of its mean (0: keep all)")
    ap.add_argument("--fluct-scale", default=None,
                    help="diagnostic: comma-separated factors a; also runs warm starts from mean + a (field - mean), "
                         "the mean being the network's average over 8 phases")
    ap.add_argument("--rho", default="net", choices=["net", "poisson"],
                    help="warm-start density: the network's or the pressure Poisson solution of its velocity")
    ap.add_argument("--train-re-min", type=float, default=40.0, help="training cases: Re ~ U[this, 200]")
    ap.add_argument("--test-re-min", type=float, default=40.0, help="held-out cases: Re ~ U[this, 200]")
    ap.add_argument("--probe", type=int, default=20,
                    help="coarse steps after a restart at which the per-case residual decides the fallback")
    ap.add_argument("--fallback-residual", type=float, default=0.15,
                    help="a warm case whose residual after --probe steps exceeds this restarts cold (measured "
                         "after 20 steps: warm starts from the network 0.04-0.10, cold starts about 0.22, a "
                         "corrupted field 0.52)")
    ap.add_argument("--ensemble", type=int, default=1,
                    help="train this many networks (seeds net-seed, net-seed+1000, ...) and average their fields")
    ap.add_argument("--load-params", default=None,
                    help="comma-separated .npz files of trained networks (their 'params'): skip snapshots and "
                         "training, evaluate the (averaged) networks on the held-out cases")
    args = ap.parse_args(argv)
    lib = load_library(False)
    storage = abi.STORE_NATIVE if args.storage == "f32" else abi.STORE_F16S
    snap_storage = {"f32": abi.STORE_NATIVE, "f16s": abi.STORE_F16S, "f16d": abi.STORE_F16D}[args.snap_storage]
    if args.epochs is None:
        args.epochs = 400 if args.model == "harmonic" else 40
    H, W = args.H, args.W
    out = {"train_cases": args.train, "test_cases": args.test, "raster": [H, W], "storage": args.storage,
           "snap_storage": args.snap_storage}
    if args.self_init:
        return self_init(lib, args, storage, out)

    if args.load_params:
        return evaluate_saved(lib, args, storage, out)
    # 1. snapshots of the training cases
    tr = c5_specs(args.train, seed=101, re_min=args.train_re_min)
    # the run goes on `tail` steps past the last snapshot so that a later zero crossing brackets it
    steps_tr = args.spin + args.nsnap * args.every + args.tail
    _, cl_tr, wall_tr, (shots, shot_t), nu_tr = run(lib, tr, steps_tr, snap_storage,
                                                    snap=(args.spin + args.every, args.every, args.nsnap, H, W),
                                                    log="train cases")
    out["snapshot_run_s"] = round(wall_tr, 2)
    out["snapshot_run_mlups"] = round(nu_tr * steps_tr / wall_tr / 1e6, 1)
    # training cases not yet on their limit cycle during the snapshot window put noise in the targets
    keep = np.ones(len(tr), bool)
    if args.settled > 0:
        for c in range(len(tr)):
            x = cl_tr[c, args.spin:]
            T = shedding_period(cl_tr[c])
            if T < 8 or float(np.mean(amplitude(x, T)[T - 1:])) <= 1e-3:
                # steady: kept when its Cl no longer moves over the snapshot window
                keep[c] = float(np.max(x) - np.min(x)) <= 1e-3
                continue
            a = amplitude(x, T)[T - 1:]
            keep[c] = np.max(np.abs(a - a.mean())) <= args.settled * a.mean()
        out["settled_cases"] = int(keep.sum())
        tr = [sp for sp, k in zip(tr, keep) if k]
        cl_tr = cl_tr[keep]
        shots = [sh[keep] for sh in shots]
    # per case: the snapshots' targets and phases
    case_t, case_ph = [], []
    for c, sp in enumerate(tr):
        x = cl_tr[c, args.spin:]
        ph = np.array([crossing_phase(x, t - 1 - args.spin) for t in shot_t])
        if not np.isfinite(ph).any() and float(np.max(x) - np.min(x)) <= 1e-3:
            # a steady case: the field does not depend on the phase, so its snapshots are spread
            # over all phases (the network learns that independence)
            ph = np.linspace(0.0, 2 * np.pi, len(shot_t), endpoint=False)
        tg = np.stack([to_target(shots[k][c:c + 1], [sp])[0] for k in range(len(shot_t))])
        ok = np.isfinite(ph) & np.isfinite(tg).all(axis=(1, 2, 3))
        case_t.append(tg[ok])
        case_ph.append(ph[ok])
    nval = max(1, args.train // 8)  # the last cases are held out for validation
    out["model"] = args.model
    if args.model == "phase":
        in_ch, out_ch = IN_CH, OUT_CH
        def samples(cases):
            X = np.concatenate([raster_inputs([tr[c]] * len(case_ph[c]), H, W, case_ph[c]) for c in cases])
            return X, np.concatenate([case_t[c] for c in cases]).astype(np.float32)
    else:
        in_ch, out_ch = STATIC_CH, NHARM * OUT_CH
        def samples(cases):
            use = [c for c in cases if len(case_ph[c]) >= 2 * NHARM]
            X = np.concatenate([static_inputs([tr[c]], H, W) for c in use])
            return X, np.stack([harmonic_fit(case_t[c], case_ph[c]) for c in use])
    Xt, Tt = samples(range(len(tr) - nval))
    Xv, Tv = samples(range(len(tr) - nval, len(tr)))
    out["samples"] = int(Xt.shape[0] + Xv.shape[0])

    # 2. training; the learning rate steps down (lr, lr/3, lr/10 over 50%, 30%, 20% of the epochs);
    # with --ensemble K, K networks from different seeds, whose predicted fields are averaged
    batch = min(args.batch, Xt.shape[0])
    t0 = time.perf_counter()
    nets, hists = [], []
    for e in range(args.ensemble):
        seed = args.net_seed + 1000 * e
        net = Net(lib, H, W, in_ch, out_ch, base=args.base, max_batch=max(batch, args.batch), seed=seed)
        hist = []
        e1 = max(1, int(0.5 * args.epochs))
        e2 = max(1, int(0.3 * args.epochs))
        for ep, lr in ((e1, args.lr), (e2, args.lr / 3), (max(1, args.epochs - e1 - e2), args.lr / 10)):
            done_ep = 0
            while done_ep < ep:  # calls of a few epochs: Adam state carries over, progress is printed
                k = min(ep - done_ep, 5 if args.model == "phase" else 25)
                hist.extend(net.train(Xt, Tt, batch=batch, epochs=k, lr=lr, seed=99 + len(hist) + seed - 2024))
                done_ep += k
                print(f"# net {e} epoch {len(hist)} loss {hist[-1]:.4e} {time.perf_counter() - t0:.1f} s",
                      file=sys.stderr, flush=True)
        nets.append(net)
        hists.append(np.array(hist))
    net = Ensemble(nets)
    hist = hists[0]
    out["train_s"] = round(time.perf_counter() - t0, 2)
    out["epochs"] = int(hist.size)
    out["net_seed"] = args.net_seed
    out["ensemble"] = args.ensemble
    out["epoch_loss_first_last"] = [[float(x[0]), float(x[-1])] for x in hists]
    # validation on the held-out cases' snapshots, in the target units of the phase model
    vcases = range(len(tr) - nval, len(tr))
    Tsnap = np.concatenate([case_t[c] for c in vcases])
    if args.model == "phase":
        Psnap = net.forward(Xv)
    else:
        pv = net.forward(np.concatenate([static_inputs([tr[c]], H, W) for c in vcases]))
        Psnap = np.concatenate([harmonic_eval(pv[i], case_ph[c]) for i, c in enumerate(vcases)])
    err = (Psnap - Tsnap) ** 2
    out["val_loss"] = float(np.mean(err))
    out["val_loss_per_channel"] = [float(v) for v in np.mean(err, axis=(0, 1, 2))]
    out["target_var_per_channel"] = [float(v) for v in np.var(Tsnap, axis=(0, 1, 2))]
    # validation error along the channel (8 bins in x), per channel, relative to the target variance
    ev = np.mean(err, axis=(0, 1)).reshape(8, -1, OUT_CH).mean(axis=1)
    out["val_rel_err_along_x"] = np.round(ev / np.var(Tsnap, axis=(0, 1, 2)), 4).T.tolist()
    print(f"# trained: loss {hist[0]:.4e} -> {hist[-1]:.4e} in {out['train_s']} s, val {out['val_loss']:.4e}",
          file=sys.stderr, flush=True)

    # 3. held-out cases, cold and warm
    te, cd_c, cl_c, cd_w, cl_w, tc, tw, fallback, pred, fin = held_out(lib, args, storage, net, out)
    out.update(early_overshoot(cl_c, cl_w))
    if args.fluct_scale:
        phs = np.linspace(0.0, 2 * np.pi, 8, endpoint=False)
        mean = np.mean([predict(net, args.model, te, H, W, p) for p in phs], axis=0)
        fsd = {}
        for a in [float(x) for x in args.fluct_scale.split(",")]:
            init_a = from_target(mean + a * (pred - mean), te).astype(np.float32)
            cd_a, cl_a, _, _, _ = run(lib, te, args.steps, storage, init=init_a, log=f"fluctuation x{a}")
            tca, twa, fba = convergence(cl_c, cl_a, cd_a)
            both_a = (tca > 0) & (twa > 0)
            r = {"median_steps": float(np.median(twa[both_a])) if both_a.any() else None, "fallbacks": fba}
            if both_a.any():
                r["speedup_median_steps"] = round(float(np.median(tca[both_a]) / np.median(twa[both_a])), 3)
                r["speedup_total_steps"] = round(float(tca[both_a].sum() / twa[both_a].sum()), 3)
            r.update(early_overshoot(cl_c, cl_a))
            fsd[str(a)] = r
            print(f"# fluctuation x{a}: {r}", file=sys.stderr, flush=True)
        out["fluct_scale"] = fsd
    if args.diagnose:
        # phase of each case's final cold state: linear from its last upward zero crossing over one period
        ph = np.zeros(len(te))
        for c in range(len(te)):
            T_c = max(shedding_period(cl_c[c]), 1)
            x = cl_c[c, -4 * T_c:] - np.mean(cl_c[c, -4 * T_c:])
            s_up = np.nonzero((x[:-1] < 0) & (x[1:] >= 0))[0]
            t_last = s_up[-1] + (-x[s_up[-1]]) / (x[s_up[-1] + 1] - x[s_up[-1]]) if s_up.size else 0.0
            ph[c] = 2 * np.pi * ((x.size - 1 - t_last) / T_c)
        exact = fin[0]
        netf = from_target(predict(net, args.model, te, H, W, ph), te).astype(np.float32)
        # the network's error on the held-out cases at their final phase, per channel, relative to
        # the variance of the exact fields (target units), over all pixels and over fluid pixels
        # within 3 diameters downstream of the cylinder
        te_t = np.stack([to_target(exact[c:c + 1], [te[c]])[0] for c in range(len(te))])
        pr_t = predict(net, args.model, te, H, W, ph)
        err = (pr_t - te_t) ** 2
        near = np.zeros(err.shape[:3], bool)
        for c, sp in enumerate(te):
            xs = np.linspace(0.0, sp.LX - 1.0, W)[None, :]
            ys = np.linspace(0.0, sp.LY - 1.0, H)[:, None]
            d = 2.0 * sp.R
            near[c] = (xs > sp.CX - d) & (xs < sp.CX + 3 * d) & (np.abs(ys - sp.CY) < 1.5 * d)
        var = np.var(te_t, axis=(0, 1, 2))
        out["test_rel_err"] = [round(float(v), 5) for v in np.mean(err, axis=(0, 1, 2)) / var]
        out["test_rel_err_near_wake"] = [round(float(v), 5) for v in err[near].mean(axis=0) / var]
        # per case: how far the cold run is from a settled limit cycle at its end (largest relative
        # deviation of the Cl amplitude over its last 8 periods), the network's near-wake error in v,
        # and the steps to the converged amplitude cold and warm
        settle = []
        for c in range(len(te)):
            T_c = max(shedding_period(cl_c[c]), 8)
            a = amplitude(cl_c[c, -8 * T_c:], T_c)[T_c - 1:]
            settle.append(float(np.max(np.abs(a - a.mean())) / max(a.mean(), 1e-12)))
        out["per_case"] = {"settle": [round(x, 4) for x in settle],
                           "v_err_near": [round(float(err[c][near[c]][:, 2].mean() / var[2]), 4) for c in range(len(te))],
                           "tc": [int(x) for x in tc], "tw": [int(x) for x in tw]}
        variants = {"net_matched_phase": netf,
                    "exact_rho_net_u": np.concatenate([exact[..., :1], netf[..., 1:]], -1),
                    "net_rho_exact_u": np.concatenate([netf[..., :1], exact[..., 1:]], -1)}
        diag = {}
        for name, init_v in variants.items():
            cd_v, cl_v, _, _, _ = run(lib, te, args.steps, storage, init=init_v, log=name)
            tcv, twv, fbv = convergence(cl_c, cl_v, cd_v)
            both_v = (tcv > 0) & (twv > 0)
            diag[name] = {"median_steps": float(np.median(twv[both_v])) if both_v.any() else None,
                          "speedup_median_steps": round(float(np.median(tcv[both_v]) / np.median(twv[both_v])), 3)
                          if both_v.any() else None, "fallbacks": fbv}
        out["diagnose"] = diag
    if args.save:
        np.savez(args.save, tc=tc, tw=tw, cl_cold=cl_c[:, ::10].astype(np.float32),
                 cl_warm=cl_w[:, ::10].astype(np.float32), loss=hist,
                 **{("params" if i == 0 else f"params{i}"): n_.get_params() for i, n_ in enumerate(net.nets)})
    net.close()
    print(json.dumps(out), flush=True)
    return out


class Ensemble:
    """Networks whose predictions are averaged (one network: that network)."""

    def __init__(self, nets):
        self.nets = nets

    def forward(self, x):
        return np.mean([n.forward(x) for n in self.nets], axis=0)

    def close(self):
        for n in self.nets:
            n.close()


def held_out(lib, args, storage, net, out):
    """The held-out cases from step 0, cold and warm; per-case residual after --probe steps of
    both (the warm one decides the fallback); convergence and the summary into `out`."""
    H, W = args.H, args.W
    te = c5_specs(args.test, seed=202, re_min=args.test_re_min)
    res_c, res_w = [], []
    cd_c, cl_c, wall_c, (fin, _), nu = run(lib, te, args.steps, storage, log="cold",
                                           snap=(args.steps, 1, 1, H, W) if args.diagnose else None,
                                           probe=args.probe, res_out=res_c)
    t0 = time.perf_counter()
    pred = predict(net, args.model, te, H, W, 0.0)
    init = from_target(pred, te).astype(np.float32)
    if args.rho == "poisson":
        init = poisson_density(init, te)
    out["rho"] = args.rho
    t_inf = time.perf_counter() - t0
    cd_w, cl_w, wall_w, _, _ = run(lib, te, args.steps, storage, init=init, log="warm", probe=args.probe,
                                   res_out=res_w)
    ms_step = 0.5 * (wall_c + wall_w) / args.steps * 1e3
    tc, tw, fallback = convergence(cl_c, cl_w, cd_w, cd_c=cd_c, res0=res_w[0], threshold=args.fallback_residual,
                                   probe=args.probe)
    out["steps_run"] = args.steps
    out["mlups"] = round(nu * args.steps / wall_c / 1e6, 1)
    out["test_re_min"] = args.test_re_min
    out["steady_cases"] = int(sum(1 for c in range(len(te)) if shedding_period(cl_c[c]) < 8 or
                                  float(np.mean(amplitude(cl_c[c], max(shedding_period(cl_c[c]), 8))[-16:])) <= 1e-3))
    out["probe_steps"] = args.probe
    out["fallback_residual"] = args.fallback_residual
    out["residual_after_probe"] = {
        "cold_median": float(np.median(res_c[0])), "cold_max": float(np.max(res_c[0])),
        "warm_median": float(np.median(res_w[0])), "warm_max": float(np.max(res_w[0]))}
    summary(out, tc, tw, fallback, ms_step, t_inf)
    return te, cd_c, cl_c, cd_w, cl_w, tc, tw, fallback, pred, fin


def evaluate_saved(lib, args, storage, out):
    """Held-out evaluation of saved networks (their predictions averaged): no snapshots, no training."""
    nets = []
    in_ch = IN_CH if args.model == "phase" else STATIC_CH
    out_ch = OUT_CH if args.model == "phase" else NHARM * OUT_CH
    for path in args.load_params.split(","):
        d = np.load(path, allow_pickle=False)
        keys = sorted(k for k in d.files if k.startswith("params"))
        for k in keys:
            n_ = Net(lib, args.H, args.W, in_ch, out_ch, base=args.base, max_batch=args.batch, seed=1)
            n_.set_params(d[k])
            nets.append(n_)
    net = Ensemble(nets)
    out["model"] = args.model
    out["ensemble"] = len(nets)
    out["loaded"] = args.load_params
    held_out(lib, args, storage, net, out)
    net.close()
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()
