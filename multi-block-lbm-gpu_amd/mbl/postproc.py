"""Post-processing of the per-step force coefficients and residuals.

The reference writes one line per coarse step to fdrag_t.dat ("t\\tCd\\tCl\\t",
lbm/force_block.h:212-215) and error_L2.dat ("t\\tresidual",
lbm/error_L2.h:85-88) and stops there. These functions read those files (or
the series returned by the library) and reduce them to the quantities the
benchmark cases report: time-averaged Cd, Cl rms and amplitude,
Strouhal number, and the time to a converged Cd.

Time is in coarse steps (dt = 1 coarse lattice unit); St = f D / U with D the
cylinder diameter and U the reference velocity in the same units.
"""
import numpy as np


def read_fdrag(path):
    """fdrag_t.dat -> (t, Cd, Cl) arrays."""
    t, cd, cl = [], [], []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) >= 3:
                t.append(int(parts[0]))
                cd.append(float(parts[1]))
                cl.append(float(parts[2]))
    return np.array(t), np.array(cd), np.array(cl)


def read_error_l2(path):
    t, r = [], []
    with open(path) as fh:
        for line in fh:
            parts = line.split()
            if len(parts) >= 2:
                t.append(int(parts[0]))
                r.append(float(parts[1]))
    return np.array(t), np.array(r)


def dominant_frequency(x, dt=1.0):
    """Frequency of the largest spectral peak of x (mean removed, Hann window),
    refined by quadratic interpolation of the log-magnitude around the peak bin."""
    x = np.asarray(x, dtype=np.float64)
    n = x.size
    if n < 8:
        return 0.0
    w = np.hanning(n)
    y = (x - x.mean()) * w
    nfft = 1 << int(np.ceil(np.log2(4 * n)))
    spec = np.abs(np.fft.rfft(y, nfft))
    k = int(np.argmax(spec[1:])) + 1
    if 1 <= k < spec.size - 1:
        a, b, c = np.log(spec[k - 1] + 1e-300), np.log(spec[k] + 1e-300), np.log(spec[k + 1] + 1e-300)
        den = a - 2 * b + c
        delta = 0.5 * (a - c) / den if den != 0 else 0.0
    else:
        delta = 0.0
    return (k + delta) / (nfft * dt)


def zero_crossing_period(x):
    """Mean period from upward zero crossings of x - mean(x), linearly interpolated."""
    x = np.asarray(x, dtype=np.float64) - np.mean(x)
    s = np.nonzero((x[:-1] < 0) & (x[1:] >= 0))[0]
    if s.size < 2:
        return float("nan")
    tc = s + (-x[s]) / (x[s + 1] - x[s])
    return float((tc[-1] - tc[0]) / (tc.size - 1))


def whole_periods(x):
    """Indices [i0, i1) spanning an integer number of periods at the end of x."""
    x = np.asarray(x, dtype=np.float64) - np.mean(x)
    s = np.nonzero((x[:-1] < 0) & (x[1:] >= 0))[0]
    if s.size < 2:
        return 0, x.size
    return int(s[0]) + 1, int(s[-1]) + 1


def force_statistics(cd, cl, D, U, dt=1.0, transient=0.5):
    """Cd mean, Cd amplitude, Cl rms, Cl amplitude and St from the series after
    the first `transient` fraction (or that many steps when > 1)."""
    cd = np.asarray(cd, dtype=np.float64)
    cl = np.asarray(cl, dtype=np.float64)
    start = int(transient) if transient > 1 else int(transient * cd.size)
    cdw, clw = cd[start:], cl[start:]
    i0, i1 = whole_periods(clw)
    if i1 - i0 < 8:
        i0, i1 = 0, clw.size
    f = dominant_frequency(clw[i0:i1], dt)
    return {
        "Cd_mean": float(np.mean(cdw[i0:i1])),
        "Cd_amp": float(0.5 * (np.max(cdw[i0:i1]) - np.min(cdw[i0:i1]))),
        "Cl_rms": float(np.sqrt(np.mean((clw[i0:i1] - np.mean(clw[i0:i1])) ** 2))),
        "Cl_amp": float(0.5 * (np.max(clw[i0:i1]) - np.min(clw[i0:i1]))),
        "St": float(f * D / U),
        "samples": int(i1 - i0),
    }


def _integral(y, a, b):
    """Integral over [a, b] (sample-index units) of the piecewise-linear interpolant of y."""
    y = np.asarray(y, dtype=np.float64)

    def at(t):
        i = min(int(np.floor(t)), y.size - 2)
        return y[i] + (t - i) * (y[i + 1] - y[i])

    ia, ib = int(np.ceil(a)), int(np.floor(b))
    if ib < ia:  # a and b inside one interval
        return 0.5 * (at(a) + at(b)) * (b - a)
    s = 0.5 * (at(a) + y[ia]) * (ia - a) + 0.5 * (y[ib] + at(b)) * (b - ib)
    if ib > ia:
        s += np.sum(0.5 * (y[ia:ib] + y[ia + 1:ib + 1]))
    return float(s)


def _peak(y, i):
    """Extremum of the parabola through samples i-1, i, i+1 of y."""
    if i <= 0 or i >= y.size - 1:
        return float(y[i])
    a, b, c = y[i - 1], y[i], y[i + 1]
    den = a - 2.0 * b + c
    if den == 0.0:
        return float(b)
    d = 0.5 * (a - c) / den
    return float(b - 0.25 * (a - c) * d)


def upward_crossings(x, level, hyst=None):
    """Times (fractional sample index) where x crosses `level` upward, linearly interpolated.
    A crossing counts only once x has been below level - hyst since the previous one
    (hysteresis, default half the standard deviation, about a third of a sinusoid's
    amplitude), so sample noise near the level
    (16-bit planes) does not add crossings."""
    x = np.asarray(x, dtype=np.float64) - level
    if hyst is None:
        hyst = 0.5 * float(np.std(x))
    cand = np.nonzero((x[:-1] < 0) & (x[1:] >= 0))[0]
    low = np.nonzero(x < -hyst)[0]
    keep = []
    last = -1
    for i in cand:
        # armed: some sample below -hyst after the previous kept crossing and at or before i
        k = np.searchsorted(low, last + 1)
        if k < low.size and low[k] <= i:
            keep.append(i)
            last = i
    s = np.array(keep, dtype=np.int64)
    return s + (-x[s]) / (x[s + 1] - x[s])


def phase_locked_statistics(cd, cl, D, U, start=0, dt=1.0, periods=None, near=None):
    """Cd mean, Cl mean/rms/amplitude and St over the whole shedding periods of the series
    from `start` on. The periods are bounded by upward crossings of Cl through its mean
    (interpolated between samples, the mean taken over whole periods and the crossings
    found again with it), the averages are integrals of the piecewise-linear series between
    the first and last crossing, the amplitude is half the peak-to-peak of each period
    (parabolic peak refinement) averaged over the periods, and St = D / (U T) with T the mean
    period. A phase shift of the signal moves the window with it, so two runs that differ
    only in phase give the same statistics; `periods` fixes the number of whole periods
    (from the first crossing after `start`) so that two runs are compared over the same
    periods, and `near` (a step index, e.g. another run's "t_first") starts the window at
    the crossing closest to it instead, so that a run whose phase differs a little still
    covers the same periods. Returns None when fewer than two periods (or fewer than
    `periods`) exist."""
    cd = np.asarray(cd, dtype=np.float64)[start:]
    cl = np.asarray(cl, dtype=np.float64)[start:]
    level = float(np.mean(cl))
    tc = None
    for _ in range(3):
        t = upward_crossings(cl, level)
        if near is not None and t.size:
            t = t[int(np.argmin(np.abs(t + start - near))):]
        if t.size < 3 or (periods is not None and t.size < periods + 1):
            return None
        tc = t if periods is None else t[:periods + 1]
        new = _integral(cl, tc[0], tc[-1]) / (tc[-1] - tc[0])
        if new == level:
            break
        level = new
    T = tc[-1] - tc[0]
    n = tc.size - 1
    cd_mean = _integral(cd, tc[0], tc[-1]) / T
    cl_mean = _integral(cl, tc[0], tc[-1]) / T
    cl_rms = np.sqrt(max(0.0, _integral((cl - cl_mean) ** 2, tc[0], tc[-1]) / T))
    amps = []
    for p in range(n):
        i0, i1 = int(np.ceil(tc[p])), int(np.floor(tc[p + 1])) + 1
        seg = cl[i0:i1]
        if seg.size < 3:
            continue
        hi = _peak(cl, i0 + int(np.argmax(seg)))
        lo = _peak(cl, i0 + int(np.argmin(seg)))
        amps.append(0.5 * (hi - lo))
    period = T / n * dt
    return {"Cd_mean": float(cd_mean), "Cl_mean": float(cl_mean), "Cl_rms": float(cl_rms),
            "Cl_amp": float(np.mean(amps)) if amps else float("nan"), "St": float(D / (U * period)),
            "period": float(period), "periods": int(n), "t_first": float(tc[0] + start)}


def estimator_spread(cd, cl, D, U, start=0):
    """The phase-locked statistics over the first and the second half of the whole periods
    after `start`: their relative difference per quantity bounds what the estimator itself
    (window choice, residual drift of the limit cycle) contributes to a comparison."""
    full = phase_locked_statistics(cd, cl, D, U, start)
    if full is None:
        return None
    cl_ = np.asarray(cl, dtype=np.float64)[start:]
    tc = upward_crossings(cl_, full["Cl_mean"])
    mid = int(tc[tc.size // 2])
    a = phase_locked_statistics(cd[:start + mid + 2], cl[:start + mid + 2], D, U, start)
    b = phase_locked_statistics(cd, cl, D, U, start + mid - 1)
    if a is None or b is None:
        return None
    return {k: abs(a[k] - b[k]) / max(abs(full[k]), 1e-300) for k in ("Cd_mean", "Cl_rms", "Cl_amp", "St")}


def time_to_converged(x, rel=0.005, window=None):
    """First step after which the running mean of x over `window` steps (or the
    value itself) stays within rel of its final value; -1 if never."""
    x = np.asarray(x, dtype=np.float64)
    if window and window > 1:
        c = np.cumsum(np.insert(x, 0, 0.0))
        x = (c[window:] - c[:-window]) / window
        offset = window - 1
    else:
        offset = 0
    final = x[-1]
    bad = np.nonzero(np.abs(x - final) > rel * abs(final))[0]
    if bad.size == 0:
        return offset
    if bad[-1] == x.size - 1:
        return -1
    return int(bad[-1] + 1 + offset)
