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
