"""Force-coefficient post-processing against synthetic signals with known answers."""
import numpy as np
import pytest

from mbl import postproc


def test_strouhal_and_amplitudes_of_a_sinusoid():
    D, U = 16.0, 0.1
    f = 0.165 * U / D  # St = 0.165
    t = np.arange(60000, dtype=float)
    cl = 0.33 * np.sin(2 * np.pi * f * t + 0.4)
    cd = 1.41 + 0.012 * np.sin(4 * np.pi * f * t)
    s = postproc.force_statistics(cd, cl, D, U, transient=0.2)
    assert s["St"] == pytest.approx(0.165, rel=2e-3)
    assert s["Cl_rms"] == pytest.approx(0.33 / np.sqrt(2), rel=2e-3)
    assert s["Cl_amp"] == pytest.approx(0.33, rel=1e-3)
    assert s["Cd_mean"] == pytest.approx(1.41, rel=1e-4)
    assert s["Cd_amp"] == pytest.approx(0.012, rel=1e-2)


def test_zero_crossing_period():
    t = np.arange(10000, dtype=float)
    x = np.sin(2 * np.pi * t / 137.5) + 0.2
    assert postproc.zero_crossing_period(x) == pytest.approx(137.5, rel=1e-3)


def test_time_to_converged():
    t = np.arange(5000, dtype=float)
    x = 1.5 + 0.5 * np.exp(-t / 300.0)
    k = postproc.time_to_converged(x, rel=0.005)
    # |x - x_final| <= 0.005 x_final from t = 300 ln(0.5 / 0.0075) on
    assert k == pytest.approx(300 * np.log(0.5 / (0.005 * x[-1])), abs=2)


def test_read_reference_formats(tmp_path):
    p = tmp_path / "fdrag_t.dat"
    p.write_text("0\t1.5\t-0.01\t\n1\t1.4\t0.02\t\n")
    t, cd, cl = postproc.read_fdrag(str(p))
    assert list(t) == [0, 1] and list(cd) == [1.5, 1.4] and list(cl) == [-0.01, 0.02]
    q = tmp_path / "error_L2.dat"
    q.write_text("0\t0.5\n1\t0.25\n")
    t, r = postproc.read_error_l2(str(q))
    assert list(r) == [0.5, 0.25]


@pytest.mark.parametrize("phase", [0.0, 1.3, 4.0])
def test_phase_locked_statistics_of_a_sinusoid(phase):
    """Whole periods between interpolated zero crossings: a sinusoid with a period of a
    non-integer number of samples gives St, the mean, rms and amplitude to far below the
    1e-5 bar, whatever its phase (a shifted signal gives the same numbers)."""
    D, U = 16.0, 0.1
    period = 981.37
    t = np.arange(50000, dtype=float)
    w = 2 * np.pi * t / period + phase
    cl = 0.02 + 0.32 * np.sin(w)
    cd = 1.318 + 0.011 * np.sin(2 * w + 0.3)
    s = postproc.phase_locked_statistics(cd, cl, D, U, start=5000)
    assert s["St"] == pytest.approx(D / (U * period), rel=1e-7)
    assert s["Cd_mean"] == pytest.approx(1.318, rel=1e-7)
    assert s["Cl_mean"] == pytest.approx(0.02, abs=1e-7)
    assert s["Cl_rms"] == pytest.approx(0.32 / np.sqrt(2), rel=1e-6)
    assert s["Cl_amp"] == pytest.approx(0.32, rel=1e-7)
    sp = postproc.estimator_spread(cd, cl, D, U, start=5000)
    assert max(sp.values()) < 1e-6


def test_phase_locked_statistics_detects_a_changed_period():
    """A 1e-4 change of the period moves St by 1e-4 (the estimator resolves it)."""
    D, U = 16.0, 0.1
    t = np.arange(50000, dtype=float)
    a = postproc.phase_locked_statistics(np.ones_like(t), np.sin(2 * np.pi * t / 981.37), D, U, 5000)
    b = postproc.phase_locked_statistics(np.ones_like(t), np.sin(2 * np.pi * t / (981.37 * 1.0001)), D, U, 5000)
    assert (a["St"] - b["St"]) / a["St"] == pytest.approx(1e-4, rel=2e-3)


def test_crossings_ignore_sample_noise():
    """Noise of a few percent of the amplitude (16-bit planes) adds no zero crossings, and the
    statistics stay within the noise's own effect."""
    D, U = 16.0, 0.1
    period = 981.37
    t = np.arange(50000, dtype=float)
    rng = np.random.default_rng(5)
    clean = 0.32 * np.sin(2 * np.pi * t / period)
    noisy = clean + 0.01 * rng.standard_normal(t.size)
    assert len(postproc.upward_crossings(noisy, 0.0)) == len(postproc.upward_crossings(clean, 0.0))
    s = postproc.phase_locked_statistics(np.ones_like(t), noisy, D, U, start=5000)
    assert s["St"] == pytest.approx(D / (U * period), rel=1e-4)
