"""Host-side pieces of the warm-start pipeline (no GPU): the amplitude
envelope and the time to a converged amplitude on signals whose answer is
known, the network inputs, and the target normalisation."""
import numpy as np

from mbl import warmstart as ws


def test_amplitude_of_a_sinusoid_is_its_amplitude():
    T = 40
    t = np.arange(2000)
    a = ws.amplitude(0.3 + 1.7 * np.sin(2 * np.pi * t / T), T)[T - 1:]
    assert np.allclose(a, 1.7, rtol=0.01)


def test_time_to_amplitude_known_envelope():
    T = 50
    t = np.arange(20000, dtype=float)
    # amplitude saturating as 1 - exp(-t / tau): within 1% after tau ln(100) steps
    tau = 1500.0
    env = 1.0 - np.exp(-t / tau)
    cl = env * np.sin(2 * np.pi * t / T)
    # the converged amplitude measured the same way (the sampled peak of the sinusoid)
    a_ref = ws.amplitude(np.sin(2 * np.pi * t / T), T)[-1]
    got = ws.time_to_amplitude(cl, T, a_ref, tol=0.01)
    want = tau * np.log(100.0)
    assert want <= got <= want + 2 * T
    # never converged: still growing at the end
    assert ws.time_to_amplitude(cl[:3000], T, 1.0) == -1
    # converged from the start
    assert ws.time_to_amplitude(np.sin(2 * np.pi * t / T), T, 1.0) == T


def test_shedding_period():
    t = np.arange(10000)
    assert ws.shedding_period(np.sin(2 * np.pi * t / 173.0)) in (172, 173, 174)
    assert ws.shedding_period(np.ones(100)) == 0


def test_inputs_and_targets():
    specs = ws.c5_specs(4, seed=1)
    assert [s.Re for s in specs] == [s.Re for s in ws.c5_specs(4, seed=1)]
    for s in specs:
        assert 80 <= s.Re <= 200 and 2 * s.R in (8, 12) and abs(s.CY - round(s.CY) - 0.25) < 1e-12
    x = ws.raster_inputs(specs, 32, 128, 0.5)
    assert x.shape == (4, 32, 128, ws.IN_CH)
    for c, s in enumerate(specs):
        frac = x[c, :, :, 0].mean()
        area = np.pi * s.R ** 2 / ((s.LX - 1) * (s.LY - 1))
        assert 0.3 * area < frac < 3 * area
        assert np.allclose(x[c, :, :, 3], np.sin(0.5)) and np.allclose(x[c, :, :, 4], np.cos(0.5))
        assert x[c, 0, 0, 5] == 0 and x[c, -1, -1, 5] == 1 and x[c, -1, -1, 6] == 1
    rng = np.random.default_rng(2)
    f = np.stack([1 + rng.normal(0, 1e-3, (32, 128)), rng.normal(0, 0.05, (32, 128)),
                  rng.normal(0, 0.05, (32, 128))], -1)[None].repeat(4, 0)
    assert np.allclose(ws.from_target(ws.to_target(f, specs), specs), f, rtol=1e-12, atol=1e-14)
