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


def test_crossing_phase_of_a_sinusoid():
    T = 97.0
    t = np.arange(3000)
    cl = 0.2 + np.sin(2 * np.pi * t / T + 0.3)
    for idx in (400, 1234, 2500):
        want = (2 * np.pi * idx / T + 0.3) % (2 * np.pi)
        got = ws.crossing_phase(cl, idx)
        assert abs(np.angle(np.exp(1j * (got - want)))) < 0.02
    assert np.isnan(ws.crossing_phase(cl, 10)) and np.isnan(ws.crossing_phase(cl, 2990))


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


def test_chapman_enskog_state_of_a_relaxed_shear_wave():
    """The non-equilibrium part that mbl_batch_init_from_raster writes,
    -(tau - 1) w_k rho / cs^2 (e_k e_k - cs^2 I) : grad u, is what a BGK run
    with pull streaming and post-collision storage settles into. A periodic
    shear wave ux = A sin(k y) runs from F = feq for 300 steps (the initial
    layer decays within a few tau) in plain numpy; the stored F minus feq of
    its own moments then matches the formula for the diagonal directions to
    O((k)^2). A factor tau instead of tau - 1, or the opposite sign, fails."""
    N, tau, A = 64, 0.8, 0.01
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
    ey = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
    w = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
    kw = 2 * np.pi / N
    y = np.arange(N)[:, None] * np.ones((1, N))

    def feq(rho, ux, uy):
        eu = ex[:, None, None] * ux + ey[:, None, None] * uy
        return w[:, None, None] * rho * (1 + 3 * eu + 4.5 * eu ** 2 - 1.5 * (ux ** 2 + uy ** 2))

    F = feq(np.ones((N, N)), A * np.sin(kw * y), np.zeros((N, N)))
    for _ in range(300):
        f = np.stack([np.roll(np.roll(F[k], ey[k], axis=0), ex[k], axis=1) for k in range(9)])  # pull
        rho = f.sum(0)
        ux = (f * ex[:, None, None]).sum(0) / rho
        uy = (f * ey[:, None, None]).sum(0) / rho
        F = f - (f - feq(rho, ux, uy)) / tau
    rho = F.sum(0)
    ux = (F * ex[:, None, None]).sum(0) / rho
    uy = (F * ey[:, None, None]).sum(0) / rho
    neq = F - feq(rho, ux, uy)
    # d ux / dy from the wave's current amplitude and phase (spectral)
    c = np.fft.rfft(ux[:, 0])[1] * 2 / N
    dux_dy = np.real(1j * kw * c * np.exp(1j * kw * y))
    for k in (5, 6, 7, 8):
        qg = ex[k] * ey[k] * dux_dy  # e_a e_b d_a u_b - cs^2 div u, div u = 0
        want = -(tau - 1) * w[k] * rho * 3 * qg
        ratio = np.sum(neq[k] * want) / np.sum(want * want)
        assert abs(ratio - 1) < 0.02, (k, ratio)


def test_poisson_density_of_plane_poiseuille_flow():
    """A parabolic channel profile without a body: lap p = 0 inside, the
    inlet gradient nu u'' = -8 nu Umax / (LY - 2)^2, p = 0 at the outlet, so the
    density falls linearly from 1 + 24 nu Umax (LX - 1) / (LY - 2)^2 to 1."""
    specs = ws.c5_specs(2, seed=1)
    H, W = 24, 96
    f = np.zeros((2, H, W, 3), np.float32)
    for c, sp in enumerate(specs):
        Y = np.arange(H)[:, None] * (sp.LY - 1.0) / (H - 1)
        f[c, :, :, 1] = np.clip(4 * sp.U * (sp.LY - 1.5 - Y) * (Y - 0.5) / (sp.LY - 2.0) ** 2, 0, None)
        sp.CX = 1e6  # the body out of the channel
    g = ws.poisson_density(f, specs)
    for c, sp in enumerate(specs):
        nu = sp.U * 2 * sp.R / sp.Re
        drop = 24 * nu * sp.U / (sp.LY - 2.0) ** 2 * (sp.LX - 1)
        X = np.arange(W) * (sp.LX - 1.0) / (W - 1)
        want = 1 + drop * (1 - X / (sp.LX - 1.0))
        assert np.allclose(g[c, :, :, 0], want[None, :], rtol=0, atol=1e-6 * drop + 1e-7)
        assert np.array_equal(g[c, :, :, 1:], f[c, :, :, 1:])


def test_harmonic_fit_recovers_mean_and_two_harmonics():
    rng = np.random.default_rng(4)
    coef = rng.normal(size=(6, 10, 3 * ws.NHARM)).astype(np.float32)
    ph = rng.uniform(0, 2 * np.pi, 16)
    snaps = ws.harmonic_eval(coef, ph)
    assert snaps.shape == (16, 6, 10, 3)
    # the snapshot at phase p is mean + a1 cos p + b1 sin p + a2 cos 2p + b2 sin 2p
    c = coef.reshape(6, 10, ws.NHARM, 3)
    p = ph[3]
    want = c[:, :, 0] + c[:, :, 1] * np.cos(p) + c[:, :, 2] * np.sin(p) + c[:, :, 3] * np.cos(2 * p) \
        + c[:, :, 4] * np.sin(2 * p)
    assert np.allclose(snaps[3], want, atol=1e-5)
    assert np.allclose(ws.harmonic_fit(snaps, ph), coef, atol=1e-4)
    noisy = snaps + rng.normal(0, 1e-3, snaps.shape)
    assert np.abs(ws.harmonic_fit(noisy, ph) - coef).max() < 5e-3
    assert ws.static_inputs(ws.c5_specs(2, seed=1), 16, 32).shape == (2, 16, 32, ws.STATIC_CH)


def test_case_set_configuration():
    """Training and held-out cases: D in {16, 24} finest cells (8 or 12 coarse), blockage
    D/H ~ U[0.1, 0.3] (H in whole coarse cells), Re ~ U[80, 200], the cylinder 0.25 coarse
    cells above the channel's centre line; the two seeded sets do not share a case."""
    tr, te = ws.c5_specs(256, seed=101), ws.c5_specs(64, seed=202)
    for s in tr + te:
        dc = int(round(2.0 * s.R))
        assert dc in (8, 12)
        assert 80.0 <= s.Re <= 200.0
        assert 0.1 - 0.01 <= dc / (s.LY - 1) <= 0.3 + 0.01
        assert s.CY - (s.LY - 1) // 2 == 0.25
    assert {int(round(4.0 * s.R)) for s in tr} == {16, 24}
    key = lambda s: (s.LX, s.LY, round(s.Re, 9))
    assert not ({key(s) for s in tr} & {key(s) for s in te})


def test_case_sets_of_the_full_range():
    """The pipeline's default sets (Re ~ U[40, 200], training and held-out) cover the steady
    range below the onset of shedding and do not share a case."""
    tr, te = ws.c5_specs(256, seed=101, re_min=40.0), ws.c5_specs(64, seed=202, re_min=40.0)
    for s in tr + te:
        assert 40.0 <= s.Re <= 200.0
    assert sum(s.Re < 60.0 for s in te) >= 3
    key = lambda s: (s.LX, s.LY, round(s.Re, 9))
    assert not ({key(s) for s in tr} & {key(s) for s in te})


def test_convergence_rules():
    """Shedding cases are judged by the Cl amplitude, steady ones by Cd within 0.5 %; a warm
    case whose residual after the probe exceeds the threshold is charged probe + cold steps."""
    t = np.arange(6000, dtype=float)
    shed = 0.3 * (1 - np.exp(-t / 800.0)) * np.sin(2 * np.pi * t / 100.0)
    steady_cd = 1.8 + 0.4 * np.exp(-t / 500.0)
    cl_c = np.stack([shed, np.zeros_like(t)])
    cd_c = np.stack([np.full_like(t, 1.3), steady_cd])
    cl_w = np.stack([0.3 * np.sin(2 * np.pi * t / 100.0), np.zeros_like(t)])
    cd_w = np.stack([np.full_like(t, 1.3), 1.8 + 0.4 * np.exp(-t / 100.0)])
    tc, tw, fb = ws.convergence(cl_c, cl_w, cd_w, cd_c=cd_c, res0=np.array([0.001, 0.001]), threshold=0.02,
                                probe=20)
    assert fb == 0 and tc[0] > tw[0] > 0 and tc[1] > tw[1] > 0
    # steady: the first step of |Cd - final| <= 0.5 % of final
    assert abs(tc[1] - 500.0 * np.log(0.4 / (0.005 * steady_cd[-1]))) <= 2
    tc2, tw2, fb2 = ws.convergence(cl_c, cl_w, cd_w, cd_c=cd_c, res0=np.array([0.5, np.nan]), threshold=0.02,
                                   probe=20)
    assert fb2 == 2 and list(tw2) == [20 + tc2[0], 20 + tc2[1]]
