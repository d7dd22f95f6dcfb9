"""Pins the FP64 oracle (oracle/oracle.c) before it is trusted as the checker.

The reference ships no tests, golden vectors or fixtures, and cannot be built
here (CUDA + cuSPARSE), so each oracle function is checked against what the
mathematics fixes: closed-form moments, an independent dense/banded solve,
exactness properties, brute-force counts, conservation laws, the analytic
Poiseuille profile, and the layout fixture tests/golden/appendix_b.json.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest
import scipy.linalg

import mbl
from mbl import abi
import oracle

GOLD = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "appendix_b.json")))
EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)


def dp(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


# ---- lbm/assit_func.h: feq and Macros ------------------------------------
def test_feq_moments(olib):
    rng = np.random.default_rng(1)
    for _ in range(20):
        rho, ux, uy = 0.8 + 0.4 * rng.random(), 0.2 * rng.standard_normal(), 0.2 * rng.standard_normal()
        f = np.array([olib.orc_feq(k, rho, ux, uy) for k in range(9)])
        assert f.sum() == pytest.approx(rho, rel=1e-14)
        assert (f * EX).sum() == pytest.approx(rho * ux, abs=1e-15)
        assert (f * EY).sum() == pytest.approx(rho * uy, abs=1e-15)
        # second moments: rho (u_a u_b + delta_ab / 3)
        assert (f * EX * EX).sum() == pytest.approx(rho * (ux * ux + 1 / 3), rel=1e-13)
        assert (f * EY * EY).sum() == pytest.approx(rho * (uy * uy + 1 / 3), rel=1e-13)
        assert (f * EX * EY).sum() == pytest.approx(rho * ux * uy, abs=1e-15)


def test_macros_of_feq(olib):
    rho, ux, uy = 1.07, 0.031, -0.052
    f = np.array([olib.orc_feq(k, rho, ux, uy) for k in range(9)])
    r, u, v = C.c_double(), C.c_double(), C.c_double()
    olib.orc_macros(C.byref(r), C.byref(u), C.byref(v), dp(f))
    assert r.value == pytest.approx(rho, rel=1e-15)
    assert u.value == pytest.approx(ux, rel=1e-14)
    assert v.value == pytest.approx(uy, rel=1e-14)


# ---- multi/init_multi.h + cusparseDgtsv2: tridiagonal solve ---------------
@pytest.mark.parametrize("N", [3, 4, 9, 97, 321, 641])
def test_tridiagonal_against_lapack(olib, N):
    l, m, u = np.zeros(N), np.zeros(N), np.zeros(N)
    olib.orc_tri_A(N, dp(l), dp(m), dp(u))
    # intended final state of init_tri_A: decoupled first and last rows
    assert l[0] == 0 and u[N - 1] == 0 and u[0] == 0 and l[1] == 0 and u[N - 2] == 0 and l[N - 1] == 0
    assert np.all(m == 2.0)
    rng = np.random.default_rng(N)
    B = rng.standard_normal((12, N))
    X = B.copy()
    olib.orc_gtsv(N, 12, dp(l), dp(m), dp(u), dp(X), N)
    ab = np.zeros((3, N))
    ab[0, 1:] = u[:-1]
    ab[1, :] = m
    ab[2, :-1] = l[1:]
    for c in range(12):
        ref = scipy.linalg.solve_banded((1, 1), ab, B[c])
        assert np.max(np.abs(X[c] - ref)) <= 1e-13 * max(1.0, np.max(np.abs(ref)))
        A = np.diag(m) + np.diag(u[:-1], 1) + np.diag(l[1:], -1)
        assert np.linalg.norm(A @ X[c] - B[c]) <= 1e-14 * np.linalg.norm(B[c]) * N


# ---- multi/FH/assit_multi.h: the as-coded spline ---------------------------
def test_spline_kat_as_coded_sign():
    kat = GOLD["spline_kat"]
    g = np.arange(9, dtype=float) ** 2
    M, out = oracle.spline_column(g, 2)
    assert np.allclose(M, kat["M"], atol=5e-5)
    assert np.allclose(out[1::2], kat["midpoints"], atol=5e-5)
    # independent derivation: interior rows solve tridiag(0.5, 2, 0.5) M = -3 * second difference
    n = 7
    A = 2 * np.eye(n) + 0.5 * np.eye(n, k=1) + 0.5 * np.eye(n, k=-1)
    rhs = -3.0 * (g[:-2] - 2 * g[1:-1] + g[2:])
    assert np.allclose(M[1:-1], np.linalg.solve(A, rhs), atol=1e-14)


def test_spline_midpoint_closed_form():
    """Midpoint = linear average - 0.375 * 0.1666667 * (M_{I-1} + M_I) for any data."""
    rng = np.random.default_rng(3)
    g = rng.standard_normal(41)
    M, out = oracle.spline_column(g, 2)
    c6 = 0.1666667
    mid = 0.5 * (g[:-1] + g[1:]) - 0.375 * c6 * (M[:-1] + M[1:])
    assert np.max(np.abs(out[1::2] - mid)) < 1e-14
    assert np.max(np.abs(out[0::2] - g)) < 1e-14  # coincident nodes are exact


def test_spline_linear_data_exact():
    g = 0.3 + 0.17 * np.arange(25)
    M, out = oracle.spline_column(g, 2)
    assert np.max(np.abs(M)) < 1e-13  # right-hand side is zero up to rounding
    xs = np.arange(out.size) / 2.0
    assert np.max(np.abs(out - (0.3 + 0.17 * xs))) < 1e-13


def test_spline3_s_endpoints(olib):
    # at xx = x_1 the cubic returns fi_1, at xx = x_1 + h it returns fi
    assert olib.orc_spline3_s(2.0, 2.0, 1.0, 0.7, 1.3, 0.2, -0.4) == pytest.approx(0.7, abs=1e-15)
    assert olib.orc_spline3_s(3.0, 2.0, 1.0, 0.7, 1.3, 0.2, -0.4) == pytest.approx(1.3, abs=1e-15)


# ---- temporal interpolation and the ring markers ---------------------------
def test_temporal_lagrange_exact_for_quadratics(olib):
    rng = np.random.default_rng(5)
    for _ in range(10):
        a, b, c = rng.standard_normal(3)
        q = lambda t: a + b * t + c * t * t  # noqa: E731
        for xm in (0.0, 0.25, 0.5, 0.75):
            v = olib.orc_temporal_value(xm, q(-1.0), q(0.0), q(1.0))
            assert v == pytest.approx(q(xm), abs=1e-14)
    assert olib.orc_temporal_value(0.0, 1.5, 2.5, 3.5) == 2.5


def tiny_channel(lib):
    return mbl.Geometry.channel(lib, 61, 25, [(10, 30, 6, 18)], R=2.5, CX=18.0, CY=12.0, Re=20.0)


def make_cfg(lib, geom, **kw):
    g = geom.g
    cfg = mbl.default_config(lib, LEVEL=g.LEVEL, number_line2=g.number_line2, LX=g.LX, LY=g.LY, R=g.R,
                             CX=geom.CX, CY=geom.CY, bc_mode=geom.bc_mode, init_mode=geom.init_mode, **kw)
    for i in range(g.LEVEL):
        cfg.mi[i] = g.mi[i]
    return cfg


def test_marker_rotation_sequence(lib32):
    geom = tiny_channel(lib32)
    o = oracle.OracleSim(geom.g, make_cfg(lib32, geom))
    seq = [o.markers(0)]
    for _ in range(3):
        o.evolve(1)
        seq.append(o.markers(0))
    assert seq == [[2, 1, 0], [1, 0, 2], [0, 2, 1], [2, 1, 0]]
    o.close()


# ---- geo/init_geo.h: the layout fixture ------------------------------------
@pytest.mark.parametrize("name", ["GEO_2_5", "GEO_4_9"])
def test_geometry_appendix_b(lib32, olib, name):
    G = mbl.Geometry.preset(lib32, name)
    g = G.g
    ad = (abi.P4 * g.number_area)()
    ar = (abi.AREA * g.number_area)()
    lc = (abi.Inf_Ff_store * g.number_line)()
    lf = (abi.Inf_Ff_store * g.number_line)()
    lcc = (abi.Inf_Ff_store * max(1, g.number_line2))()
    olib.orc_geo_init(C.byref(g), ad, ar, lc, lf, lcc)
    want = GOLD[name]
    got = [[a.m, a.nx, a.ny, round(a.tau, 6)] for a in ar]
    assert got == [[m, nx, ny, pytest.approx(t)] for m, nx, ny, t in want["areas"]]
    assert sum(a.nx * a.ny * a.m for a in ar) == want["node_updates_per_coarse_step"]
    if "area_define" in want:
        assert [[p.xstart, p.xend, p.ystart, p.yend] for p in ad] == want["area_define"]
    lines = [[l.areac_n, l.areaf_n, l.Nc, l.Nf] for l in lc]
    for L in want.get("cf_lines", []):
        assert L in lines
    for L in want.get("cf_lines_subset", []):
        assert L in lines
    if "cc_nc" in want:
        assert [l.Nc for l in lcc] == want["cc_nc"]


def test_c1_layout_fixture(lib32):
    geom = mbl.c1_geometry(lib32)
    o = oracle.OracleSim(geom.g, make_cfg(lib32, geom))
    want = GOLD["C1"]
    assert [[a.m, a.nx, a.ny, round(a.tau, 6)] for a in o.areas] == want["areas"]
    assert sum(a.nx * a.ny * a.m for a in o.areas) == want["node_updates_per_coarse_step"]
    assert [[o.line(0, b).areac_n, o.line(0, b).areaf_n, o.line(0, b).Nc, o.line(0, b).Nf]
            for b in range(4)] == want["cf_lines"]
    assert [o.line(1, c).Nc for c in range(8)] == want["cc_nc"]
    # brute-force link count from the flags: solid node, in-block fluid neighbour x - e_k
    solid = o.get(4, "solid").astype(int)
    ny, nx = solid.shape
    n = 0
    for j in range(ny):
        for i in range(nx):
            if solid[j, i] != abi.SOLID_SBB:
                continue
            for k in range(9):
                jd, id_ = j - EY[k], i - EX[k]
                if 0 <= id_ < nx and 0 <= jd < ny and solid[jd, id_] == abi.FLUID:
                    n += 1
    assert n == want["drag_links"] == len(o.links())
    o.close()


# ---- stream/collide invariants (multi/FH/evolution_block.h) ---------------
def test_periodic_equilibrium_fixed_point():
    o = oracle.OraclePeriodic(24, 16, tau=0.7, sigma=0.0)  # u = 0 everywhere
    F0 = np.stack([o.get(("F", k)) for k in range(9)])
    o.run(5)
    F1 = np.stack([o.get(("F", k)) for k in range(9)])
    assert np.max(np.abs(F1 - F0)) < 2e-16
    assert np.allclose(F0[:, 0, 0], W, atol=1e-16)


def test_periodic_conservation_and_streaming():
    o = oracle.OraclePeriodic(30, 20, tau=0.8)
    F0 = np.stack([o.get(("F", k)) for k in range(9)])
    o.run(1)
    f1 = np.stack([o.get(("f", k)) for k in range(9)])
    F1 = np.stack([o.get(("F", k)) for k in range(9)])
    # pull streaming: f_k(x) = F_k(x - e_k) on the periodic torus
    for k in range(9):
        assert np.array_equal(f1[k], np.roll(np.roll(F0[k], EY[k], axis=0), EX[k], axis=1))
    assert F1.sum() == pytest.approx(F0.sum(), rel=1e-14)
    assert (F1 * EX[:, None, None]).sum() == pytest.approx((F0 * EX[:, None, None]).sum(), abs=1e-12)
    # BGK relaxes towards the local equilibrium: F = f - (f - feq)/tau
    rho = f1.sum(0)
    ux, uy = (f1 * EX[:, None, None]).sum(0) / rho, (f1 * EY[:, None, None]).sum(0) / rho
    lib = oracle.load()
    for (j, i) in [(0, 0), (7, 11), (19, 29)]:
        for k in range(9):
            fe = lib.orc_feq(k, rho[j, i], ux[j, i], uy[j, i])
            assert F1[k, j, i] == pytest.approx(f1[k, j, i] - (f1[k, j, i] - fe) / 0.8, abs=1e-15)


def test_random_velocity_deterministic_and_normal(olib):
    ux, uy = C.c_double(), C.c_double()
    vals = []
    for j in range(64):
        for i in range(64):
            olib.orc_random_velocity(7, 0.01, i, j, C.byref(ux), C.byref(uy))
            vals.append((ux.value, uy.value))
    v = np.array(vals)
    assert abs(v.mean()) < 1e-3 * 1.0 and abs(v.std() - 0.01) < 5e-4
    olib.orc_random_velocity(7, 0.01, 5, 9, C.byref(ux), C.byref(uy))
    assert (ux.value, uy.value) == tuple(v[9 * 64 + 5])


def test_bounce_back_and_zero_drag_at_rest(lib32):
    """u = 0 equilibrium around a symmetric body: BB reversal and zero net force."""
    geom = tiny_channel(lib32)
    cfg = make_cfg(lib32, geom, rho0=1.0, U0=(0.1, 0.0))
    o = oracle.OracleSim(geom.g, cfg)
    d = o.diag()
    assert abs(d[1]) < 1e-15 and abs(d[2]) < 1e-15
    o.evolve(1)
    solid = o.get(4, "solid").astype(int)
    f = [o.get(4, ("f", k)) for k in range(9)]
    F = [o.get(4, ("F", k)) for k in range(9)]
    opp = [0, 3, 4, 1, 2, 7, 8, 5, 6]
    js, is_ = np.nonzero(solid == abi.SOLID_SBB)
    for k in range(9):
        assert np.array_equal(F[k][js, is_], f[opp[k]][js, is_])
    o.close()


def test_error_l2_closed_form(lib32):
    geom = tiny_channel(lib32)
    o = oracle.OracleSim(geom.g, make_cfg(lib32, geom))
    rng = np.random.default_rng(11)
    num = den = 0.0
    for a in range(o.na):
        shp = o.shape(a)
        u, v, pu, pv = (rng.standard_normal(shp) for _ in range(4))
        o.set(a, "ux", u)
        o.set(a, "uy", v)
        o.set(a, "upx", pu)
        o.set(a, "upy", pv)
        num += ((u - pu) ** 2 + (v - pv) ** 2).sum() / o.areas[a].m
        den += (u * u + v * v).sum() / o.areas[a].m
    d = o.diag()
    assert d[0] == pytest.approx(np.sqrt(num) / np.sqrt(den), rel=1e-13)
    assert np.array_equal(o.get(0, "upx"), o.get(0, "ux"))  # u_prev <- u
    o.close()


def test_channel_poiseuille_profile(lib32):
    """Walls (full-way bounce-back), parabolic velocity inlet and pressure outlet
    (both by non-equilibrium extrapolation) on one block. The developed flow
    has the analytic profile shape u(y) ~ (y - 0.5)(LY - 1.5 - y): at
    tau = 1/2 + sqrt(3/16) the bounce-back wall sits exactly half-way. Mass flux
    is the same through every cross-section and the outlet holds rho0."""
    tau = 0.5 + np.sqrt(3.0 / 16.0)
    nu = (tau - 0.5) / 3.0
    U, R = 0.02, 1.0
    Re = U * 2 * R / nu
    LX, LY = 121, 17
    geom = mbl.Geometry.channel(lib32, LX, LY, [], R=R, CX=-1000.0, CY=-1000.0, Re=Re, U=U)
    cfg = make_cfg(lib32, geom, U0=(U, 0.0))
    o = oracle.OracleSim(geom.g, cfg)
    assert o.areas[0].tau == pytest.approx(tau)
    o.evolve(6000)
    ux, uy, rho = o.get(0, "ux"), o.get(0, "uy"), o.get(0, "rho")
    y = np.arange(LY, dtype=float)
    shape = (LY - 1.5 - y[1:-1]) * (y[1:-1] - 0.5)
    for x in (30, 60, 90):
        col = ux[1:-1, x]
        fit = col @ shape / (shape @ shape)
        assert np.max(np.abs(col - fit * shape)) < 1e-4 * U
        assert np.max(np.abs(uy[1:-1, x])) < 1e-4 * U
    flux = (rho[1:-1, 1:-1] * ux[1:-1, 1:-1]).sum(axis=0)
    assert np.max(np.abs(flux - flux.mean())) < 1e-6 * flux.mean()
    # the inlet node carries the parabola; the flux through the first face matches its sum to
    # within the O(1) discrete error of the extrapolated inlet
    inflow = (4 * U * (LY - 1.5 - y[1:-1]) * (y[1:-1] - 0.5) / (LY - 2.0) ** 2).sum()
    assert flux.mean() == pytest.approx(inflow * rho[1:-1, 1].mean(), rel=1.5e-2)
    assert np.allclose(ux[1:-1, 0], 4 * U * (LY - 1.5 - y[1:-1]) * (y[1:-1] - 0.5) / (LY - 2.0) ** 2, atol=1e-17)
    assert np.allclose(rho[1:-1, -1], 1.0, atol=1e-15)
    o.close()


def test_fp32_emulation_gate():
    """The oracle is FP64 throughout: feq of values that differ below FP32
    resolution stay distinguishable (guards against a float build)."""
    lib = oracle.load()
    a = lib.orc_feq(1, 1.0, 1e-9, 0.0)
    b = lib.orc_feq(1, 1.0, 0.0, 0.0)
    assert a - b == pytest.approx(1 / 9 * 3e-9, rel=1e-6)
