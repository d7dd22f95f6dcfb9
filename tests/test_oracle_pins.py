"""Checks that pin three parts of the FP64 oracle to what the mathematics fixes, so that a
plausible mistake in the restatement (and therefore in the HIP path, which the GPU tests
hold to the oracle) fails here:

* the non-equilibrium rescale of the coarse/fine transfers, (tau_f - 1)/(r (tau_c - 1))
  coarse -> fine (multi/FH/assit_multi.h:176) and r (tau_c - 1)/(tau_f - 1) fine -> coarse
  (:304): a Chapman-Enskog shear state of one level must become the Chapman-Enskog state of
  the other (the gradient per lattice unit scales with the level's time step);
* the momentum-exchange sum f[k][x_s] e_k - f[opp k][x_f] e_opp (lbm/force_block.h:155-156):
  in a periodic box with an obstacle and a uniform body force (the forcing term of
  multi/FH/evolution_block.h:102-110), the steady force on the obstacle equals the momentum
  the force puts into the fluid per step;
* the time semantics of the refinement cycle (evolution.h:9-63): with every block's state
  driven as a known function of time, the fine edge a fine substep streams from is the
  coarse solution at that substep's time (ring slot order, rotation after the spline, the
  Lagrange weights at x = i / mi), and the fine edge's rho, u are those of the coarse state
  at t_{n+1} (assit_multi.h:172-174).

CPU only (the oracle is test infrastructure); the library is loaded for its host geometry
builder and config defaults only.
"""
import ctypes as C

import numpy as np
import pytest

import mbl
from mbl import abi
import oracle

EX = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])
EY = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])
W = np.array([4 / 9] + [1 / 9] * 4 + [1 / 36] * 4)
OPP = [0, 3, 4, 1, 2, 7, 8, 5, 6]


def feq(rho, ux, uy):
    """[9, ...] equilibrium of lbm/assit_func.h:2-9 (numpy)."""
    eu = EX.reshape(9, *([1] * np.ndim(ux))) * ux + EY.reshape(9, *([1] * np.ndim(ux))) * uy
    w = W.reshape(9, *([1] * np.ndim(ux)))
    return w * rho * (1.0 + 3.0 * eu + 4.5 * eu * eu - 1.5 * (ux * ux + uy * uy))


def oracle_sim(lib, geom, **kw):
    g = geom.g
    cfg = mbl.default_config(lib, LEVEL=g.LEVEL, number_line2=g.number_line2, LX=g.LX, LY=g.LY, R=g.R,
                             CX=geom.CX, CY=geom.CY, bc_mode=kw.get("bc_mode", geom.bc_mode),
                             init_mode=kw.get("init_mode", geom.init_mode), write_files=0,
                             number_area=g.number_area)
    for i in range(g.LEVEL):
        cfg.mi[i] = g.mi[i]
    return oracle.OracleSim(g, cfg)


def coords(o, a):
    ar = o.areas[a]
    ny, nx = o.shape(a)
    j, i = np.mgrid[0:ny, 0:nx]
    return ar.base_x + i / ar.m, ar.base_y + j / ar.m


def line_nodes(P, nx):
    """Plane indices of the nodes of line P (vertical when xstart == xend)."""
    n = (P.yend - P.ystart + 1) if P.xstart == P.xend else (P.xend - P.xstart + 1)
    ip = np.arange(n)
    if P.xstart == P.xend:
        return P.xstart + (P.ystart + ip) * nx
    return (P.xstart + ip) + P.ystart * nx


# ---------------------------------------------------------------------------
# non-equilibrium rescale
# ---------------------------------------------------------------------------
def ce_state(X, Y, tau, m, gamma, vertical):
    """Chapman-Enskog state of a steady simple shear with rho = 1: u = (gamma Y, 0) (or
    (0, gamma X) when `vertical`), gamma per coarse length. Post-collision storage on a level
    of refinement m: F = feq - (tau - 1) w_k rho / cs^2 (e_k e_k - cs^2 I) : grad u, with the
    velocity gradient per lattice unit of that level, gamma / m."""
    ux, uy = (np.zeros_like(X), gamma * X) if vertical else (gamma * Y, np.zeros_like(Y))
    g = gamma / m
    # (e e - I/3) : grad u for a pure shear (du_x/dy or du_y/dx): 2 e_x e_y g / 2 ... the
    # trace part vanishes, and both shears give e_x e_y g
    qg = (EX * EY).reshape(9, 1, 1) * g
    F = feq(1.0, ux, uy) - (tau - 1.0) * W.reshape(9, 1, 1) * 3.0 * qg
    return F, ux, uy


@pytest.fixture(scope="module")
def two_level(lib32):
    # tau_c = 0.575, tau_f = 0.65, refinement ratio 2
    return mbl.Geometry.channel(lib32, 61, 25, [(10, 30, 6, 18)], R=2.5, CX=18.0, CY=12.0, Re=20.0)


@pytest.mark.parametrize("vertical", [False, True])
def test_coarse_to_fine_rescale_gives_fine_chapman_enskog_state(lib32, two_level, vertical):
    gamma = 2e-3
    o = oracle_sim(lib32, two_level)
    for a in range(o.na):
        X, Y = coords(o, a)
        F, ux, uy = ce_state(X, Y, o.areas[a].tau, o.areas[a].m, gamma, vertical)
        o.set(a, "rho", np.ones_like(X))
        o.set(a, "ux", ux)
        o.set(a, "uy", uy)
        for k in range(9):
            o.set(a, ("F", k), F[k])
    checked = 0
    for b in range(o.lib.orc_sim_num_lines(o.p, 0)):
        L = o.line(0, b)
        if (L.line_F.xstart == L.line_F.xend) != vertical:
            continue  # along the line the shear state is constant, so the spline is exact
        o.spline_line(b)
        slot = o.markers(b)[0]
        af = L.areaf_n
        Xf, Yf = coords(o, af)
        idx = line_nodes(L.line_F, o.shape(af)[1])
        Fexp, uxe, uye = ce_state(Xf.ravel()[idx][None, :], Yf.ravel()[idx][None, :], o.areas[af].tau,
                                  o.areas[af].m, gamma, vertical)
        for k in range(9):
            got = o.ring(b, slot, k, L.Nf)
            assert np.max(np.abs(got - Fexp[k, 0])) < 1e-15, (b, k)
        # the fine edge's rho and u are the interpolated coarse macros
        assert np.max(np.abs(o.get(af, "ux").ravel()[idx] - uxe[0])) < 1e-16
        assert np.max(np.abs(o.get(af, "uy").ravel()[idx] - uye[0])) < 1e-16
        checked += 1
    assert checked == 2
    o.close()


def test_fine_to_coarse_rescale_gives_coarse_chapman_enskog_state(lib32, two_level):
    gamma = 2e-3
    o = oracle_sim(lib32, two_level)
    fine = [a for a in range(o.na) if o.areas[a].m == 2]
    for a in range(o.na):
        X, Y = coords(o, a)
        if a in fine:
            F, ux, uy = ce_state(X, Y, o.areas[a].tau, o.areas[a].m, gamma, False)
        else:  # coarse blocks start from a marker value, so the written nodes show
            F, ux, uy = np.full((9,) + X.shape, -7.0), np.full(X.shape, -7.0), np.full(X.shape, -7.0)
        o.set(a, "rho", np.ones_like(X))
        o.set(a, "ux", ux)
        o.set(a, "uy", uy)
        for k in range(9):
            o.set(a, ("F", k), F[k])
    nc = 0
    for b in range(o.lib.orc_sim_num_lines(o.p, 0)):
        o.transfer_line(b, 1)
        nc += o.line(2, b).Nc
    written = 0
    for a in range(o.na):
        if a in fine:
            continue
        X, Y = coords(o, a)
        Fexp, uxe, _ = ce_state(X, Y, o.areas[a].tau, o.areas[a].m, gamma, False)
        ux = o.get(a, "ux")
        mask = ux != -7.0
        written += int(mask.sum())
        assert np.max(np.abs(ux[mask] - uxe[mask])) < 1e-16
        for k in range(9):
            got = o.get(a, ("F", k))
            assert np.max(np.abs(got[mask] - Fexp[k][mask])) < 1e-15, (a, k)
    # every coarse ghost node of every line was written once (corners shared by two lines)
    assert 0 < written <= nc
    o.close()


def test_rescale_round_trip_and_equilibrium_preserved(lib32, two_level):
    """Coarse -> fine -> coarse returns the coarse state (para_cf * para_fc = 1), and an
    equilibrium state crosses the interface unchanged."""
    o = oracle_sim(lib32, two_level)
    rng = np.random.default_rng(3)
    for a in range(o.na):
        X, _ = coords(o, a)
        rho = 1.0 + 0.01 * rng.standard_normal(X.shape)
        ux = 0.05 + 0.01 * rng.standard_normal(X.shape)
        uy = 0.01 * rng.standard_normal(X.shape)
        F = feq(rho, ux, uy) + (1e-4 * rng.standard_normal((9,) + X.shape) if a == 4 else 0.0)
        o.set(a, "rho", rho)
        o.set(a, "ux", ux)
        o.set(a, "uy", uy)
        for k in range(9):
            o.set(a, ("F", k), F[k])
    # fine -> coarse on the equilibrium-plus-noise fine block, then compare the coarse ghost's
    # non-equilibrium part with the fine one scaled by r (tau_c - 1) / (tau_f - 1), and check
    # that rescaling it back by (tau_f - 1) / (r (tau_c - 1)) restores the fine values
    tc, tf = o.areas[0].tau, o.areas[4].tau
    for b in range(o.lib.orc_sim_num_lines(o.p, 0)):
        L = o.line(2, b)
        o.transfer_line(b, 1)
        nxc, nxf = o.shape(L.areac_n)[1], o.shape(L.areaf_n)[1]
        ic = line_nodes(L.line_C, nxc)
        ifn = line_nodes(L.line_F, nxf)[::2]
        rf, uf, vf = (o.get(L.areaf_n, f).ravel()[ifn] for f in ("rho", "ux", "uy"))
        fe = feq(rf, uf, vf)
        for k in range(9):
            Ff = o.get(L.areaf_n, ("F", k)).ravel()[ifn]
            Fc = o.get(L.areac_n, ("F", k)).ravel()[ic]
            back = fe[k] + (tf - 1.0) / (2 * (tc - 1.0)) * (Fc - fe[k])
            assert np.max(np.abs(back - Ff)) < 1e-15
    # an equilibrium coarse block crosses coarse -> fine unchanged
    for a in range(4):
        X, _ = coords(o, a)
        rho, ux, uy = 1.0 + 0 * X, 0.04 + 0 * X, -0.01 + 0 * X
        F = feq(rho, ux, uy)
        o.set(a, "rho", rho)
        o.set(a, "ux", ux)
        o.set(a, "uy", uy)
        for k in range(9):
            o.set(a, ("F", k), F[k])
    fe = feq(1.0, 0.04, -0.01)
    for b in range(o.lib.orc_sim_num_lines(o.p, 0)):
        o.spline_line(b)
        L = o.line(0, b)
        for k in range(9):
            assert np.max(np.abs(o.ring(b, o.markers(b)[0], k, L.Nf) - fe[k])) < 1e-16
    o.close()


# ---------------------------------------------------------------------------
# momentum exchange
# ---------------------------------------------------------------------------
def test_momentum_exchange_balances_body_force(lib32):
    """Periodic 41 x 31 box, cylinder R = 4, uniform body force g on the fluid nodes: at steady
    state the momentum-exchange force on the cylinder equals N_fluid (1 - 1/(2 tau)) g (the
    forcing term adds (1 - 1/(2 tau)) g per node and step, evolution_block.h:102-110).
    Summing the same links with the solid and fluid nodes swapped, or with the returning
    population's sign flipped, misses the balance by orders of magnitude more than the
    tolerance."""
    geom = mbl.Geometry.channel(lib32, 41, 31, [], R=4.0, CX=20.0, CY=15.3, Re=4.0)
    o = oracle_sim(lib32, geom, bc_mode=abi.BC_PERIODIC, init_mode=abi.INIT_REST)
    flag = o.get(0, "solid")
    fluid = flag == 0
    nf = int(fluid.sum())
    gx, gy = 2e-6, -1e-6
    o.set(0, "forcex", np.where(fluid, gx, 0.0))
    o.set(0, "forcey", np.where(fluid, gy, 0.0))
    tau = o.areas[0].tau
    o.run(15000)  # balance within 3e-6 after 6,000 steps, 4e-9 after 9,000, 1e-12 after 12,000
    fx, fy = o.series(1)[-1], o.series(2)[-1]
    ex, ey = nf * (1 - 0.5 / tau) * gx, nf * (1 - 0.5 / tau) * gy
    assert abs(fx - ex) / abs(ex) < 1e-8, (fx, ex)
    assert abs(fy - ey) / abs(ey) < 1e-8, (fy, ey)
    # the same sum restated from the post-stream planes, and two index mistakes
    f = [o.get(0, ("f", k)) for k in range(9)]
    good = np.zeros(2)
    swapped = np.zeros(2)
    flipped = np.zeros(2)
    for (i, j, k) in o.links():
        kd, jf, i_f = OPP[k], j - EY[k], i - EX[k]
        e_k, e_d = np.array([EX[k], EY[k]]), np.array([EX[kd], EY[kd]])
        good += f[k][j, i] * e_k - f[kd][jf, i_f] * e_d
        swapped += f[k][jf, i_f] * e_k - f[kd][j, i] * e_d
        flipped += f[k][j, i] * e_k + f[kd][jf, i_f] * e_d
    assert np.allclose(good, [fx, fy], rtol=1e-12, atol=0)
    for bad in (swapped, flipped):
        assert abs(bad[0] - ex) / abs(ex) > 1e-2
    o.close()


# ---------------------------------------------------------------------------
# refinement-cycle time semantics
# ---------------------------------------------------------------------------
def driven(t):
    """Spatially uniform state at time t (coarse steps): rho = 1, u linear in t, F = feq, so
    every population is quadratic in t and the 3-point Lagrange interpolation is exact."""
    ux, uy = 0.05 + 0.004 * t, -0.02 + 0.003 * t
    return ux, uy, feq(1.0, ux, uy)


def test_refinement_cycle_time_semantics(lib32):
    """Three levels (ratio 2 each, mi = 1, 2, 2). Every block's stream-collide is replaced by
    'record the fine edge, then set the block's updated nodes (all but the C_BOUNDARY ghost and
    edge nodes, which the stream-collide carries over) to the driven state at the end of this
    substep'. From the third coarse step on (the ring holds three driven slots), the edge a
    substep starts from must be the driven state at that substep's start time, and the edge's
    rho, u the coarse state at the end of the coarse substep that spawned it."""
    geom = mbl.Geometry.channel(lib32, 81, 33, [(12, 44, 6, 26), (16, 36, 10, 22)], R=2.0, CX=22.0, CY=16.0,
                                Re=20.0)
    o = oracle_sim(lib32, geom)
    nl = o.lib.orc_sim_num_lines(o.p, 0)
    lines = [o.line(0, b) for b in range(nl)]
    keep = {a: o.get(a, "solid") == 9 for a in range(o.na)}  # C_BOUNDARY
    state = {"i1": 0}
    errs = []  # (what, level, substep, coarse step, error); asserted after the run (a ctypes
    checks = {1: 0, 2: 0}  # callback cannot raise into the caller)

    def hook(a, level, i):
        n = o.lib.orc_sim_steps_done(o.p)
        if level == 1:
            state["i1"] = i
        if level == 0:
            t0, t1 = n, n + 1.0
        elif level == 1:
            t0, t1 = n + i / 2.0, n + (i + 1) / 2.0
        else:
            base = n + state["i1"] / 2.0
            t0, t1 = base + i / 4.0, base + (i + 1) / 4.0
        if level > 0 and n >= 3:
            tp = n + 1.0 if level == 1 else n + (state["i1"] + 1) / 2.0
            uxp, uyp, _ = driven(tp)
            _, _, F_e = driven(t0)
            nx = o.shape(a)[1]
            for L in lines:
                if L.areaf_n != a:
                    continue
                idx = line_nodes(L.line_F, nx)
                checks[level] += 1
                for k in range(9):
                    errs.append(("F", level, i, n, float(np.max(np.abs(o.get(a, ("F", k)).ravel()[idx] - F_e[k])))))
                errs.append(("ux", level, i, n, float(np.max(np.abs(o.get(a, "ux").ravel()[idx] - uxp)))))
                errs.append(("uy", level, i, n, float(np.max(np.abs(o.get(a, "uy").ravel()[idx] - uyp)))))
        ux, uy, F = driven(t1)
        kp = keep[a]

        def put(field, v):
            o.set(a, field, np.where(kp, o.get(a, field), v))

        put("rho", 1.0)
        put("ux", ux)
        put("uy", uy)
        for k in range(9):
            put(("F", k), F[k])

    o.set_block_hook(hook)
    o.evolve(6)
    o.set_block_hook(None)
    # 3 checked coarse steps: level 1, 8 lines x 2 substeps; level 2, 4 lines x 4 substeps
    assert checks == {1: 48, 2: 48}
    worst = max(errs, key=lambda e: e[4])
    assert worst[4] < 1e-14, worst
    o.close()
