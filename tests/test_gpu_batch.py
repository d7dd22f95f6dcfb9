"""Batched independent cases (mbl_batch_*): a batch of uneven two-level cases
reproduces the same cases run one at a time through evolution()/getDrag_block,
and the autotuner returns a valid launch shape."""
import numpy as np
import pytest

import mbl
from mbl import abi

pytestmark = pytest.mark.gpu


def single_geometry(lib, sp):
    r = sp.region
    return mbl.Geometry.channel(lib, sp.LX, sp.LY, [(r.xstart, r.xend, r.ystart, r.yend)], R=sp.R, CX=sp.CX,
                                CY=sp.CY, Re=sp.Re, U=sp.U)


def test_batch_equals_individual_runs(lib64):
    specs = [mbl.case_spec(5, 0.2, 60.0), mbl.case_spec(8, 0.25, 100.0), mbl.case_spec(6, 0.15, 45.0)]
    steps = 30
    b = mbl.Batch(lib64, specs)
    b.step(steps)
    cd_b, cl_b = b.forces(steps)
    b.close()
    for c, sp in enumerate(specs):
        sim = mbl.MultiBlockSim(lib64, single_geometry(lib64, sp), U0=(sp.U, 0.0), init_mode=abi.INIT_PROFILE)
        sim.step(steps)
        _, drag = sim.series()
        sim.close()
        assert np.max(np.abs(cd_b[c] - drag["Cd"])) <= 1e-12 * np.max(np.abs(drag["Cd"]))
        assert np.max(np.abs(cl_b[c] - drag["Cl"])) <= 1e-12 * max(1e-9, np.max(np.abs(drag["Cl"])))


def test_c4_specs_deterministic():
    a, b = mbl.c4_specs(16, seed=7), mbl.c4_specs(16, seed=7)
    assert [(s.LX, s.LY, s.Re) for s in a] == [(s.LX, s.LY, s.Re) for s in b]
    for s in a:
        D = 2 * s.R
        assert 40 <= s.Re <= 200 and 2 * D in (10, 16, 24, 32, 40)
        assert 0.09 < D / (s.LY - 1) < 0.31


def test_autotune_and_persistent_grid(lib32):
    specs = mbl.c4_specs(12, seed=3)
    b = mbl.Batch(lib32, specs)
    ref = mbl.Batch(lib32, specs)
    t = b.autotune(trial_steps=3)
    assert t["vx"] in (1, 2, 4) and t["tpb"] in (64, 128, 256) and t["ms_per_step"] > 0
    # the tuned batch has advanced 36 * 4 steps; bring the reference to the same step count
    ref.step(3 * 3 * 4 * 4)
    b.sync()
    cd_a, _ = b.forces(1)
    cd_r, _ = ref.forces(1)
    assert np.allclose(cd_a, cd_r, rtol=1e-5, atol=1e-7)
    b.close()
    ref.close()


def oracle_for(lib, sp):
    import oracle
    geom = single_geometry(lib, sp)
    g = geom.g
    cfg = mbl.default_config(lib, LEVEL=g.LEVEL, number_line2=g.number_line2, LX=g.LX, LY=g.LY, R=g.R, CX=sp.CX,
                             CY=sp.CY, U0=(sp.U, 0.0), bc_mode=abi.BC_CHANNEL, init_mode=abi.INIT_PROFILE,
                             write_files=0, number_area=g.number_area)
    for i in range(g.LEVEL):
        cfg.mi[i] = g.mi[i]
    return oracle.OracleSim(g, cfg)


def uneven_specs():
    specs = [mbl.case_spec(5, 0.2, 60.0), mbl.case_spec(8, 0.25, 150.0), mbl.case_spec(6, 0.12, 45.0),
             mbl.case_spec(4, 0.3, 100.0)]
    specs[1].CY += 0.25
    return specs


@pytest.mark.parametrize("persistent", [False, True])
def test_batch_matches_oracle_fp64(lib64, persistent):
    """Four uneven cases in one batch (FP64 build) against the FP64 oracle, each case on its
    own: every plane of every block to 1e-12 and the drag series to 1e-10 after 40 coarse
    steps (the batch's refinement cycle is the reference's evolution.h:2-65 per case), for the
    captured-graph path and for the one-launch path (one workgroup per case)."""
    specs = uneven_specs()
    steps = 40
    b = mbl.Batch(lib64, specs, persistent=persistent)
    b.step(steps)
    cd_b, cl_b = b.forces(steps)
    for c, sp in enumerate(specs):
        o = oracle_for(lib64, sp)
        o.run(steps)
        for q in range(5):
            for fld in ("rho", "ux", "uy"):
                g, r = b.read(c, q, fld), o.get(q, fld)
                assert np.max(np.abs(g - r)) < 1e-12, (c, q, fld)
            for k in range(9):
                g, r = b.read(c, q, f"F{k}"), o.get(q, ("F", k))
                assert np.max(np.abs(g - r)) < 1e-12, (c, q, k)
        scale = max(1.0, np.max(np.abs(o.series(3))))
        assert np.max(np.abs(cd_b[c] - o.series(3))) / scale < 1e-10
        assert np.max(np.abs(cl_b[c] - o.series(4))) / scale < 1e-10
        o.close()
    b.close()


def batch_state(b, ncases):
    out = []
    for c in range(ncases):
        for q in range(5):
            for fld in ["rho", "ux", "uy"] + [f"F{k}" for k in range(9)]:
                out.append(b.read(c, q, fld))
    return out


@pytest.mark.parametrize("fp64", [True, False])
def test_persistent_equals_graph(lib64, lib32, fp64):
    """The one-launch path (mbl_batch_set_persistent: one workgroup per case runs its refinement
    cycles in the reference order) against the captured-graph path on the same uneven cases:
    every plane and macro field and every force, bit for bit, after step() calls of 1, 7, 5 and
    13 coarse steps (launches that start at different steps of the phase period), then with the
    two batches' paths swapped for 4 more steps."""
    lib = lib64 if fp64 else lib32
    specs = uneven_specs()
    g = mbl.Batch(lib, specs)
    p = mbl.Batch(lib, specs, persistent=True)
    total = 0
    for n in (1, 7, 5, 13):
        g.step(n)
        p.step(n)
        total += n
    for x, y in zip(batch_state(g, len(specs)), batch_state(p, len(specs))):
        assert np.array_equal(x, y)
    cg, lg = g.forces(total)
    cp, lp = p.forces(total)
    assert cg.shape == (len(specs), total)
    assert np.array_equal(cg, cp) and np.array_equal(lg, lp)
    g.set_persistent(True)
    p.set_persistent(False)
    g.step(4)
    p.step(4)
    for x, y in zip(batch_state(g, len(specs)), batch_state(p, len(specs))):
        assert np.array_equal(x, y)
    cg, lg = g.forces(total + 4)
    cp, lp = p.forces(total + 4)
    assert np.array_equal(cg, cp) and np.array_equal(lg, lp)
    g.close()
    p.close()


def test_persistent_equals_graph_many_lines(lib32):
    """32 cases of the sweep (128 spline lines: the graph's spline launches spread a line over 3
    workgroups, the one-launch path solves it in one) give the same bits on both paths."""
    specs = mbl.c4_specs(32, seed=7)
    g = mbl.Batch(lib32, specs)
    p = mbl.Batch(lib32, specs, persistent=True)
    g.step(9)
    p.step(9)
    for x, y in zip(batch_state(g, len(specs)), batch_state(p, len(specs))):
        assert np.array_equal(x, y)
    cg, lg = g.forces(9)
    cp, lp = p.forces(9)
    assert np.array_equal(cg, cp) and np.array_equal(lg, lp)
    g.close()
    p.close()


def test_batch_residual_matches_error_l2(lib64):
    """The per-case residual of a batch (mbl_batch_residual) equals error_L2 of the same case
    run alone, step after step (FP64)."""
    specs = [mbl.case_spec(5, 0.2, 60.0), mbl.case_spec(7, 0.15, 120.0)]
    b = mbl.Batch(lib64, specs)
    res = []
    for _ in range(6):
        b.step(1)
        res.append(b.residual())
    res = np.array(res)
    b.close()
    for c, sp in enumerate(specs):
        sim = mbl.MultiBlockSim(lib64, single_geometry(lib64, sp), U0=(sp.U, 0.0), init_mode=abi.INIT_PROFILE)
        sim.step(6)
        l2, _ = sim.series()
        sim.close()
        assert np.max(np.abs(res[:, c] - l2["residual"]) / np.abs(l2["residual"])) < 1e-12


@pytest.mark.parametrize("persistent", [False, True])
def test_batch_256_cases(lib32, persistent):
    """The 256-case sweep (seeded Re ~ U[40, 200], D in {10 .. 40} finest cells, blockage
    ~ U[0.1, 0.3]) at reduced steps: every case's forces finite; its mass (mean density over
    its channel, each node weighted by its cell area 1/m^2, the overlaps of the blocks counted
    once per block) within 2 % of the initial 1 through the inlet/outlet exchange (measured
    1.0 %); no node's
    density more than 15 % from 1 (measured 7.1 %: the pressure waves of the start from rest);
    and its residual finite. Captured-graph and one-launch paths."""
    specs = mbl.c4_specs(256)
    b = mbl.Batch(lib32, specs, persistent=persistent)
    b.step(300)
    cd, cl = b.forces(300)
    assert cd.shape == (256, 300) and np.all(np.isfinite(cd)) and np.all(np.isfinite(cl))
    r = b.residual()
    assert np.all(np.isfinite(r)) and np.all(r > 0)
    worst_node, worst_mass = 0.0, 0.0
    for c in range(256):
        mass = area = 0.0
        for q in range(5):
            a, _ = b.block(c, q)
            rho = b.read(c, q, "rho").astype(np.float64)
            assert np.all(np.isfinite(rho)), (c, q)
            worst_node = max(worst_node, float(np.max(np.abs(rho - 1.0))))
            w = 1.0 / (a.m * a.m)
            mass += w * float(rho.sum())
            area += w * rho.size
        worst_mass = max(worst_mass, abs(mass / area - 1.0))
    print("largest |rho - 1|", worst_node, "largest |mean rho - 1| of a case", worst_mass)
    assert worst_node < 0.15
    assert worst_mass < 0.02
