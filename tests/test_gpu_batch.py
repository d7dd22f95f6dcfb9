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
