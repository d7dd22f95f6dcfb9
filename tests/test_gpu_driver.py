"""The C host driver (driver/mbl_driver.c): the reference's main loop written
against include/mbl.h only. Run as a child process in a scratch directory, it
must write the reference's diagnostic files (error_L2.h:85-88 "%d\\t%g\\n",
force_block.h:212-215 "%d\\t%g\\t%g\\t\\n", one line per coarse step) and, for
the FP64 build, the same residual and forces as the Python host path through
the same C ABI."""
import os
import subprocess

import numpy as np
import pytest

import mbl

pytestmark = pytest.mark.gpu

BIN = os.path.join(mbl.PKG_DIR, "bin")


def read_dat(path, ncol):
    rows = []
    with open(path) as fh:
        for line in fh:
            parts = line.rstrip("\n").split("\t")
            if ncol == 2:
                assert len(parts) == 2, line
            else:
                assert len(parts) == 4 and parts[3] == "", line
            rows.append([float(x) for x in parts[:ncol]])
    return np.array(rows)


@pytest.mark.parametrize("exe", ["mbl_driver", "mbl_driver64"])
def test_driver_writes_reference_files(tmp_path, exe):
    steps = 12
    r = subprocess.run([os.path.join(BIN, exe), "GEO_2_5", str(steps), "6"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    l2 = read_dat(tmp_path / "error_L2.dat", 2)
    fd = read_dat(tmp_path / "fdrag_t.dat", 3)
    assert l2.shape == (steps, 2) and fd.shape == (steps, 3)
    assert np.array_equal(l2[:, 0], np.arange(steps)) and np.array_equal(fd[:, 0], np.arange(steps))
    assert np.all(np.isfinite(l2)) and np.all(np.isfinite(fd))
    assert (tmp_path / "time_t.dat").exists()


def test_driver_matches_python_host_fp64(tmp_path, lib64):
    steps = 8
    r = subprocess.run([os.path.join(BIN, "mbl_driver64"), "GEO_2_5", str(steps), "8"], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    l2 = read_dat(tmp_path / "error_L2.dat", 2)
    fd = read_dat(tmp_path / "fdrag_t.dat", 3)
    sim = mbl.MultiBlockSim(lib64, mbl.Geometry.preset(lib64, "GEO_2_5"))
    sim.step(steps)
    res, drag = sim.series()
    sim.close()
    # the files carry "%g" (6 significant digits)
    np.testing.assert_allclose(l2[:, 1], res["residual"], rtol=1e-5)
    np.testing.assert_allclose(fd[:, 1], drag["Cd"], rtol=1e-5, atol=1e-9)
    np.testing.assert_allclose(fd[:, 2], drag["Cl"], rtol=1e-5, atol=1e-9)


def oracle_series(geom, steps, **cfg_kw):
    """error_L2 residual, Cd and Cl of the FP64 oracle's restatement of Multi-block-LBM.cu:main."""
    import oracle
    lib = mbl.load_library(False)
    g = geom.g
    cfg = mbl.default_config(lib, LEVEL=g.LEVEL, number_line2=g.number_line2, LX=g.LX, LY=g.LY, R=g.R,
                             CX=geom.CX, CY=geom.CY, bc_mode=geom.bc_mode, init_mode=geom.init_mode,
                             write_files=0, number_area=g.number_area, **cfg_kw)
    for i in range(g.LEVEL):
        cfg.mi[i] = g.mi[i]
    o = oracle.OracleSim(g, cfg)
    o.run(steps)
    out = o.series(0), o.series(3), o.series(4)
    o.close()
    return out


@pytest.mark.parametrize("exe,layout,steps,fused", [("mbl_driver64", "GEO_2_5", 12, 0), ("mbl_driver64", "C1", 400, 0),
                                                    ("mbl_driver", "C1", 400, 0), ("mbl_driver64", "C1", 400, 1),
                                                    ("mbl_driver", "GEO_2_5", 12, 1)])
def test_driver_files_match_oracle(tmp_path, lib32, exe, layout, steps, fused):
    """The C driver's error_L2.dat and fdrag_t.dat (written with the reference's "%g", six
    significant digits) against the FP64 oracle's series of the same layout: every line
    within the format's rounding (5e-6 relative), the FP32 build's forces too (its residual, a
    difference of consecutive velocity fields, carries FP32 rounding of u and is compared at
    2e-3). fused: the driver's chunks as mbl_evolution_step calls."""
    r = subprocess.run([os.path.join(BIN, exe), layout, str(steps), str(steps), "0", str(fused)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    l2 = read_dat(tmp_path / "error_L2.dat", 2)
    fd = read_dat(tmp_path / "fdrag_t.dat", 3)
    geom = mbl.c1_geometry(lib32) if layout == "C1" else mbl.Geometry.preset(lib32, layout)
    res, cd, cl = oracle_series(geom, steps)
    assert l2.shape == (steps, 2) and fd.shape == (steps, 3)
    rtol_res = 5e-6 if exe == "mbl_driver64" else 2e-3
    np.testing.assert_allclose(l2[:, 1], res, rtol=rtol_res, atol=0)
    scale = max(1.0, np.max(np.abs(cd)))
    np.testing.assert_allclose(fd[:, 1], cd, rtol=5e-6, atol=5e-6 * scale)
    np.testing.assert_allclose(fd[:, 2], cl, rtol=5e-6, atol=5e-6 * scale)
