"""mbl_evolution_step: the reference driver's loop body (evolution -> error_L2 ->
getDrag_block, Multi-block-LBM.cu:144-158) in one call, with the residual summed
by the last stream-collide of each level and one launch after the cycle for the
remaining nodes, the drag and the final sums.

Against the three separate calls on the same layout: every plane, macro and
u_prev array bit for bit (the node updates are the same), the residual and the
forces to rounding (another summation order). Against the FP64 oracle through
the other multi-block tests, which step through the fused call by default.
"""
import numpy as np
import pytest

import mbl

pytestmark = pytest.mark.gpu


def tiny3(lib):
    return mbl.Geometry.channel(lib, 81, 33, [(12, 44, 6, 26), (16, 36, 10, 22)], R=2.0, CX=22.0, CY=16.0,
                                Re=20.0)


def layouts(lib):
    return {"tiny3": tiny3(lib), "GEO_2_5": mbl.Geometry.preset(lib, "GEO_2_5"), "C1": mbl.c1_geometry(lib),
            "C2b": mbl.c2b_geometry(lib)}


def fields_equal(a, b):
    for k in range(a.na):
        for fld in ("rho", "ux", "uy", "upx", "upy"):
            assert np.array_equal(a.read(k, fld), b.read(k, fld)), (k, fld)
        for q in range(9):
            assert np.array_equal(a.read(k, ("F", q)), b.read(k, ("F", q))), (k, q)


def series_close(a, b, rtol):
    la, da = a.series()
    lb, db = b.series()
    assert np.array_equal(la["t"], lb["t"]) and np.array_equal(da["t"], db["t"])
    np.testing.assert_allclose(la["residual"], lb["residual"], rtol=rtol, atol=0)
    scale = max(1.0, float(np.max(np.abs(db["Cd"]))))
    np.testing.assert_allclose(da["Cd"], db["Cd"], rtol=0, atol=rtol * scale)
    np.testing.assert_allclose(da["Cl"], db["Cl"], rtol=0, atol=rtol * scale)


@pytest.mark.parametrize("fp64", [True, False])
@pytest.mark.parametrize("name", ["tiny3", "GEO_2_5", "C1", "C2b"])
def test_fused_step_equals_three_calls(name, fp64):
    lib = mbl.load_library(fp64)
    geom = layouts(lib)[name]
    steps = 37
    sf = mbl.MultiBlockSim(lib, geom, fused=True)
    s3 = mbl.MultiBlockSim(lib, geom, fused=False)
    sf.step(steps)
    s3.step(steps)
    assert lib.mbl_evolution_step_fused(sf.h) == 1
    print(name, "fp64" if fp64 else "fp32", "launches", lib.mbl_cycle_launches(sf.h, 0, None, None, None))
    fields_equal(sf, s3)
    series_close(sf, s3, 1e-12 if fp64 else 2e-6)
    sf.close()
    s3.close()


def test_fused_step_chunks_bitwise_equal(lib32):
    """One call of n steps and n calls of one step: the same bits, series included."""
    geom = tiny3(lib32)
    a = mbl.MultiBlockSim(lib32, geom)
    b = mbl.MultiBlockSim(lib32, geom)
    a.step(24)
    for _ in range(24):
        b.step(1)
    assert lib32.mbl_evolution_step_fused(a.h) == 1
    fields_equal(a, b)
    la, da = a.series()
    lb, db = b.series()
    assert np.array_equal(la["residual"], lb["residual"])
    assert np.array_equal(da["Cd"], db["Cd"]) and np.array_equal(da["Cl"], db["Cl"])
    a.close()
    b.close()


def test_fused_step_falls_back_to_three_calls(lib32):
    """Two nodes per thread: the collides do not sum the residual, so the call makes the three
    reference calls; the results are those of the three calls, bit for bit."""
    geom = tiny3(lib32)
    a = mbl.MultiBlockSim(lib32, geom, fused=True, tuning=(2, 256))
    b = mbl.MultiBlockSim(lib32, geom, fused=False, tuning=(2, 256))
    a.step(15)
    b.step(15)
    assert lib32.mbl_evolution_step_fused(a.h) == 0
    fields_equal(a, b)
    la, da = a.series()
    lb, db = b.series()
    assert np.array_equal(la["residual"], lb["residual"])
    assert np.array_equal(da["Cd"], db["Cd"])
    a.close()
    b.close()


def test_fused_step_rejects_missing_u_prev(lib32):
    import ctypes as C
    geom = tiny3(lib32)
    s = mbl.MultiBlockSim(lib32, geom)
    null = (C.c_void_p * s.na)()
    rc = lib32.mbl_evolution_step(s.h, 0, 1, s.asp, s.ls3, s.ltc, s.lccp, s.block, s.grid, s.areas, s.pps_h,
                                  s.pps_d, s.lc, s.lf, s.lcc, s.t3, s.s3, null, null, s.d_links, s.n_drag, None)
    assert rc == -1  # MBL_E_INVALID
    s.close()
