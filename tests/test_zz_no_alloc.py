"""mbl_error_L2_noalloc (the allocation-free variant of error_L2) against error_L2 through the
C ABI: two identical runs, one recording its residual with error_L2, the other with the
allocation-free entry point after every coarse step, on a three-level layout and then, in
the same process, on a smaller two-level one. The residual series must be equal bit for bit
(same kernels, same summation order), and so must the previous-velocity planes both update."""
import ctypes as C

import numpy as np
import pytest

import mbl

pytestmark = pytest.mark.gpu


def tiny2(lib):
    return mbl.Geometry.channel(lib, 61, 25, [(10, 30, 6, 18)], R=2.5, CX=18.0, CY=12.0, Re=20.0)


def tiny3(lib):
    return mbl.Geometry.channel(lib, 81, 33, [(12, 44, 6, 26), (16, 36, 10, 22)], R=2.0, CX=22.0, CY=16.0,
                                Re=20.0)


def compare(lib, geom, steps):
    # a steps through the reference calls (evolution, error_L2, getDrag_block): its residual is
    # error_L2's, the one the allocation-free entry point reproduces bit for bit
    a = mbl.MultiBlockSim(lib, geom, flush_interval=4096, fused=False)
    b = mbl.MultiBlockSim(lib, geom, flush_interval=4096)
    d_out = lib.mbl_malloc(4 * 8)
    rec = np.zeros(4)
    got = []
    try:
        for _ in range(steps):
            a.step(1)
            b.step(1, diagnostics=False)
            rc = lib.mbl_error_L2_noalloc(b.h, b.na, b.areas, b.pps_d, b.upx, b.upy, d_out)
            assert rc == 0
            assert lib.mbl_memcpy_d2h(rec.ctypes.data, d_out, 32) == 0
            got.append(rec.copy())
        l2, _ = a.series()
        got = np.array(got)
        assert np.array_equal(got[:, 0], l2["residual"])
        assert np.all(got[:, 3] > 0.9) and np.all(got[:, 3] < 1.1)  # smallest density
        for blk in range(a.na):
            assert np.array_equal(a.read(blk, "upx"), b.read(blk, "upx"))
            assert np.array_equal(a.read(blk, "upy"), b.read(blk, "upy"))
    finally:
        lib.mbl_free(d_out)
        a.close()
        b.close()


@pytest.mark.parametrize("fp64", [False, True])
def test_noalloc_residual_equals_error_l2(fp64):
    lib = mbl.load_library(fp64)
    compare(lib, tiny3(lib), 12)  # first call: three levels, 9 blocks
    compare(lib, tiny2(lib), 8)   # second call, smaller input: two levels, 5 blocks


def test_noalloc_rejects_oversized_layout():
    """A layout beyond the scratch array's capacity is refused before any work (the plane
    pointers of the oversized block are never dereferenced)."""
    lib = mbl.load_library(False)
    geom = tiny2(lib)
    sim = mbl.MultiBlockSim(lib, geom)
    try:
        areas = (mbl.AREA * 1)()
        areas[0].m, areas[0].nx, areas[0].ny = 1, 16384, 16384
        pps = (mbl.D2Q9 * 1)()
        pps[0] = sim.pps_d[0]
        upx = (C.c_void_p * 1)(sim.upx[0])
        upy = (C.c_void_p * 1)(sim.upy[0])
        d_out = lib.mbl_malloc(32)
        assert lib.mbl_error_L2_noalloc(sim.h, 1, areas, pps, upx, upy, d_out) == -1
        assert lib.mbl_error_L2_noalloc(sim.h, 65, sim.areas, sim.pps_d, sim.upx, sim.upy, d_out) == -1
        lib.mbl_free(d_out)
        assert sim.last_error() == 0
    finally:
        sim.close()
