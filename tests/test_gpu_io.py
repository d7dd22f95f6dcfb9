"""Field output (output/output.h:2-38 format) and checkpoint / resume through
the C ABI."""
import numpy as np
import pytest

import mbl
from mbl import abi

pytestmark = pytest.mark.gpu


def tiny_geometry(lib):
    """Two-level channel of 60 x 24 coarse cells, ratio-2 block around a D = 6 cylinder."""
    return mbl.Geometry.channel(lib, 61, 25, [(10, 30, 6, 18)], R=3.0, CX=18.0, CY=12.25, Re=60.0, U=0.1)


@pytest.mark.parametrize("fp64,storage", [(True, abi.STORE_NATIVE), (False, abi.STORE_F16S)])
def test_resume_reproduces_the_run(lib32, lib64, tmp_path, fp64, storage):
    lib = lib64 if fp64 else lib32
    kw = dict(storage=storage, U0=(0.1, 0.0), init_mode=abi.INIT_PROFILE)
    a = mbl.MultiBlockSim(lib, tiny_geometry(lib), **kw)
    a.step(37)
    path = tmp_path / "state.ckpt"
    a.checkpoint(path)
    a.step(41)
    la, da = a.series()
    fa = [a.read(q, ("F", k)) for q in range(a.na) for k in range(9)]

    b = mbl.MultiBlockSim(lib, tiny_geometry(lib), **kw)
    b.step(5)  # a different state and marker phase before the restore
    b.flush()  # the 5 steps' records leave the device before the series is cleared
    b.restore(path)
    assert b.t == 37
    lib.mbl_diagnostics_clear(b.h)
    b.step(41)
    lb, db = b.series()
    fb = [b.read(q, ("F", k)) for q in range(b.na) for k in range(9)]
    np.testing.assert_array_equal(da["Cd"][-41:], db["Cd"])
    np.testing.assert_array_equal(da["Cl"][-41:], db["Cl"])
    np.testing.assert_array_equal(la["residual"][-41:], lb["residual"])
    for x, y in zip(fa, fb):
        np.testing.assert_array_equal(x, y)
    for q in range(b.na):
        assert np.array_equal(a.read(q, "ux"), b.read(q, "ux"))
    a.close()
    b.close()


def test_checkpoint_of_another_storage_is_rejected(lib32, tmp_path):
    a = mbl.MultiBlockSim(lib32, tiny_geometry(lib32), storage=abi.STORE_F16S)
    a.step(3)
    path = tmp_path / "s.ckpt"
    a.checkpoint(path)
    a.close()
    b = mbl.MultiBlockSim(lib32, tiny_geometry(lib32), storage=abi.STORE_NATIVE)
    with pytest.raises(mbl.MblError):
        b.restore(path)
    b.close()


def test_output_writes_the_reference_point_file(lib32, tmp_path):
    sim = mbl.MultiBlockSim(lib32, tiny_geometry(lib32), out_dir=str(tmp_path), U0=(0.1, 0.0),
                            init_mode=abi.INIT_PROFILE)
    sim.step(20)
    sim.output(20)
    text = (tmp_path / "steady_20.dat").read_text().splitlines()
    assert text[0] == 'Title= "LBM 2D"'
    assert text[1] == 'VARIABLES= "X","Y","solid","rho","U","V",'
    line = 2
    for a in range(sim.na):
        ny, nx = sim.shape(a)
        assert text[line] == f'ZONE T= "zone_level{a}",I= {nx},J= {ny},F=POINT'
        rows = np.array([[float(v) for v in r.split("\t")] for r in text[line + 1:line + 1 + nx * ny]])
        line += 1 + nx * ny
        assert rows.shape == (nx * ny, 6)
        for col, field in ((0, "x"), (1, "y"), (2, "solid"), (3, "rho"), (4, "ux"), (5, "uy")):
            want = sim.read(a, field).ravel().astype(np.float64)
            np.testing.assert_allclose(rows[:, col], want, rtol=1e-5, atol=1e-12)  # "%g": 6 significant digits
    assert line == len(text)
    sim.close()
