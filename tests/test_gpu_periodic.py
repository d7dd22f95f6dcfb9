"""Single periodic block (the bandwidth-benchmark workload) on the GPU against
the FP64 oracle (oracle/oracle.c: stb1/stb2/stb3 of
multi/FH/evolution_block.h:3-126 with periodic pulls).

Tolerances, after 20 steps at tau = 0.6 from u ~ N(0, 0.01^2), about 5x the measured
maximum error of each storage (in brackets):
  FP64 planes            |F - F_oracle| <= 1e-13
  FP32 planes (f)        <= 2e-7   [3.6e-8: binary32 rounding of f ~ 0.03-0.45 at the store]
  FP32 planes (f - w)    <= 1e-8   [2.2e-9]
  FP16 planes (f - w)    <= 1.5e-5 [2.7e-6: binary16, 11-bit significand of |f - w| <~ 0.01]
  BF16 planes (f - w)    <= 1.5e-4 [2.6e-5: bfloat16, 8-bit significand]
  1-4-11 planes (f - w)  <= 1e-5   [1.6e-6: 12-bit significand]
  1-3-12 planes (f - w)  <= 1e-4   [1.6e-5: most |f - w| here lie below 2^-6, where the
                                    format has a fixed step of 2^-18]
"""
import numpy as np
import pytest

import mbl
from mbl import abi
import oracle

pytestmark = pytest.mark.gpu

NX, NY, STEPS, TAU = 96, 64, 20, 0.6


@pytest.fixture(scope="module")
def ref():
    o = oracle.OraclePeriodic(NX, NY, tau=TAU)
    o.run(STEPS)
    F = np.stack([o.get(("F", k)) for k in range(9)])
    o.close()
    return F


def run_gpu(lib, storage=abi.STORE_NATIVE, tuning=None, nx=NX, ny=NY, steps=STEPS):
    b = mbl.PeriodicBlock(lib, nx, ny, tau=TAU, storage=storage, tuning=tuning)
    b.step(steps)
    b.sync()
    F = np.stack([b.read_F(k) for k in range(9)])
    b.close()
    return F


def test_init_matches_oracle(lib64):
    o = oracle.OraclePeriodic(NX, NY, tau=TAU)
    F0 = np.stack([o.get(("F", k)) for k in range(9)])
    g = run_gpu(lib64, steps=0)
    assert np.max(np.abs(g - F0)) < 1e-15


def test_fp64_matches_oracle(lib64, ref):
    g = run_gpu(lib64)
    err = np.max(np.abs(g - ref))
    print("fp64 max err", err)
    assert err < 1e-13


@pytest.mark.parametrize("storage,tol", [(abi.STORE_NATIVE, 2e-7), (abi.STORE_F32S, 1e-8),
                                         (abi.STORE_F16S, 1.5e-5), (abi.STORE_BF16S, 1.5e-4),
                                         (abi.STORE_F16C, 1e-5), (abi.STORE_F16D, 1e-4)])
def test_fp32_storage_modes(lib32, ref, storage, tol):
    g = run_gpu(lib32, storage=storage)
    err = np.max(np.abs(g - ref))
    print("storage", storage, "max err", err)
    assert err < tol


@pytest.mark.parametrize("vx", [1, 2, 4])
@pytest.mark.parametrize("tpb", [64, 128, 256])
def test_tilings_bitwise_equal(lib32, vx, tpb):
    base = run_gpu(lib32, tuning=(1, 256), nx=128, ny=40, steps=7)
    g = run_gpu(lib32, tuning=(vx, tpb), nx=128, ny=40, steps=7)
    assert np.array_equal(base, g)


@pytest.mark.parametrize("storage", [abi.STORE_F16C, abi.STORE_F16D])
@pytest.mark.parametrize("nx,ny", [(96, 61), (128, 40)])
def test_custom_16bit_row_tiles(lib32, storage, nx, ny):
    """The 1-4-11 and 1-3-12 planes take four rows per wave in a single block
    (tiles of 4, 8 or 16 rows for 64, 128 or 256 threads), so 61 and 40 rows
    leave a partial last tile: every workgroup size gives the same bits, and
    the result stays within the format's rounding of the FP64 oracle."""
    o = oracle.OraclePeriodic(nx, ny, tau=TAU)
    o.run(STEPS)
    F = np.stack([o.get(("F", k)) for k in range(9)])
    o.close()
    runs = [run_gpu(lib32, storage=storage, tuning=(1, tpb), nx=nx, ny=ny) for tpb in (64, 128, 256)]
    for g in runs[1:]:
        assert np.array_equal(runs[0], g)
    err = np.max(np.abs(runs[0] - F))
    print("storage", storage, nx, ny, "max err", err)
    assert err < 1e-4


def test_mass_and_momentum_conserved(lib64):
    b = mbl.PeriodicBlock(lib64, NX, NY, tau=TAU)
    F0 = np.stack([b.read_F(k) for k in range(9)])
    b.step(50)
    b.sync()
    F1 = np.stack([b.read_F(k) for k in range(9)])
    b.close()
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])[:, None, None]
    ey = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])[:, None, None]
    assert abs(F1.sum() - F0.sum()) < 1e-10 * F0.sum()
    assert abs((F1 * ex).sum() - (F0 * ex).sum()) < 1e-12 * NX * NY
    assert abs((F1 * ey).sum() - (F0 * ey).sum()) < 1e-12 * NX * NY


@pytest.mark.parametrize("fp64,storage", [(True, 0), (False, 0), (False, 2), (False, 5)])
def test_device_moments_match_host_sums(lib32, lib64, fp64, storage):
    """mbl_block_moments (the full-size invariant bench.py reports) against the same sums of
    the planes read back: mass and momentum, before and after 20 steps, and no non-finite
    entry."""
    lib = lib64 if fp64 else lib32
    b = mbl.PeriodicBlock(lib, NX, NY, tau=TAU, storage=storage)
    ex = np.array([0, 1, 0, -1, 0, 1, -1, -1, 1])[:, None, None]
    ey = np.array([0, 0, 1, 0, -1, 1, 1, -1, -1])[:, None, None]
    for steps in (0, 20):
        b.step(steps)
        b.sync()
        F = np.stack([b.read_F(k) for k in range(9)])
        m = b.moments()
        assert m[3] == 0
        assert abs(m[0] - F.sum()) < 1e-12 * F.sum()
        assert abs(m[1] - (F * ex).sum()) < 1e-12 * NX * NY
        assert abs(m[2] - (F * ey).sum()) < 1e-12 * NX * NY
    b.close()


def test_odd_width_periodic(lib32, lib64):
    """nx not a multiple of the vector width falls back to one node per thread."""
    o = oracle.OraclePeriodic(37, 23, tau=TAU)
    o.run(9)
    F = np.stack([o.get(("F", k)) for k in range(9)])
    o.close()
    g = run_gpu(lib64, nx=37, ny=23, steps=9)
    assert np.max(np.abs(g - F)) < 1e-13


def test_bandwidth_probes(lib32):
    """The streaming-copy probes bench.py reports next to the roofline: both
    return a rate between a tenth of the 8 TB/s peak and the peak itself, and
    reject a zero repetition count."""
    import ctypes
    cfg = mbl.default_config(lib32, LEVEL=1, write_files=0)
    cfg.mi[0] = 1
    h = ctypes.c_void_p()
    assert lib32.mbl_handle_create(ctypes.byref(h), ctypes.byref(cfg)) == 0
    g = ctypes.c_double()
    assert lib32.mbl_bandwidth_probe(h, 256 << 20, 3, ctypes.byref(g)) == 0
    assert 800.0 < g.value < 8000.0, g.value
    assert lib32.mbl_bandwidth_probe_planes(h, 32 << 20, 3, ctypes.byref(g)) == 0
    assert 800.0 < g.value < 8000.0, g.value
    assert lib32.mbl_bandwidth_probe(h, 256 << 20, 0, ctypes.byref(g)) != 0
    lib32.mbl_handle_destroy(h)
