"""mbl_step_block on a periodic all-fluid block: two or three time steps per launch (the
intermediate steps kept in LDS) against one step per launch (mbl_set_step_pairs(h, 0)), every
storage mode and the FP64 build: the state in F bit for bit after step counts of every remainder,
and a block whose shape does not allow pairs stepping one step per launch either way."""
import numpy as np
import pytest

import mbl
from mbl import abi

pytestmark = pytest.mark.gpu


def states(lib, n, ny, storage, steps, pairs, sigma=0.02):
    """pairs: 0 single steps, 1 paired tiles, 2 paired strips (where the shape allows), 3 three
    steps per launch."""
    blk = mbl.PeriodicBlock(lib, n, ny, tau=0.6, storage=storage, sigma=sigma)
    mbl.check(lib, lib.mbl_set_step_pairs(blk.h, int(pairs)), "pairs")
    out = []
    done = 0
    for s in steps:
        blk.step(s - done)
        done = s
        out.append([blk.read_F(k) for k in range(9)])
    blk.close()
    return out


@pytest.mark.parametrize("storage", [abi.STORE_NATIVE, abi.STORE_F32S, abi.STORE_F16S, abi.STORE_F16D])
def test_pairs_equal_single_steps(lib32, storage):
    steps = [2, 7, 12]  # 1 pair; then 2 pairs + 1 single; then 2 pairs + 1 single
    a = states(lib32, 128, 64, storage, steps, True)
    b = states(lib32, 128, 64, storage, steps, False)
    for sa, sb in zip(a, b):
        for k in range(9):
            assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("storage", [abi.STORE_NATIVE, abi.STORE_F16D])
@pytest.mark.parametrize("shape", [(300, 128), (512, 64)])
def test_strips_equal_single_steps(lib32, storage, shape):
    """Paired strips (254 output columns and 64 rows per workgroup, a 3-row window of step n+1):
    a width that leaves a partial last strip (300) and one that does not (512)."""
    steps = [2, 5]
    a = states(lib32, shape[0], shape[1], storage, steps, 2)
    b = states(lib32, shape[0], shape[1], storage, steps, 0)
    for sa, sb in zip(a, b):
        for k in range(9):
            assert np.array_equal(sa[k], sb[k]), k


def test_pairs_equal_single_steps_fp64(lib64):
    steps = [4, 9]
    a = states(lib64, 64, 40, abi.STORE_NATIVE, steps, True)
    b = states(lib64, 64, 40, abi.STORE_NATIVE, steps, False)
    for sa, sb in zip(a, b):
        for k in range(9):
            assert np.array_equal(sa[k], sb[k]), k


@pytest.mark.parametrize("storage", [abi.STORE_NATIVE, abi.STORE_F32S, abi.STORE_F16S, abi.STORE_F16D])
def test_triples_equal_single_steps(lib32, storage):
    """Mode 3: three steps per launch for 4-byte planes (step n+1 with a two-node ring and step
    n+2 with a one-node ring in LDS), pairs for 16-bit planes: 3 steps (one launch), 7 (two
    launches + one single step), 11 (three launches + one pair), after which the F planes hold the
    single-step state bit for bit."""
    steps = [3, 7, 11]
    a = states(lib32, 128, 64, storage, steps, 3)
    b = states(lib32, 128, 64, storage, steps, 0)
    for sa, sb in zip(a, b):
        for k in range(9):
            assert np.array_equal(sa[k], sb[k]), k


def test_triples_equal_single_steps_fp64(lib64):
    """Mode 3 on FP64 planes (three steps per launch, 64 x 8 tiles; a remainder of two as a pair)."""
    steps = [3, 8]
    a = states(lib64, 64, 40, abi.STORE_NATIVE, steps, 3)
    b = states(lib64, 64, 40, abi.STORE_NATIVE, steps, 0)
    for sa, sb in zip(a, b):
        for k in range(9):
            assert np.array_equal(sa[k], sb[k]), k


def test_triples_launch_count(lib32):
    """FP32 planes: three steps per launch by default, remainders as one pair or one single step;
    16-bit planes: pairs."""
    blk = mbl.PeriodicBlock(lib32, 128, 64, tau=0.6, sigma=0.02)
    blk.step(1)
    counts = [blk.launches(n) for n in (3, 4, 5, 6, 11)]
    blk.close()
    assert counts == [1, 2, 2, 2, 4]
    blk = mbl.PeriodicBlock(lib32, 128, 64, tau=0.6, storage=abi.STORE_F16D, sigma=0.02)
    blk.step(1)
    counts = [blk.launches(n) for n in (3, 4, 6)]
    blk.close()
    assert counts == [2, 2, 3]


def test_shape_without_pairs(lib32):
    """96 x 40: the width is not a multiple of 64, so every launch is a single step."""
    a = states(lib32, 96, 40, abi.STORE_NATIVE, [6], True)
    b = states(lib32, 96, 40, abi.STORE_NATIVE, [6], False)
    for k in range(9):
        assert np.array_equal(a[0][k], b[0][k])


def test_pairs_conserve_mass(lib32):
    blk = mbl.PeriodicBlock(lib32, 256, 128, tau=0.6, sigma=0.02)
    m0 = blk.moments()
    blk.step(40)
    m1 = blk.moments()
    blk.close()
    assert m1[3] == 0
    # FP32 planes: each step rounds every value once (4e-10 relative measured over 40 steps)
    assert abs(m1[0] - m0[0]) / m0[0] < 1e-8
