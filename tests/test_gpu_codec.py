"""The custom 16-bit storage formats (MBL_STORE_F16C: 1-4-11, MBL_STORE_F16D:
1-3-12) through the C ABI.

Decoding is checked on all 65536 codes against the format's definition
(sign, EB exponent bits with bias 2^EB - 1 and no infinities, 15 - EB
fraction bits, subnormals below 2^(2 - 2^EB), 0xFFFF = NaN). Encoding is
checked against a brute-force search for the nearest code (ties to the even
code) over random values, exact ties, the subnormal range and values beyond
the range (saturation)."""
import numpy as np
import pytest

import mbl
from mbl import abi

pytestmark = pytest.mark.gpu

N = 256  # a 256 x 256 block holds one plane of 65536 entries
W0 = 4.0 / 9.0  # planes hold f - w_k; plane k = 0 is read and written here
FORMATS = [(abi.STORE_F16C, 4), (abi.STORE_F16D, 3)]


def definition(codes, eb):
    fb, bias = 15 - eb, (1 << eb) - 1
    c = codes.astype(np.int64)
    s = np.where(c & 0x8000, -1.0, 1.0)
    e = (c >> fb) & ((1 << eb) - 1)
    m = c & ((1 << fb) - 1)
    v = np.where(e > 0, np.ldexp(1.0 + m / 2.0 ** fb, e - bias), np.ldexp(m / 2.0 ** fb, 1 - bias))
    v = s * v
    v[c == 0xFFFF] = np.nan
    return v


def nearest_code(x, eb):
    """Brute force: the code whose value is nearest to x, ties to the even code;
    beyond the largest value the largest code of that sign (magnitude code
    0x7FFE: 0x7FFF is never produced, so no sign gives the NaN code 0xFFFF)."""
    vals = definition(np.arange(0x7FFE + 1), eb)  # increasing for the positive codes
    out = np.empty(x.size, np.int64)
    for i, xv in enumerate(x):
        a = abs(float(xv))
        j = int(np.searchsorted(vals, a))
        if j >= vals.size:
            c = 0x7FFE
        elif j == 0:
            c = 0
        else:
            lo, hi = vals[j - 1], vals[j]
            if a - lo < hi - a:
                c = j - 1
            elif a - lo > hi - a:
                c = j
            else:
                c = j - 1 if (j - 1) % 2 == 0 else j
        if np.signbit(xv):
            c |= 0x8000
        out[i] = c
    return out


@pytest.fixture(scope="module", params=FORMATS, ids=["1-4-11", "1-3-12"])
def blk(request, lib32):
    storage, eb = request.param
    b = mbl.PeriodicBlock(lib32, N, N, storage=storage)
    b.eb = eb
    yield b
    b.close()


def raw_plane(blk):
    out = np.zeros(N * N, np.uint16)
    blk.sync()
    mbl.check(blk.lib, blk.lib.mbl_memcpy_d2h(out.ctypes.data, blk.pp.F[0], out.nbytes), "d2h")
    return out


def test_decode_every_code(blk):
    codes = np.arange(65536, dtype=np.uint16)
    blk.sync()
    mbl.check(blk.lib, blk.lib.mbl_memcpy_h2d(blk.pp.F[0], codes.ctypes.data, codes.nbytes), "h2d")
    got = blk.read_F(0).ravel()  # decoded value + w_0, added in double
    want = definition(codes, blk.eb)
    assert np.isnan(got[0xFFFF])
    ok = np.arange(65536) != 0xFFFF
    np.testing.assert_array_equal(got[ok], want[ok] + W0)


def test_encode_is_nearest_even(blk):
    import ctypes as C

    eb = blk.eb
    rng = np.random.default_rng(7)
    n = N * N
    x = np.concatenate([
        rng.uniform(-2.2, 2.2, 20000),
        rng.normal(0, 1e-3, 20000),
        np.ldexp(rng.uniform(-1, 1, 15000), rng.integers(-30, -4, 15000)),  # around the subnormal range
    ])
    pos = definition(np.arange(0x0100, 0x0900), eb)
    ties = (pos[:-1] + pos[1:]) / 2  # exact ties between neighbouring codes
    x = np.concatenate([x, ties, -ties, [0.0, 1.99951171875, 2.5, -3.0, 1e-9, -1e-9]])
    f = np.resize(x, n) + W0
    # what the kernel encodes: f - w_0 in double, rounded to binary32
    xe = (f - W0).astype(np.float32)
    mbl.check(blk.lib, blk.lib.mbl_write_distribution(blk.h, C.byref(blk.pp), 0, 0, n,
                                                      f.ctypes.data_as(C.POINTER(C.c_double))), "write")
    got = raw_plane(blk).astype(np.int64)
    want = nearest_code(xe, eb)
    bad = np.nonzero(got != want)[0]
    assert bad.size == 0, [(float(xe[i]), hex(got[i]), hex(want[i])) for i in bad[:5]]
    # the exact ties survived the binary32 rounding and were exercised
    assert np.isin(ties.astype(np.float32), xe).sum() > 100
