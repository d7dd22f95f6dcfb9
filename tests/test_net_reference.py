"""Pins the CPU reference of the warm-start network (tests/net_reference.py).

- its convolution is torch's conv2d; both are checked against a loop over
  taps, so a flipped kernel or a wrong padding fails;
- its gradients (autograd) against central finite differences in float64;
- its BF16 rounding against a bit-level round-to-nearest-even restatement,
  including exact ties;
- the parameter layout against the closed-form parameter count.
"""
import os
import sys

import numpy as np
import pytest
import torch

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", "multi-block-lbm-gpu_amd"))
sys.path.insert(0, os.path.dirname(__file__))

import net_reference as ref  # noqa: E402
from mbl.net import layer_table, unpack  # noqa: E402


def random_layers(in_ch, out_ch, base, seed=0):
    table, n = layer_table(in_ch, out_ch, base)
    rng = np.random.default_rng(seed)
    p = np.zeros(n, np.float32)
    for name, cin, cr, cout, cor, k, woff, boff in table:
        w = np.zeros((cout, k, k, cin), np.float32)
        w[:cor, :, :, :cr] = rng.normal(0, np.sqrt(2.0 / (cr * k * k)), (cor, k, k, cr))
        p[woff:woff + w.size] = w.ravel()
        p[boff:boff + cor] = rng.normal(0, 0.1, cor)
    return p, unpack(p, in_ch, out_ch, base)


def test_param_count_closed_form():
    for in_ch, out_ch, base in [(5, 3, 32), (7, 3, 64), (1, 1, 32)]:
        c = [base, 2 * base, 4 * base, 4 * base]
        n = 32 * 9 * c[0] + c[0] + 9 * c[0] * c[0] + c[0]
        for lv in range(1, 4):
            n += 9 * c[lv - 1] * c[lv] + c[lv] + 9 * c[lv] * c[lv] + c[lv]
        for lv in (2, 1, 0):
            n += 9 * (c[lv + 1] + c[lv]) * c[lv] + c[lv] + 9 * c[lv] * c[lv] + c[lv]
        n += c[0] * 32 + 32
        assert layer_table(in_ch, out_ch, base)[1] == n


def test_conv_matches_tap_loop():
    rng = np.random.default_rng(1)
    x = rng.normal(size=(2, 3, 6, 5))
    for k in (1, 3):
        w = rng.normal(size=(4, 3, k, k))
        b = rng.normal(size=4)
        y = torch.nn.functional.conv2d(torch.tensor(x), torch.tensor(w), torch.tensor(b), padding=k // 2).numpy()
        np.testing.assert_allclose(y, ref.conv_direct(x, w, b), rtol=1e-12, atol=1e-12)
    # an asymmetric kernel: a flipped (convolution instead of correlation) kernel differs
    w = np.zeros((1, 1, 3, 3))
    w[0, 0, 0, 2] = 1.0  # reads x[y-1][x+1]
    img = np.arange(30.0).reshape(1, 1, 5, 6)
    y = ref.conv_direct(img, w, np.zeros(1))
    assert y[0, 0, 2, 2] == img[0, 0, 1, 3]


def test_gradients_match_finite_differences():
    in_ch, out_ch, base = 3, 2, 4
    _, layers = random_layers(in_ch, out_ch, base, seed=3)
    rng = np.random.default_rng(4)
    x = rng.normal(size=(2, 8, 16, in_ch))
    t = rng.normal(size=(2, 8, 16, out_ch))
    loss, grads = ref.loss_and_grads(layers, x, t)
    eps = 1e-6
    for name in ["e0a", "e2b", "e3a", "d1a", "d0b", "head"]:
        w, b = layers[name]
        for idx in [tuple(rng.integers(0, s) for s in w.shape) for _ in range(3)]:
            wp = {k: (v[0].copy(), v[1].copy()) for k, v in layers.items()}
            wm = {k: (v[0].copy(), v[1].copy()) for k, v in layers.items()}
            wp[name][0][idx] += eps
            wm[name][0][idx] -= eps
            fd = (ref.loss_and_grads(wp, x, t)[0] - ref.loss_and_grads(wm, x, t)[0]) / (2 * eps)
            assert abs(fd - grads[name][0][idx]) <= 1e-6 * max(1.0, abs(fd)), (name, idx, fd, grads[name][0][idx])
        j = int(rng.integers(0, b.size))
        bp = {k: (v[0].copy(), v[1].copy()) for k, v in layers.items()}
        bm = {k: (v[0].copy(), v[1].copy()) for k, v in layers.items()}
        bp[name][1][j] += eps
        bm[name][1][j] -= eps
        fd = (ref.loss_and_grads(bp, x, t)[0] - ref.loss_and_grads(bm, x, t)[0]) / (2 * eps)
        assert abs(fd - grads[name][1][j]) <= 1e-6 * max(1.0, abs(fd))


def bf16_rne_bits(x):
    """Round float32 values to BF16 (nearest, ties to even) by integer arithmetic."""
    u = np.asarray(x, np.float32).view(np.uint32).astype(np.uint64)
    lower = u & 0xFFFF
    upper = u >> 16
    up = (lower > 0x8000) | ((lower == 0x8000) & ((upper & 1) == 1))
    r = (upper + up.astype(np.uint64)) << 16
    return r.astype(np.uint32).view(np.float32)


def test_bf16_rounding_is_nearest_even():
    rng = np.random.default_rng(5)
    x = rng.normal(size=10000).astype(np.float32) * np.float32(10.0) ** rng.integers(-6, 6, 10000)
    # exact ties: lower half 0x8000, both parities of the kept bit
    ties = (np.arange(1, 200, dtype=np.uint32) << 16 | 0x8000).view(np.float32)
    for v in (x, ties):
        got = ref.bf16_round(torch.tensor(v, dtype=torch.float64)).numpy().astype(np.float32)
        np.testing.assert_array_equal(got, bf16_rne_bits(v))


@pytest.mark.parametrize("bf16", [False, True])
def test_forward_shapes_and_skip_order(bf16):
    in_ch, out_ch, base = 2, 3, 4
    _, layers = random_layers(in_ch, out_ch, base, seed=6)
    x = np.random.default_rng(7).normal(size=(1, 8, 8, in_ch))
    y = ref.output(layers, x, bf16=bf16)
    assert y.shape == (1, 8, 8, out_ch) and np.isfinite(y).all()
    # d0a reads the concatenation of the upsampled level-1 output and the skip
    assert layers["d0a"][0].shape[1] == 2 * base + base
