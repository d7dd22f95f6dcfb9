"""CPU reference of the warm-start network (test infrastructure only).

The same encoder-decoder as the GPU library, in torch on the CPU in float64,
from the per-layer weights that mbl.net.unpack returns. With bf16=True the
values the GPU stores in BF16 are rounded to BF16 (round to nearest even) at
the same places: the packed input, the weights of every convolution and each
convolution's ReLU output; the head output stays unrounded, as the GPU keeps
it in FP32. Rounding passes gradients straight through, so the backward pass
is that of the rounded forward.
"""
import numpy as np
import torch
import torch.nn.functional as F

ORDER = ["e0a", "e0b", "e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "d2a", "d2b", "d1a", "d1b", "d0a", "d0b", "head"]


def bf16_round(t):
    """Round to the nearest BF16 value, ties to even, gradient passed through."""
    r = t.detach().to(torch.float32).to(torch.bfloat16).to(t.dtype)
    return t + (r - t).detach()


class _GradRound(torch.autograd.Function):
    """Identity forward; the incoming gradient rounded to BF16 backward."""

    @staticmethod
    def forward(ctx, t):
        return t.view_as(t)

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.float32).to(torch.bfloat16).to(g.dtype)


def grad_round(t):
    return _GradRound.apply(t)


def make_params(layers, requires_grad=True):
    """{name: (W, b)} numpy -> {name: (W, b)} float64 torch leaves."""
    out = {}
    for k, (w, b) in layers.items():
        out[k] = (torch.tensor(w, dtype=torch.float64, requires_grad=requires_grad),
                  torch.tensor(b, dtype=torch.float64, requires_grad=requires_grad))
    return out


def forward(params, x, bf16=False, grad_bf16=False):
    """x: [N, H, W, in_ch] -> [N, H, W, out_ch] (NHWC in and out).

    grad_bf16 also rounds gradients to BF16 where the GPU stores them: the
    gradient of the network output, of every convolution's input (the GPU
    applies the ReLU mask in the same pass), of each upsampled tensor after
    the sum over its four children, and of each encoder output after the skip
    and pooling parts are added."""
    rnd = bf16_round if bf16 else (lambda t: t)
    grnd = grad_round if grad_bf16 else (lambda t: t)
    h = x.permute(0, 3, 1, 2)
    h = rnd(h)

    def conv(h, name, relu=True):
        w, b = params[name]
        y = F.conv2d(grnd(h), rnd(w), b, padding=w.shape[-1] // 2)
        return rnd(F.relu(y)) if relu else y

    skips = []
    for lv in range(4):
        h = conv(h, f"e{lv}a")
        h = grnd(conv(h, f"e{lv}b"))
        if lv < 3:
            skips.append(h)
            h = F.max_pool2d(h, 2)
    for lv in (2, 1, 0):
        h = torch.cat([F.interpolate(grnd(h), scale_factor=2, mode="nearest"), skips[lv]], dim=1)
        h = conv(h, f"d{lv}a")
        h = conv(h, f"d{lv}b")
    y = grnd(conv(h, "head", relu=False))
    return y.permute(0, 2, 3, 1)


def loss_and_grads(layers, x, t, bf16=False, grad_bf16=False):
    """Mean squared error and {name: (dW, db)} as numpy float64."""
    p = make_params(layers)
    y = forward(p, torch.tensor(x, dtype=torch.float64), bf16=bf16, grad_bf16=grad_bf16)
    loss = ((y - torch.tensor(t, dtype=torch.float64)) ** 2).mean()
    loss.backward()
    return float(loss.detach()), {k: (w.grad.numpy(), b.grad.numpy()) for k, (w, b) in p.items()}


def output(layers, x, bf16=False):
    with torch.no_grad():
        return forward(make_params(layers, False), torch.tensor(x, dtype=torch.float64), bf16=bf16).numpy()


def conv_direct(x, w, b):
    """Zero-padded 'same' cross-correlation by explicit loops over taps (NCHW)."""
    N, Ci, H, W = x.shape
    Co, _, K, _ = w.shape
    h = K // 2
    xp = np.zeros((N, Ci, H + 2 * h, W + 2 * h))
    xp[:, :, h:h + H, h:h + W] = x
    y = np.zeros((N, Co, H, W)) + b[None, :, None, None]
    for ky in range(K):
        for kx in range(K):
            y += np.einsum("nchw,oc->nohw", xp[:, :, ky:ky + H, kx:kx + W], w[:, :, ky, kx])
    return y
