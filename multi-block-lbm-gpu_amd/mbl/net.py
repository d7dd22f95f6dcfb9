"""The warm-start network through the C ABI (mbl_net_*).

A 4-level convolutional encoder-decoder over an H x W raster: levels of
base, 2 base, 4 base, 4 base channels, two 3x3 convolutions with ReLU per
level, 2x2 max pooling down, nearest x2 upsampling with skip concatenation up,
and a linear 1x1 head. BF16 activations, BF16 MFMA with FP32 accumulation,
FP32 master weights and Adam, all on the GPU.

The flat FP32 parameter vector holds, per layer in the order
e0a e0b e1a e1b e2a e2b e3a e3b d2a d2b d1a d1b d0a d0b head,
the weights [cout][ksize*ksize][cin] and then the bias [cout]. The first layer
reads the input padded to 32 channels and the head writes 32 channels of
which the first out_ch are the outputs; the padding weights are zero.
"""
import ctypes as C

import numpy as np

from . import MblError, check

F32P = C.POINTER(C.c_float)


def layer_table(in_ch, out_ch, base):
    """[(name, cin, cin_real, cout, cout_real, ksize, woff, boff)] in parameter order."""
    c = [base, 2 * base, 4 * base, 4 * base]
    spec = [("e0a", 32, in_ch, c[0], c[0], 3), ("e0b", c[0], c[0], c[0], c[0], 3)]
    for lv in range(1, 4):
        spec.append((f"e{lv}a", c[lv - 1], c[lv - 1], c[lv], c[lv], 3))
        spec.append((f"e{lv}b", c[lv], c[lv], c[lv], c[lv], 3))
    for lv in (2, 1, 0):
        spec.append((f"d{lv}a", c[lv + 1] + c[lv], c[lv + 1] + c[lv], c[lv], c[lv], 3))
        spec.append((f"d{lv}b", c[lv], c[lv], c[lv], c[lv], 3))
    spec.append(("head", c[0], c[0], 32, out_ch, 1))
    out, off = [], 0
    for name, cin, cr, cout, cor, k in spec:
        woff = off
        off += cout * k * k * cin
        boff = off
        off += cout
        out.append((name, cin, cr, cout, cor, k, woff, boff))
    return out, off


def unpack(params, in_ch, out_ch, base):
    """{name: (W[cout][cin][ky][kx], b[cout])} as float64 arrays, padding removed."""
    table, n = layer_table(in_ch, out_ch, base)
    assert params.size == n
    out = {}
    for name, cin, cr, cout, cor, k, woff, boff in table:
        w = params[woff:woff + cout * k * k * cin].reshape(cout, k, k, cin).astype(np.float64)
        out[name] = (np.ascontiguousarray(w.transpose(0, 3, 1, 2)[:cor, :cr]), params[boff:boff + cor].astype(np.float64))
    return out


class Net:
    def __init__(self, lib, H, W, in_ch, out_ch, base=32, max_batch=16, seed=1):
        self.lib = lib
        self.H, self.W, self.in_ch, self.out_ch, self.base, self.max_batch = H, W, in_ch, out_ch, base, max_batch
        n = C.c_void_p()
        check(lib, lib.mbl_net_create(None, H, W, in_ch, out_ch, base, max_batch, seed, C.byref(n)), "net create")
        self.n = n

    def num_params(self):
        return int(self.lib.mbl_net_num_params(self.n))

    def get_params(self):
        p = np.zeros(self.num_params(), np.float32)
        check(self.lib, self.lib.mbl_net_params(self.n, p.ctypes.data_as(F32P), 1), "net params")
        return p

    def set_params(self, p):
        p = np.ascontiguousarray(p, np.float32)
        if p.size != self.num_params():
            raise ValueError("parameter count")
        check(self.lib, self.lib.mbl_net_params(self.n, p.ctypes.data_as(F32P), 0), "net params")

    def _in(self, x):
        x = np.ascontiguousarray(x, np.float32)
        if x.ndim != 4 or x.shape[1:] != (self.H, self.W, self.in_ch):
            raise ValueError(f"input must be [N, {self.H}, {self.W}, {self.in_ch}], got {x.shape}")
        return x

    def forward(self, x):
        x = self._in(x)
        out = np.zeros((x.shape[0], self.H, self.W, self.out_ch), np.float32)
        for b0 in range(0, x.shape[0], self.max_batch):
            xb = np.ascontiguousarray(x[b0:b0 + self.max_batch])
            ob = np.zeros((xb.shape[0], self.H, self.W, self.out_ch), np.float32)
            check(self.lib, self.lib.mbl_net_forward(self.n, xb.ctypes.data_as(F32P), xb.shape[0],
                                                     ob.ctypes.data_as(F32P)), "net forward")
            out[b0:b0 + xb.shape[0]] = ob
        return out

    def loss_grad(self, x, t):
        x = self._in(x)
        t = np.ascontiguousarray(t, np.float32)
        loss = C.c_double()
        g = np.zeros(self.num_params(), np.float32)
        check(self.lib, self.lib.mbl_net_loss_grad(self.n, x.ctypes.data_as(F32P), t.ctypes.data_as(F32P), x.shape[0],
                                                   C.byref(loss), g.ctypes.data_as(F32P)), "net loss/grad")
        return loss.value, g

    def train(self, x, t, batch, epochs, lr=1e-3, seed=7):
        x = self._in(x)
        t = np.ascontiguousarray(t, np.float32)
        if t.shape != x.shape[:3] + (self.out_ch,):
            raise ValueError("target shape")
        hist = np.zeros(epochs)
        check(self.lib, self.lib.mbl_net_train(self.n, x.ctypes.data_as(F32P), t.ctypes.data_as(F32P), x.shape[0],
                                               batch, epochs, lr, seed,
                                               hist.ctypes.data_as(C.POINTER(C.c_double))), "net train")
        return hist

    def close(self):
        if getattr(self, "n", None):
            self.lib.mbl_net_destroy(self.n)
            self.n = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


__all__ = ["Net", "layer_table", "unpack", "MblError"]
