"""Throughput of the warm-start network's BF16 MFMA kernels.

    python -m mbl.netbench [--H 48] [--W 192] [--batch 32] [--samples 256] [--epochs 3]

Trains on random inputs/targets (the arithmetic does not depend on the
values) and times forward passes; prints one JSON line with samples per
second and the convolution FLOP rate (2 multiply-adds per weight per output
pixel, forward; backward counts twice the forward).
"""
import argparse
import json
import time

import numpy as np

from . import load_library
from .net import Net, layer_table


def conv_flops(H, W, in_ch, out_ch, base):
    """Forward multiply-add FLOPs of one sample (padded channel counts, as computed)."""
    table, _ = layer_table(in_ch, out_ch, base)
    level = [0, 0, 1, 1, 2, 2, 3, 3, 2, 2, 1, 1, 0, 0, 0]
    f = 0
    for (name, cin, cr, cout, cor, k, woff, boff), lv in zip(table, level):
        f += 2 * (H >> lv) * (W >> lv) * cin * cout * k * k
    return f


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--H", type=int, default=48)
    ap.add_argument("--W", type=int, default=192)
    ap.add_argument("--in-ch", type=int, default=7)
    ap.add_argument("--out-ch", type=int, default=3)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--batch", type=int, default=32)
    ap.add_argument("--samples", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=3)
    args = ap.parse_args(argv)
    lib = load_library(False)
    rng = np.random.default_rng(0)
    x = rng.normal(size=(args.samples, args.H, args.W, args.in_ch)).astype(np.float32)
    t = rng.normal(size=(args.samples, args.H, args.W, args.out_ch)).astype(np.float32)
    net = Net(lib, args.H, args.W, args.in_ch, args.out_ch, base=args.base, max_batch=args.batch, seed=1)
    net.train(x[:args.batch], t[:args.batch], batch=args.batch, epochs=1)  # warm-up
    t0 = time.perf_counter()
    net.train(x, t, batch=args.batch, epochs=args.epochs)
    train_s = time.perf_counter() - t0
    net.forward(x[:args.batch])
    t0 = time.perf_counter()
    net.forward(x)
    fwd_s = time.perf_counter() - t0
    f = conv_flops(args.H, args.W, args.in_ch, args.out_ch, args.base)
    n_train = args.samples // args.batch * args.batch * args.epochs
    out = {"raster": [args.H, args.W], "base": args.base, "params": net.num_params(), "batch": args.batch,
           "fwd_gflop_per_sample": round(f / 1e9, 4),
           "train_samples_per_s": round(n_train / train_s, 1),
           "train_tflops": round(3 * f * n_train / train_s / 1e12, 2),
           "forward_samples_per_s": round(args.samples / fwd_s, 1),
           "forward_tflops": round(f * args.samples / fwd_s / 1e12, 2)}
    net.close()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
