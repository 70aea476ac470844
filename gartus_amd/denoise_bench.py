"""Throughput of the denoiser's kernels: inference on a 1920x1080 frame and
training steps on a 16 x 256 x 256 batch, with synthetic inputs.

Reports FLOP/s against the MI355X dense bf16 matrix peak (2.5 PFLOP/s).
FLOPs per convolution = 2 * 9 * Cin * Cout * pixels of its level (padded
sizes, the work the kernels do). A training step is forward + data gradient
+ weight gradient, 3x the forward.

    python -m gartus_amd.denoise_bench [--reps 10]
"""
from __future__ import annotations

import argparse
import json
import sys
import time

import numpy as np

from . import api
from .denoise_train import DeviceArray
from .denoiser import Denoiser

BF16_PEAK = 2.5e15


def forward_flops(levels: int, channels, n: int, h: int, w: int) -> float:
    m = 1 << (levels - 1)
    hp, wp = -(-h // m) * m, -(-w // m) * m
    ch = list(channels)[:levels]
    total = 0.0
    for l in range(levels):
        px = n * (hp >> l) * (wp >> l)
        cin = 16 if l == 0 else ch[l - 1]
        total += 2 * 9 * px * (cin * ch[l] + ch[l] * ch[l])
        if l < levels - 1:
            total += 2 * 9 * px * ((ch[l + 1] + ch[l]) * ch[l] + ch[l] * ch[l])
    total += 2 * 9 * n * hp * wp * ch[0] * 16
    return total


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--reps", type=int, default=10)
    args = ap.parse_args(argv)
    ctx = api.default_context()
    dn = Denoiser(ctx)
    out = {}
    rng = np.random.default_rng(0)

    def arrays(n, h, w):
        res = []
        for k in range(4):
            d = DeviceArray(ctx, (n, h, w, 3))
            host = rng.uniform(0, 1, (n, h, w, 3)).astype(np.float32)
            ctx.lib.gt_memcpy_to_device(ctx.handle, d.ptr, host.ctypes.data, host.nbytes)
            res.append(d)
        return res

    # inference, 1080p
    b, a, nr, o = arrays(1, 1080, 1920)
    dn.infer(b, a, nr, out=o)
    t0 = time.perf_counter()
    for _ in range(args.reps):
        dn.infer(b, a, nr, out=o)
    t = (time.perf_counter() - t0) / args.reps
    f = forward_flops(dn.levels, dn.channels, 1, 1080, 1920)
    out["inference_1080p"] = {"ms": t * 1e3, "gflop": f / 1e9, "tflops": f / t / 1e12, "frac_of_bf16_peak": f / t / BF16_PEAK}
    print(f"inference 1920x1080: {t * 1e3:.2f} ms, {f / 1e9:.0f} GFLOP, {f / t / 1e12:.1f} TFLOP/s", flush=True)
    del b, a, nr, o
    # training, 16 x 256 x 256
    b, a, nr, tg = arrays(16, 256, 256)
    dn.train_step(b, a, nr, tg, lr=1e-6)
    t0 = time.perf_counter()
    for _ in range(args.reps):
        dn.train_step(b, a, nr, tg, lr=1e-6)
    t = (time.perf_counter() - t0) / args.reps
    f = 3 * forward_flops(dn.levels, dn.channels, 16, 256, 256)
    out["train_step_16x256x256"] = {"ms": t * 1e3, "gflop": f / 1e9, "tflops": f / t / 1e12,
                                    "frac_of_bf16_peak": f / t / BF16_PEAK}
    print(f"train step 16x256x256: {t * 1e3:.2f} ms, {f / 1e9:.0f} GFLOP, {f / t / 1e12:.1f} TFLOP/s", flush=True)
    print(json.dumps(out), flush=True)
    dn.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
