"""Time to a target relMSE: the plain path tracer against path tracer + denoiser.

On the 1920x1080 GGX/environment frame:
  - reference: 16,384 spp, accumulated in FP64 on the GPU;
  - plain: renders at increasing spp, each timed (device work plus launch,
    synchronised) and scored by relMSE against the reference;
  - denoised: a low-spp render with its albedo/normal AOVs plus one network
    inference, timed together and scored the same way.
The error is relMSE (eps 1e-2) in two forms: over every pixel, and without
the 0.1% of pixels with the largest error. The scene's 5e4 sun seen through
grazing GGX and specular paths gives the reference's estimator single
samples of 1e8 and more (the f64 oracle reproduces them), so the plain mean
is dominated by a handful of pixels and does not fall as 1/spp; the trimmed
form is the one the time to target uses. Time to relMSE 0.01 is interpolated
on log(time) against log(relMSE) between the two measured points that
bracket it. When every plain render is
above 0.01, the Monte Carlo rate (relMSE proportional to 1/spp) extrapolates
from the best point, and the result says so.

    python -m gartus_amd.denoise_eval --weights W.npy [--ref-spp 16384] [--json out.json]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np
# torch (device memory only) must load its HIP runtime before libgartus_hip.so does: the
# library then binds to the same runtime instead of opening the device a second time
import torch  # noqa: F401

from . import api, scenes
from .denoise_train import SceneHandle, relmse, relmse_trimmed
from .denoiser import Denoiser

TARGET = 0.01


def _timed(fn, reps: int = 1) -> float:
    import torch
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / reps


def time_to_target(points, target: float = TARGET):
    """points: [(time_s, relmse)] ordered by cost. Returns (time, how)."""
    pts = sorted(points)
    for (t0, e0), (t1, e1) in zip(pts, pts[1:]):
        if e0 > target >= e1:
            f = (math.log(e0) - math.log(target)) / (math.log(e0) - math.log(e1))
            return math.exp(math.log(t0) + f * (math.log(t1) - math.log(t0))), "interpolated"
    hit = [t for t, e in pts if e <= target]
    if hit:
        return min(hit), "measured"
    t, e = min(pts, key=lambda p: p[1])
    return t * e / target, "extrapolated (relMSE ~ 1/spp)"


def evaluate(dn: Denoiser, ref_spp: int = 16384, plain_spp=(4, 16, 64, 256, 1024, 4096),
             denoised_spp=(1, 2, 4, 8, 16), width: int = 1920, log=print):
    import torch
    ctx = dn.ctx
    data, env, cam = scenes.ggx_environment()
    cam["width"] = width
    sh = SceneHandle(ctx, data, env, cam)
    h, w = sh.h, sh.w
    ref = torch.empty((h, w, 3), dtype=torch.float64, device="cuda")
    t0 = time.perf_counter()
    sh.render(ref, ref_spp, seed=987_654_321, double=True)
    log(f"reference: {ref_spp} spp in {time.perf_counter() - t0:.1f} s")
    img = torch.empty((h, w, 3), dtype=torch.float32, device="cuda")
    plain = []
    sh.render(img, 1, seed=5)  # warm
    for spp in plain_spp:
        t = _timed(lambda: sh.render(img, spp, seed=1000 + spp))
        e, et = relmse(img, ref), relmse_trimmed(img, ref)
        plain.append({"spp": spp, "time_s": t, "relmse": e, "relmse_trimmed": et})
        log(f"plain {spp:5d} spp: {t * 1e3:8.2f} ms relMSE {e:.5g} trimmed {et:.5f}")
    b = torch.empty((1, h, w, 3), dtype=torch.float32, device="cuda")
    a, nr, out = torch.empty_like(b), torch.empty_like(b), torch.empty_like(b)
    den = []
    sh.render_aovs(b[0], a[0], nr[0], 1, seed=7)  # warm
    dn.infer(b, a, nr, out=out)
    for spp in denoised_spp:
        def run():
            sh.render_aovs(b[0], a[0], nr[0], spp, seed=2000 + spp)
            dn.infer(b, a, nr, out=out)
        t = _timed(run)
        t_inf = _timed(lambda: dn.infer(b, a, nr, out=out))
        e, et = relmse(out[0], ref), relmse_trimmed(out[0], ref)
        e_in = relmse_trimmed(b[0], ref)
        den.append({"spp": spp, "time_s": t, "infer_s": t_inf, "relmse": e, "relmse_trimmed": et,
                    "relmse_trimmed_input": e_in})
        log(f"denoised {spp:3d} spp: {t * 1e3:8.2f} ms (inference {t_inf * 1e3:.2f} ms) relMSE {e:.5g} "
            f"trimmed {et:.5f} (input trimmed {e_in:.5f})")
    tp, how_p = time_to_target([(p["time_s"], p["relmse_trimmed"]) for p in plain])
    td, how_d = time_to_target([(p["time_s"], p["relmse_trimmed"]) for p in den])
    result = {
        "frame": f"{w}x{h} GGX/HDR-environment, depth 10",
        "reference_spp": ref_spp,
        "relmse_eps": 1e-2,
        "relmse_trim_fraction": 1e-3,
        "target_relmse": TARGET,
        "target_on": "relmse_trimmed",
        "plain": plain,
        "denoised": den,
        "plain_time_to_target_s": tp,
        "plain_how": how_p,
        "denoised_time_to_target_s": td,
        "denoised_how": how_d,
        "speedup": tp / td,
    }
    log(f"time to relMSE {TARGET}: plain {tp:.3f} s ({how_p}), denoised {td:.4f} s ({how_d}), "
        f"speedup {tp / td:.1f}x")
    return result


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", required=True)
    ap.add_argument("--ref-spp", type=int, default=16384)
    ap.add_argument("--json", default="")
    args = ap.parse_args(argv)
    dn = Denoiser(api.default_context())
    dn.load(args.weights)
    res = evaluate(dn, ref_spp=args.ref_spp, log=lambda m: print(m, flush=True))
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
