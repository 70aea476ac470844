"""Time to a target relMSE: the plain path tracer against path tracer + denoiser.

On the 1920x1080 GGX/environment frame:
  - reference: the average of two independent 16,384-spp renders,
    accumulated in FP64 on the GPU. Their disagreement gives the reference's
    own error: relMSE(a, b) / 4 for the average;
  - plain: renders at increasing spp, each timed (device work plus launch,
    synchronised) and scored by relMSE against the reference;
  - denoised: a low-spp render with its albedo/normal AOVs plus one network
    inference, timed together and scored the same way.

Error is relMSE (eps 1e-2) in two forms: over every pixel, and without the
0.1% of pixels with the largest error. The time to target uses the second.
The sun is 5e4 over 0.5 degrees, and specular or glossy paths that reach it
carry single samples of 5e4 or more. Those caustics make the plain mean
converge slowly and leave the reference with an error of its own, reported
beside the results (with heavy tails relMSE(a, b) / 4 is only indicative:
relMSE is asymmetric in its arguments). Time to a target is
interpolated on log(time) against log(relMSE) between the two measured
points that bracket it.

With the reference's estimator, below-horizon GGX samples fall back to a
retro-reflection and compound (gt_camera flags); that evaluation runs by
default. --reject-ggx-fallback evaluates the well-posed estimator.

    python -m gartus_amd.denoise_eval --weights W.npy [--ref-spp 16384] [--json out.json]
"""
from __future__ import annotations

import argparse
import json
import math
import sys
import time

import numpy as np

from . import abi, api, scenes
from .denoise_train import DeviceArray, SceneHandle, relmse, relmse_trimmed
from .denoiser import Denoiser

TARGET = 0.01


def _timed(fn, reps: int = 1) -> float:
    """Wall time of fn; every call used here returns after its device work."""
    t0 = time.perf_counter()
    for _ in range(reps):
        fn()
    return (time.perf_counter() - t0) / reps


def time_to_target(points, target: float = TARGET):
    """points: [(time_s, relmse)] ordered by cost. Returns (time, how); the
    time is None when no measured point reaches the target (no
    extrapolation: trimmed relMSE is not monotone in spp on this scene)."""
    pts = sorted(points)
    for (t0, e0), (t1, e1) in zip(pts, pts[1:]):
        if e0 > target >= e1:
            f = (math.log(e0) - math.log(target)) / (math.log(e0) - math.log(e1))
            return math.exp(math.log(t0) + f * (math.log(t1) - math.log(t0))), "interpolated"
    hit = [t for t, e in pts if e <= target]
    if hit:
        return min(hit), "measured"
    return None, "not reached by any measured point"


def _reached(points, target):
    """Interpolated time at which the measured series first reaches target
    (None when no measured point does)."""
    pts = sorted(points)
    if pts and pts[0][1] <= target:
        return pts[0][0]
    for (t0, e0), (t1, e1) in zip(pts, pts[1:]):
        if e0 > target >= e1:
            f = (math.log(e0) - math.log(target)) / (math.log(e0) - math.log(e1))
            return math.exp(math.log(t0) + f * (math.log(t1) - math.log(t0)))
    return None


def evaluate(dn: Denoiser, ref_spp: int = 16384, plain_spp=(4, 16, 64, 256, 1024, 4096),
             denoised_spp=(1, 2, 4, 8, 16), width: int = 1920, flags: int = 0, log=print):
    ctx = dn.ctx
    data, env, cam = scenes.ggx_environment()
    cam["width"] = width
    sh = SceneHandle(ctx, data, env, cam, flags)
    h, w = sh.h, sh.w
    # two independent references; their average is the reference, and their
    # disagreement measures its own error: for independent halves a and b,
    # relMSE(a, b) ~ 2 e(N) and the average has e(2N) ~ relMSE(a, b) / 4
    ref_d = DeviceArray(ctx, (h, w, 3), np.float64)
    t0 = time.perf_counter()
    sh.render(ref_d, ref_spp, seed=987_654_321, double=True)
    ref_a = ref_d.numpy()
    sh.render(ref_d, ref_spp, seed=123_456_789, double=True)
    ref_b = ref_d.numpy()
    ref = 0.5 * (ref_a + ref_b)
    floor = relmse(ref_a, ref_b) / 4.0
    floor_t = relmse_trimmed(ref_a, ref_b) / 4.0
    del ref_a, ref_b
    log(f"reference: 2 x {ref_spp} spp in {time.perf_counter() - t0:.1f} s; its own error: relMSE {floor:.4g}, "
        f"trimmed {floor_t:.4g}")
    img = DeviceArray(ctx, (h, w, 3))
    plain = []
    sh.render(img, 1, seed=5)  # warm
    for spp in plain_spp:
        t = _timed(lambda: sh.render(img, spp, seed=1000 + spp))
        x = img.numpy()
        e, et = relmse(x, ref), relmse_trimmed(x, ref)
        plain.append({"spp": spp, "time_s": t, "relmse": e, "relmse_trimmed": et})
        log(f"plain {spp:5d} spp: {t * 1e3:8.2f} ms relMSE {e:.5g} trimmed {et:.5f}")
    b, a, nr, out = (DeviceArray(ctx, (1, h, w, 3)) for _ in range(4))
    den = []
    sh.render_aovs(b.sub(0), a.sub(0), nr.sub(0), 1, seed=7)  # warm
    dn.infer(b, a, nr, out=out)
    for spp in denoised_spp:
        def run():
            sh.render_aovs(b.sub(0), a.sub(0), nr.sub(0), spp, seed=2000 + spp)
            dn.infer(b, a, nr, out=out)
        t = _timed(run)
        t_inf = _timed(lambda: dn.infer(b, a, nr, out=out))
        y, x = out.numpy()[0], b.numpy()[0]
        e, et = relmse(y, ref), relmse_trimmed(y, ref)
        e_in = relmse_trimmed(x, ref)
        den.append({"spp": spp, "time_s": t, "infer_s": t_inf, "relmse": e, "relmse_trimmed": et,
                    "relmse_trimmed_input": e_in})
        log(f"denoised {spp:3d} spp: {t * 1e3:8.2f} ms (inference {t_inf * 1e3:.2f} ms) relMSE {e:.5g} "
            f"trimmed {et:.5f} (input trimmed {e_in:.5f})")
    tp, how_p = time_to_target([(p["time_s"], p["relmse_trimmed"]) for p in plain])
    td, how_d = time_to_target([(p["time_s"], p["relmse_trimmed"]) for p in den])
    # every target either series reaches; the reference's own error (above)
    # bounds what differences are meaningful
    targets = {}
    for target in (0.01, 0.02, 0.03, 0.04, 0.05, 0.06, 0.08, 0.1):
        a = _reached([(p["time_s"], p["relmse_trimmed"]) for p in plain], target)
        b = _reached([(p["time_s"], p["relmse_trimmed"]) for p in den], target)
        targets[str(target)] = {"plain_s": a, "denoised_s": b,
                                "speedup": (a / b) if (a is not None and b is not None) else None}
    result = {
        "frame": f"{w}x{h} GGX/HDR-environment, depth 10",
        "camera_flags": flags,
        "reference_spp": 2 * ref_spp,
        "reference_error_relmse": floor,
        "reference_error_relmse_trimmed": floor_t,
        "time_to_targets": targets,
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
        "speedup": (tp / td) if (tp is not None and td is not None) else None,
    }
    log(f"time to trimmed relMSE {TARGET}: plain {tp} s ({how_p}), denoised {td} s ({how_d}); "
        f"the reference's own error, estimated from its two halves: trimmed relMSE {floor_t:.4f}")
    for k, v in targets.items():
        log(f"time to trimmed relMSE {k}: plain {v['plain_s']}, denoised {v['denoised_s']}, speedup {v['speedup']}")
    return result


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--weights", required=True)
    ap.add_argument("--ref-spp", type=int, default=16384)
    ap.add_argument("--json", default="")
    ap.add_argument("--reject-ggx-fallback", action="store_true",
                    help="render with GT_FLAG_GGX_REJECT_BELOW_HORIZON (the well-posed estimator)")
    args = ap.parse_args(argv)
    dn = Denoiser(api.default_context())
    dn.load(args.weights)
    flags = abi.GT_FLAG_GGX_REJECT_BELOW_HORIZON if args.reject_ggx_fallback else 0
    res = evaluate(dn, ref_spp=args.ref_spp, flags=flags, log=lambda m: print(m, flush=True))
    if args.json:
        with open(args.json, "w") as f:
            json.dump(res, f, indent=1)
    return 0


if __name__ == "__main__":
    sys.exit(main())
