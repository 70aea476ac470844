#!/usr/bin/env python3
"""Benchmark of the MI355X path tracer (one GPU).

A step is one full render of the workload frame. The default workload is the
1920x1080 GGX/environment scene: 1,101 spheres (100 GGX microfacet spheres,
roughness 0.02-1.0, metallic 0/1 by row, plus 1,000 random Dielectric, Metal
and Lambertian spheres on a ground sphere), lit by a 2048x1024 FP32 HDR sky
with a 0.5 degree sun of radiance 5e4, environment importance sampling with
MIS, 16 spp, max depth 10, Russian roulette from bounce 5.

value = (extension rays + shadow rays traced) / wall time of the K timed
steps, in Mrays/s, with the scene, environment and output already resident
in HBM. The JSON line also carries the roofline of the dominant kernel
(extend, the closest-hit traversal) and a CPU baseline: the f64 CPU oracle
(the reference's algorithm, reference RNG stream) on a bounded row sample of
the same frame, timed on this host's cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# HBM bytes per extend launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE, scaled by the same counters on a plain copy of known size
TRAFFIC_FILE = ROOT / "profiles" / "extend_hbm_traffic.json"


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def prim_bytes(data) -> float:
    """Algorithmic bytes of one primitive test: the geometry plus a 4-byte
    material id (triangle 36, quad 48, sphere 16, moving sphere 28)."""
    from gartus_amd.scene_data import MOVING_SPHERE, QUAD, SPHERE, TRIANGLE
    size = {SPHERE: 16 + 4, MOVING_SPHERE: 28 + 4, TRIANGLE: 36 + 4, QUAD: 48 + 4}
    n = sum(len(d) for _, d, _ in data.chunks)
    return sum(size[k] * len(d) for k, d, _ in data.chunks) / max(n, 1)


def build_workload(name: str):
    from gartus_amd import api, scenes
    if name == "ggx_env_1080p":
        data, env, cam = scenes.ggx_environment()
        return data, None, api.EnvironmentLight.from_rgb32f(env), env, cam, \
            "1920x1080 GGX/HDR-environment, 1101 spheres, 16 spp, depth 10"
    if name == "cornell_256":
        data, lights, cam = scenes.cornell()
        L = api.SamplingTargetList()
        for k, q, w in lights:
            L._add(k, q, w)
        return data, L, None, None, cam, "Cornell box 256x256, 64 spp stratified, depth 8, NEE"
    if name == "terrain_720p":
        data, cam = scenes.terrain()
        return data, None, None, None, cam, "1280x720 Perlin terrain 524,288 triangles + 2,000 spheres, 32 spp, depth 6"
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(data, env_rgb, lights_gpu, cam_spec, seconds: float) -> dict:
    """The CPU oracle on a bounded row sample of the same frame: reference
    RNG stream, all host cores, 16x16 tiles (render_values_tiled)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import cpu_ref
    from gartus_amd import abi, scenes
    scene = cpu_ref.Scene(data)
    env = cpu_ref.Env(rgb32f=env_rgb) if env_rgb is not None else None
    lights = cpu_ref.Lights(lights_gpu.targets) if lights_gpu is not None else None
    cam = scenes.camera_record(cam_spec)
    cam.rng_mode = abi.GT_RNG_PIXEL_STREAM
    threads = cpu_ref.cpu_count()
    W = cam.image_width
    H = max(int(W / cam.aspect_ratio), 1)
    # pilot: one row near the middle, then enough rows for ~`seconds`
    y0 = H // 2
    t = time.perf_counter()
    _, c1 = scene.render(cam, lights, env, threads=threads, crop=(0, y0, W, 1))
    pilot = max(time.perf_counter() - t, 1e-3)
    rows = int(max(1, min(H - y0, seconds / pilot)))
    t = time.perf_counter()
    _, c = scene.render(cam, lights, env, threads=threads, crop=(0, y0 - rows // 2, W, rows))
    dt = time.perf_counter() - t
    rays = c["extension_rays"] + c["shadow_rays"]
    return {
        "value": rays / dt / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{W}x{rows} rows centred on the frame, full spp, f64 oracle, reference RNG stream; "
                  f"{c['paths'] / dt:.4g} samples/s over {dt:.1f} s ({rays} rays)",
    }


def calibrate_copy(mib: int, reps: int) -> int:
    """A plain device copy of known size (out = -in over `mib` MiB of
    float32): profiled with the same counters as the extend kernel, it gives
    the factor between counter readings and bytes moved."""
    import torch
    n = mib * (1 << 20) // 4
    a = torch.ones(n, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(reps):
        torch.neg(a, out=b)
    torch.cuda.synchronize()
    log(f"copy calibration: {reps} x neg over {mib} MiB read + {mib} MiB written")
    return 0


def summarize_counters(paths: list[str]) -> int:
    """Per-kernel mean of each counter over the dispatches in rocprofv3 CSV
    counter files (counter_collection.csv)."""
    import csv
    acc: dict[tuple[str, str], list[float]] = {}
    for p in paths:
        with open(p, newline="") as f:
            for row in csv.DictReader(f):
                key = (row["Kernel_Name"], row["Counter_Name"])
                acc.setdefault(key, []).append(float(row["Counter_Value"]))
    out = {f"{k} | {c}": {"dispatches": len(v), "mean": sum(v) / len(v)} for (k, c), v in sorted(acc.items())}
    print(json.dumps(out, indent=1))
    return 0


def traffic_for(workload: str):
    try:
        rec = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    w = rec.get(workload)
    return None if w is None else w.get("bytes_per_launch")


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ggx_env_1080p", choices=["ggx_env_1080p", "cornell_256", "terrain_720p"])
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--bvh-width", type=int, default=0)
    ap.add_argument("--max-paths", type=int, default=0)
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (counter calibration)")
    ap.add_argument("--summarize-counters", nargs="+", metavar="CSV",
                    help="print per-kernel counter means from rocprofv3 CSV files and exit")
    args = ap.parse_args()
    if args.summarize_counters:
        return summarize_counters(args.summarize_counters)
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy, max(args.steps, 1))
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (use --gpus 1)", file=sys.stderr)
        return 2

    from gartus_amd import abi, api, scenes
    t0 = time.perf_counter()
    data, lights, env, env_rgb, cam_spec, desc = build_workload(args.workload)
    log(f"workload {args.workload}: {desc}; generated in {time.perf_counter() - t0:.1f} s")
    ctx = api.Context(0)
    kw = {}
    if args.block:
        kw["block_size"] = args.block
    if args.bvh_width:
        kw["bvh_width"] = args.bvh_width
    if args.max_paths:
        kw["max_paths"] = args.max_paths
    if kw:
        ctx.set_tuning(**kw)
    t0 = time.perf_counter()
    rs = api.RayScene(data, ctx).with_bvh()
    info = rs.info()
    log(f"scene compiled in {time.perf_counter() - t0:.2f} s: {info.primitive_count} prims, "
        f"{info.bvh_node_count} nodes, depth {info.bvh_depth}, {info.device_bytes / 1e6:.1f} MB on device")
    camera = api.RayCamera.__new__(api.RayCamera)
    camera.raw = scenes.camera_record(cam_spec)
    if env is not None or lights is not None:
        camera = camera.with_direct_lighting_mode(abi.GT_LIGHTING_NEE)
    W, H = camera.image_width(), camera.image_height()
    spp = camera.effective_samples_per_pixel()
    lib = ctx.lib
    sh = rs._compile()
    lh = lights._compile(ctx) if lights is not None else None
    eh = env._compile(ctx) if env is not None else None
    dout = C.c_void_p()
    abi.check(lib.gt_device_malloc(ctx.handle, W * H * 3 * 4, C.byref(dout)))

    def render(count: bool = False):
        if count:
            ctx.set_tuning(count_traversal=1)
        st = abi.GtRenderStats()
        abi.check(lib.gt_render_hdr_device(ctx.handle, sh, C.byref(camera.raw), lh, eh, dout, 0, C.byref(st)))
        if count:
            ctx.set_tuning(count_traversal=0)
        return st

    for i in range(args.warmup):
        st = render()
        log(f"warmup {i}: {st.device_ms:.1f} ms device, {st.extension_rays + st.shadow_rays} rays")
    ctx.synchronize()
    t0 = time.perf_counter()
    stats = [render() for _ in range(args.steps)]
    ctx.synchronize()
    wall = time.perf_counter() - t0
    for i, st in enumerate(stats):
        log(f"step {i}: {st.device_ms:.1f} ms device, extend {st.extend_ms:.1f} ms, shade {st.shade_ms:.1f} ms, "
            f"connect {st.connect_ms:.1f} ms")
    rays = sum(s.extension_rays + s.shadow_rays for s in stats)
    ext_rays = stats[0].extension_rays
    value = rays / wall / 1e6
    # traversal counts are deterministic: one counting render gives them
    cst = render(count=True)
    assert cst.extension_rays == ext_rays, "renders are not deterministic"
    launches = stats[0].extend_launches
    pb = prim_bytes(data)
    ext_bytes = ext_rays * (32 + 16) + 64 * cst.ext_node_fetches + pb * cst.ext_prim_tests
    avg_launch_ms = sum(s.extend_ms for s in stats) / sum(s.extend_launches for s in stats)
    achieved = (ext_bytes / launches) / (avg_launch_ms * 1e-3) / 1e9
    paths = stats[0].camera_paths
    traffic = traffic_for(args.workload)
    out = np.zeros(W * H * 3, np.float32)
    abi.check(lib.gt_memcpy_to_host(ctx.handle, out.ctypes.data_as(C.c_void_p), dout, out.nbytes))
    log(f"image mean {out.reshape(-1, 3).mean(0)}; samples/s {paths * args.steps / wall:.4g}; "
        f"ext nodes/ray {cst.ext_node_fetches / ext_rays:.2f} prims/ray {cst.ext_prim_tests / ext_rays:.2f}")
    lib.gt_device_free(ctx.handle, dout)
    result = {
        "metric": "Mrays/s",
        "value": round(value, 2),
        "unit": "Mrays/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded procedural scene and sky)",
        "config": {
            "workload": desc,
            "width": W, "height": H, "spp": spp, "max_depth": camera.max_depth(),
            "primitives": int(info.primitive_count),
            "paths_per_step": int(paths),
            "rays_per_step": int(rays // args.steps),
            "samples_per_s": round(paths * args.steps / wall, 1),
        },
        "roofline": {
            "bound": "hbm",
            "achieved": round(achieved, 1),
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4),
            "traffic": traffic,
            "kernel": "k_extend (closest-hit traversal)",
            "bytes_per_launch": round(ext_bytes / launches),
            "avg_launch_ms": round(avg_launch_ms, 4),
            # HBM bytes actually moved per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)
            # over the live launch time: the BVH and primitives are served from L2, so this is
            # far below the algorithmic figure
            "hbm_measured_gbs": None if traffic is None else round(traffic / (avg_launch_ms * 1e-3) / 1e9, 1),
        },
    }
    if not args.no_cpu_baseline:
        log("CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(data, env_rgb, lights, cam_spec, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
