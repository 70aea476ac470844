#!/usr/bin/env python3
"""Benchmark of the MI355X path tracer (one GPU).

A step is one full render of the workload. The default workload is the
1920x1080 GGX/environment scene: 1,101 spheres (100 GGX microfacet spheres,
roughness 0.02-1.0, metallic 0/1 by row, plus 1,000 random Dielectric, Metal
and Lambertian spheres on a ground sphere), lit by a 2048x1024 FP32 HDR sky
with a 0.5 degree sun of radiance 5e4, environment importance sampling with
MIS, 16 spp, max depth 10, Russian roulette from bounce 5.

Other workloads: cornell_256 (the Cornell box, 256x256, 64 spp stratified,
depth 8, NEE), terrain_720p (524,288-triangle Perlin terrain + 2,000
spheres, 1280x720, 32 spp, depth 6) and batch64 (64 seeded scenes of 1k-2M
triangles at 256^2 to 1080p, 8 spp, depth 8, NEE to 1-8 quad lights each,
all in one batched render with the tuning picked by the autotuner).

value = (extension rays + shadow rays traced) / wall time of the K timed
steps, in Mrays/s, with scenes, environment and outputs already resident in
HBM. The JSON line also carries the roofline of the dominant kernel (extend,
the closest-hit traversal) and a CPU baseline: the f64 CPU oracle (the
reference's algorithm, reference RNG stream) on a bounded row sample of the
first frame, timed on this host's cores.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E
# HBM bytes per extend launch measured with rocprofv3 --pmc FETCH_SIZE and
# WRITE_SIZE, scaled by the same counters on a plain copy of known size, with
# the L2 hit rate and shader-sequencer counters; each record carries the hash
# of the kernel sources it was measured with
TRAFFIC_FILE = ROOT / "profiles" / "extend_hbm_traffic.json"
KERNEL_SOURCES = ["gartus_amd/csrc/kernels.hip", "gartus_amd/csrc/kernels.h", "gartus_amd/csrc/device.h",
                  "gartus_amd/csrc/layout.h"]
WORKLOADS = ["ggx_env_1080p", "cornell_256", "terrain_720p", "batch64"]


def kernel_sources_sha() -> str:
    """First 16 hex digits of the SHA-256 of the tracer kernels' sources: a
    traffic record applies only to the kernels it was measured on."""
    import hashlib
    h = hashlib.sha256()
    for rel in KERNEL_SOURCES:
        h.update((ROOT / rel).read_bytes())
    return h.hexdigest()[:16]


def log(msg: str) -> None:
    print(msg, file=sys.stderr, flush=True)


def prim_bytes(frames) -> float:
    """Algorithmic bytes of one primitive test, averaged over the primitives:
    the geometry plus a 4-byte material id (triangle 36, quad 48, sphere 16,
    moving sphere 28)."""
    from gartus_amd.scene_data import MOVING_SPHERE, QUAD, SPHERE, TRIANGLE
    size = {SPHERE: 16 + 4, MOVING_SPHERE: 28 + 4, TRIANGLE: 36 + 4, QUAD: 48 + 4}
    n = b = 0
    for fr in frames:
        for k, d, _ in fr["data"].chunks:
            n += len(d)
            b += size[k] * len(d)
    return b / max(n, 1)


def _lights(targets):
    from gartus_amd import api
    L = api.SamplingTargetList()
    for k, q, w in targets:
        L._add(k, q, w)
    return L


def build_workload(name: str):
    """Returns (frames, description); a frame is a dict of data, lights,
    env, env_rgb, cam (camera spec) and targets."""
    from gartus_amd import api, scenes
    if name == "ggx_env_1080p":
        data, env, cam = scenes.ggx_environment()
        cam["nee"] = True
        return [dict(data=data, lights=None, env=api.EnvironmentLight.from_rgb32f(env), env_rgb=env, cam=cam,
                     targets=None)], "1920x1080 GGX/HDR-environment, 1101 spheres, 16 spp, depth 10"
    if name == "cornell_256":
        data, targets, cam = scenes.cornell()
        return [dict(data=data, lights=_lights(targets), env=None, env_rgb=None, cam=cam, targets=targets)], \
            "Cornell box 256x256, 64 spp stratified, depth 8, NEE"
    if name == "terrain_720p":
        data, cam = scenes.terrain()
        return [dict(data=data, lights=None, env=None, env_rgb=None, cam=cam, targets=None)], \
            "1280x720 Perlin terrain 524,288 triangles + 2,000 spheres, 32 spp, depth 6"
    if name == "batch64":
        frames = []
        for data, targets, cam in scenes.batch_scenes(64):
            frames.append(dict(data=data, lights=_lights(targets), env=None, env_rgb=None, cam=cam,
                               targets=targets))
        prims = sum(len(f["data"]) for f in frames)
        return frames, f"64 seeded scenes, {prims:,} primitives (1k-2M triangles each), 256^2-1080p, 8 spp, " \
                       f"depth 8, NEE, one batched render"
    raise SystemExit(f"unknown workload {name}")


def cpu_baseline(frames, seconds: float) -> dict:
    """The CPU oracle on a bounded sample of the workload: reference RNG
    stream, all host cores, 16x16 tiles (render_values_tiled). For each frame
    in order, a band of rows around the middle sized from a one-row pilot,
    until about `seconds` of rendering is spent (one frame is enough for a
    one-frame workload; a batch takes as many of its frames as fit)."""
    sys.path.insert(0, str(ROOT / "oracle"))
    import cpu_ref
    from gartus_amd import abi, scenes
    threads = cpu_ref.cpu_count()
    rays = paths = 0
    spent = 0.0
    parts = []
    for i, frame in enumerate(frames):
        scene = cpu_ref.Scene(frame["data"])
        env = cpu_ref.Env(rgb32f=frame["env_rgb"]) if frame["env_rgb"] is not None else None
        lights = cpu_ref.Lights(frame["targets"]) if frame["targets"] is not None else None
        cam = scenes.camera_record(frame["cam"])
        cam.rng_mode = abi.GT_RNG_PIXEL_STREAM
        W = cam.image_width
        H = max(int(W / cam.aspect_ratio), 1)
        y0 = H // 2
        t = time.perf_counter()
        _, c = scene.render(cam, lights, env, threads=threads, crop=(0, y0, W, 1))
        pilot = max(time.perf_counter() - t, 1e-4)
        rays += c["extension_rays"] + c["shadow_rays"]
        paths += c["paths"]
        spent += pilot
        left = seconds - spent
        rows = int(max(0, min(H - 1, left / pilot)))
        if rows > 0:
            t = time.perf_counter()
            _, c = scene.render(cam, lights, env, threads=threads, crop=(0, max(y0 - rows, 0), W, rows))
            spent += time.perf_counter() - t
            rays += c["extension_rays"] + c["shadow_rays"]
            paths += c["paths"]
        parts.append(f"{W}x{rows + 1}")
        if i == 0:  # one core on a 64-pixel-wide strip of the same frame
            one = one_core(scene, cam, lights, env, W, H, min(3.0, 0.25 * seconds))
        del scene
        if spent >= seconds:
            break
    what = (f"frame 0, {parts[0]} rows around the middle" if len(parts) == 1 else
            f"frames 0-{len(parts) - 1}, row bands {', '.join(parts)} around their middles")
    return {
        "value": rays / spent / 1e6,
        "unit": "Mrays/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{what}, full spp, f64 oracle, reference RNG stream; "
                  f"{paths / spent:.4g} samples/s over {spent:.1f} s ({rays} rays)",
        "cpu_model": cpu_model(),
        "one_core": one,
    }


def one_core(scene, cam, lights, env, W, H, seconds: float) -> dict:
    """The oracle on one thread: a 64-pixel-wide strip of rows around the
    middle of the frame, sized from a one-row pilot to about `seconds`."""
    w = min(64, W)
    x0, y0 = (W - w) // 2, H // 2
    t = time.perf_counter()
    _, c = scene.render(cam, lights, env, threads=1, crop=(x0, y0, w, 1))
    spent = max(time.perf_counter() - t, 1e-4)
    rays, paths = c["extension_rays"] + c["shadow_rays"], c["paths"]
    rows = int(max(0, min(y0, (seconds - spent) / spent)))
    if rows > 0:
        t = time.perf_counter()
        _, c = scene.render(cam, lights, env, threads=1, crop=(x0, y0 - rows, w, rows))
        spent += time.perf_counter() - t
        rays += c["extension_rays"] + c["shadow_rays"]
        paths += c["paths"]
    return {"value": rays / spent / 1e6, "unit": "Mrays/s", "samples_per_s": paths / spent,
            "sample": f"frame 0, {w}x{rows + 1} pixels around the middle, {spent:.1f} s"}


def time_progressive(ctx, scene, cam, lights, env, wh, tile: int, reps: int) -> dict:
    """Wall time of gt_render_progressive (render_world*_tiled_progressive,
    camera.rs:3172-3294) into a host canvas, one tile callback at a time, after
    one untimed call."""
    from gartus_amd import abi
    W, H = wh
    canvas = np.zeros((H, W, 3), np.uint8)
    count = [0]

    def on_tile(user, update):
        count[0] += 1
        return 0

    fn = abi.PROGRESS_FN(on_tile)
    ptr = canvas.ctypes.data_as(C.POINTER(C.c_uint8))
    args = (ctx.handle, scene, C.byref(cam), lights, env, tile, fn, None, ptr)
    abi.check(ctx.lib.gt_render_progressive(*args))
    count[0] = 0
    t0 = time.perf_counter()
    for _ in range(reps):
        abi.check(ctx.lib.gt_render_progressive(*args))
    ms = (time.perf_counter() - t0) / reps * 1e3
    return {"tile": tile, "ms_per_frame": round(ms, 3), "tiles": count[0] // reps}


def cpu_model() -> str:
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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


def _short_kernel(name: str) -> str:
    """'void gtk::(anonymous namespace)::k_extend<false, true, false>(gtk::Frame, ...)' -> 'k_extend<false, true, false>'."""
    name = name.replace("void ", "").replace("gtk::(anonymous namespace)::", "")
    depth = 0
    for i, ch in enumerate(name):
        depth += ch == "<"
        depth -= ch == ">"
        if ch == "(" and depth == 0:
            return name[:i]
    return name


def _counter_sums(path: Path) -> dict:
    """{kernel: {counter: [values per dispatch]}} of a rocprofv3 counter_collection.csv."""
    import csv
    out: dict = {}
    with open(path, newline="") as f:
        for row in csv.DictReader(f):
            out.setdefault(_short_kernel(row["Kernel_Name"]), {}).setdefault(row["Counter_Name"], []).append(
                float(row["Counter_Value"]))
    return out


# the counter passes behind a profile record, each a run of its own (OUT = the output folder)
PROFILE_PASSES = {
    "sq": "SQ_WAVE_CYCLES SQ_WAIT_ANY SQ_ACTIVE_INST_VALU SQ_THREAD_CYCLES_VALU SQ_INSTS_VALU SQ_BUSY_CYCLES SQ_WAVES "
          "GRBM_GUI_ACTIVE",
    "fetch": "FETCH_SIZE",
    "write": "WRITE_SIZE TCC_HIT_sum TCC_MISS_sum",
}
TUNING_ARGS = {"batch64": "--block 64 --bvh-width 4 --max-paths 67108864"}


def _profile_commands(workload: str) -> list[str]:
    base = "python -u bench.py --steps 1 --warmup 0 --no-cpu-baseline --no-denoise-eval --graph off"
    extra = (" " + TUNING_ARGS[workload]) if workload in TUNING_ARGS else ""
    cmds = [f"rocprofv3 --pmc {ctrs} --output-format csv -d OUT/{p}_{workload} -o run -- {base} --workload {workload}"
            f"{extra}" for p, ctrs in PROFILE_PASSES.items()]
    cmds.append(f"rocprofv3 --kernel-trace --stats --output-format csv -d OUT/trace_{workload} -o run -- python -u "
                f"bench.py --steps 5 --warmup 1 --no-cpu-baseline --no-denoise-eval --graph off --workload {workload}"
                f"{extra}")
    for c in ("FETCH_SIZE", "WRITE_SIZE"):
        cmds.append(f"rocprofv3 --pmc {c} --output-format csv -d OUT/cal_{c.split('_')[0].lower()} -o run -- "
                    f"python -u bench.py --calibrate-copy 1024 --steps 3")
    cmds.append(f"python bench.py --record-profile OUT --workload {workload}")
    return cmds


def record_profile(out: Path, workload: str, calibrate_mib: int = 1024, dest: Path | None = None) -> int:
    """Turns one profiling set of a workload (the passes of _profile_commands,
    under `out`) into its traffic record in profiles/extend_hbm_traffic.json,
    stamped with the kernel sources' hash, and a summary
    profiles/<workload>_counters.md. The extend kernel's figures average
    every timed (non-counting) extend dispatch, bounce 0 and later bounces
    alike, as bench.py's roofline.avg_launch_ms does."""
    import csv
    sq = _counter_sums(out / f"sq_{workload}" / "run_counter_collection.csv")
    fe = _counter_sums(out / f"fetch_{workload}" / "run_counter_collection.csv")
    wr = _counter_sums(out / f"write_{workload}" / "run_counter_collection.csv")
    cal_f = _counter_sums(out / "cal_fetch" / "run_counter_collection.csv")
    cal_w = _counter_sums(out / "cal_write" / "run_counter_collection.csv")

    def copy_mean(c, ctr):
        vals = [v for k, d in c.items() if "neg_kernel" in k for v in d.get(ctr, [])]
        return sum(vals) / len(vals)
    kib = calibrate_mib * 1024.0
    f_fetch = kib / copy_mean(cal_f, "FETCH_SIZE")
    f_write = kib / copy_mean(cal_w, "WRITE_SIZE")

    def family(prefix):
        return [k for k in fe if k.startswith(prefix)]

    def tot(c, kernels, ctr):
        return sum(sum(c.get(k, {}).get(ctr, [])) for k in kernels)

    def ndisp(c, kernels, ctr):
        return sum(len(c.get(k, {}).get(ctr, [])) for k in kernels)

    def row(kernels):
        n = ndisp(fe, kernels, "FETCH_SIZE")
        rd = tot(fe, kernels, "FETCH_SIZE") * f_fetch * 1024 / n
        wt = tot(wr, kernels, "WRITE_SIZE") * f_write * 1024 / n
        hit, miss = tot(wr, kernels, "TCC_HIT_sum"), tot(wr, kernels, "TCC_MISS_sum")
        wave = tot(sq, kernels, "SQ_WAVE_CYCLES")
        valu = tot(sq, kernels, "SQ_ACTIVE_INST_VALU")
        return {"dispatches": n, "read_bytes": rd, "write_bytes": wt,
                "l2_hit_rate": hit / (hit + miss) if hit + miss else None,
                "wave_wait_frac": tot(sq, kernels, "SQ_WAIT_ANY") / wave if wave else None,
                "valu_busy_frac": valu / wave if wave else None,
                "valu_lane_util": tot(sq, kernels, "SQ_THREAD_CYCLES_VALU") / (64.0 * valu) if valu else None}

    # the timed kernels: counting instantiations (first template argument true) left out
    timed = {k for k in fe if k.startswith(("k_extend<false", "k_shade", "k_connect<false", "k_accumulate"))}
    ext = sorted(k for k in timed if k.startswith("k_extend<false"))
    e = row(ext)
    # the same kernels' durations over the timed steps of the trace run
    trace = out / f"trace_{workload}"
    durs: dict = {}
    with open(trace / "run_kernel_trace.csv", newline="") as f:
        rows = sorted(csv.DictReader(f), key=lambda r: int(r["Dispatch_Id"]))
    for r in rows:
        durs.setdefault(_short_kernel(r["Kernel_Name"]), []).append(
            (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
    bench_line = None
    log_file = out / f"trace_{workload}.log"
    if log_file.exists():
        for line in log_file.read_text().splitlines():
            if line.startswith("{") and '"metric"' in line:
                bench_line = json.loads(line)
    if bench_line is None:
        raise SystemExit(f"{log_file}: no bench.py line (the trace run's output); it names the kernel sources")
    measured_sha = bench_line["roofline"].get("kernel_sources_sha16")
    if measured_sha != kernel_sources_sha():
        raise SystemExit(f"the profiles were measured on kernel sources {measured_sha}, these are "
                         f"{kernel_sources_sha()}: profile them again")
    steps, warmup = bench_line["steps"], bench_line["warmup"]
    renders = 1 + warmup + steps  # the statistics render, the warm-up, the timed steps (the counting render aside)
    ext_seq = [(int(r["Dispatch_Id"]), (int(r["End_Timestamp"]) - int(r["Start_Timestamp"])) * 1e-3)
               for r in rows if _short_kernel(r["Kernel_Name"]) in ext]
    per_render = len(ext_seq) // renders
    timed_ext = [d for _, d in ext_seq[-steps * per_render:]]
    avg_us = sum(timed_ext) / len(timed_ext)
    bytes_per_launch = e["read_bytes"] + e["write_bytes"]
    dest = dest or ROOT / "profiles"
    traffic_file = dest / TRAFFIC_FILE.name
    rec = json.loads(traffic_file.read_text()) if traffic_file.exists() else {}
    rec[workload] = {
        "kernel": "k_extend<false, *, *> (bounce 0 and later bounces)",
        "kernel_sources_sha16": kernel_sources_sha(),
        "bytes_per_launch": round(bytes_per_launch),
        "fetch_factor": round(f_fetch, 4),
        "write_factor": round(f_write, 4),
        "l2_hit_rate": round(e["l2_hit_rate"], 4),
        "wave_wait_frac": round(e["wave_wait_frac"], 4),
        "valu_busy_frac": round(e["valu_busy_frac"], 4),
        "valu_lane_util": round(e["valu_lane_util"], 4),
        "rocprof_avg_launch_us": round(avg_us, 1),
        "source": f"profiles/{workload}_counters.md",
        "commands": _profile_commands(workload),
    }
    traffic_file.write_text(json.dumps(rec, indent=1) + "\n")

    stats = {}
    with open(trace / "run_kernel_stats.csv", newline="") as f:
        for r in csv.DictReader(f):
            stats[_short_kernel(r["Name"])] = r
    lines = [f"# Counters per kernel: {workload}", "",
             "Each pass is a run of its own (no trace domains with `--pmc`), from the repository root on one "
             "MI355X with `TMPDIR=/tmp`; OUT is the output folder:", ""]
    lines += ["    " + c for c in _profile_commands(workload)]
    lines += ["", f"Calibration on a plain copy of {calibrate_mib} MiB read + {calibrate_mib} MiB written: "
              f"FETCH_SIZE factor {f_fetch:.4f}, WRITE_SIZE factor {f_write:.4f} (HBM bytes = counter KiB x "
              "1024 x factor). FETCH_SIZE counts what L2 requests from the memory side, so lines the "
              "Infinity Cache serves are counted too: an upper bound on HBM reads.", "",
              "Per dispatch means over the counter runs' timed renders (the counting render left out). Wait = "
              "SQ_WAIT_ANY / SQ_WAVE_CYCLES (share of wave time waiting on anything, mostly memory); VALU busy = "
              "SQ_ACTIVE_INST_VALU / SQ_WAVE_CYCLES; lane use = SQ_THREAD_CYCLES_VALU / (64 x "
              "SQ_ACTIVE_INST_VALU), the active-lane share of VALU issue (divergence). Durations: mean of the "
              "trace run's dispatches (all renders, from run_kernel_stats.csv).", "",
              "| kernel | dispatches | mean us | HBM read MB | HBM write MB | GB/s | L2 hit | wait | VALU busy | "
              "lane use |", "|---|---:|---:|---:|---:|---:|---:|---:|---:|---:|"]
    order = sorted(timed, key=lambda k: -float(stats.get(k, {}).get("TotalDurationNs", 0) or 0))
    for k in order:
        r = row([k])
        us = float(stats[k]["AverageNs"]) * 1e-3 if k in stats else float("nan")
        gbs = (r["read_bytes"] + r["write_bytes"]) / (us * 1e-6) / 1e9 if us == us else float("nan")
        f3 = (lambda v: "-" if v is None else f"{v:.3f}")
        lines.append(f"| {k} | {r['dispatches']} | {us:.1f} | {r['read_bytes'] / 1e6:.1f} | "
                     f"{r['write_bytes'] / 1e6:.1f} | {gbs:.0f} | {f3(r['l2_hit_rate'])} | "
                     f"{f3(r['wave_wait_frac'])} | {f3(r['valu_busy_frac'])} | {f3(r['valu_lane_util'])} |")
    lines += ["", "## The roofline kernel (extend)", "",
              f"- HBM traffic per launch: {bytes_per_launch:,.0f} bytes (reads {e['read_bytes']:,.0f}, writes "
              f"{e['write_bytes']:,.0f}), over {e['dispatches']} dispatches; L2 hit rate {e['l2_hit_rate']:.3f}.",
              f"- Shader sequencer: waves wait {100 * e['wave_wait_frac']:.1f} % of their cycles, issue VALU "
              f"{100 * e['valu_busy_frac']:.1f} %, with {100 * e['valu_lane_util']:.1f} % of lanes active.",
              f"- Launch time over the trace run's {steps} timed steps ({len(timed_ext)} dispatches): rocprofv3 "
              f"{avg_us:.1f} us; bench.py's own HIP events in the same run: roofline.avg_launch_ms = "
              f"{1e3 * bench_line['roofline']['avg_launch_ms']:.1f} us; measured HBM rate {bytes_per_launch / (avg_us * 1e-6) / 1e9:,.0f} GB/s, "
              f"{bytes_per_launch / (avg_us * 1e-6) / 1e9 / HBM_PEAK_GBS:.3f} of the 8 TB/s peak.",
              f"- Kernel sources: sha256[:16] {kernel_sources_sha()} of {', '.join(KERNEL_SOURCES)}."]
    lines += ["", "bench.py's line in the trace run (eager, instrumented):", "",
              f"    {bench_line['value']} Mrays/s, {bench_line['ms_per_step']} ms per step, "
              f"{bench_line['config']['workload']}"]
    (dest / f"{workload}_counters.md").write_text("\n".join(lines) + "\n")
    # the trace run's per-kernel statistics
    t = [f"# Kernel trace: {workload}", "", "Command (from the repository root, on one MI355X, `TMPDIR=/tmp`):", "",
         "    " + _profile_commands(workload)[3], "",
         f"The run renders the workload {renders + 1} times: a statistics render, {warmup} warm-up, {steps} timed "
         "steps and a traversal-counting render (the instantiations whose first template argument is true). "
         "k_extend<COUNT, MULTI, CAMERA>, k_connect<COUNT, MULTI>, k_shade<MULTI, CAMERA>; MULTI is the batched "
         "(frame-table) form, CAMERA the bounce of the camera rays. Durations in microseconds, from "
         "run_kernel_stats.csv.", "", "| kernel | calls | total us | average us | % |", "|---|---:|---:|---:|---:|"]
    for k, r in sorted(stats.items(), key=lambda kv: -float(kv[1]["TotalDurationNs"])):
        t.append(f"| {k} | {r['Calls']} | {float(r['TotalDurationNs']) * 1e-3:.1f} | {float(r['AverageNs']) * 1e-3:.2f} "
                 f"| {float(r['Percentage']):.2f} |")
    t += ["", f"Extend over the {steps} timed steps ({len(timed_ext)} dispatches, the last {steps} renders before "
          f"the counting one): rocprofv3 {avg_us:.1f} us per launch; bench.py's HIP events in the same run "
          f"(roofline.avg_launch_ms) {1e3 * bench_line['roofline']['avg_launch_ms']:.1f} us.", "",
          "bench.py's per-step lines in the same run:", ""]
    t += ["    " + ln for ln in log_file.read_text().splitlines() if ln.startswith(("step ", "eager:", "graph"))]
    (dest / f"{workload}_kernel_trace.md").write_text("\n".join(t) + "\n")
    log(f"wrote {traffic_file} [{workload}] and {dest / (workload + '_counters.md')}")
    return 0


DENOISER_WEIGHTS = ROOT / "gartus_amd" / "weights" / "denoiser_c3.npy"


def denoise_summary(ctx) -> dict:
    """Time to relMSE targets on this frame, plain path tracing against path
    tracing + the trained denoiser (gartus_amd.denoise_eval): scored on 16
    fixed 64x64 crops against their committed 2 x 2^20-spp FP64 reference,
    with GT_FLAG_GGX_REJECT_BELOW_HORIZON (the reference's own estimator has
    unbounded variance on this scene)."""
    from gartus_amd.denoise_eval import evaluate
    from gartus_amd.denoiser import Denoiser
    dn = Denoiser.from_file(str(DENOISER_WEIGHTS), ctx)
    r = evaluate(dn, log=log)
    dn.close()
    t01 = r["time_to_targets"].get("0.01", {})
    full = t01.get("relmse", {}) if t01.get("measurable") else {}
    return {
        "metric": "time to relMSE (eps 1e-2) on 16 fixed 64x64 crops of the frame, full and trimmed (0.1%)",
        "estimator": "GT_FLAG_GGX_REJECT_BELOW_HORIZON",
        "reference_spp": r["reference_spp"],
        "reference_error_relmse": round(r["reference_error_relmse"], 6),
        "reference_error_relmse_trimmed": round(r["reference_error_relmse_trimmed"], 6),
        "time_to_targets_s": r["time_to_targets"],
        "relmse_0.01": {"plain_s": full.get("plain_s"), "denoised_s": full.get("denoised_s"),
                        "denoised_over_plain": full.get("denoised_over_plain"), "target_max_ratio": 0.125},
        "plain": [{k: p[k] for k in ("spp", "time_s", "relmse", "relmse_trimmed")} for p in r["plain"]],
        "denoised": [{k: d[k] for k in ("spp", "time_s", "infer_s", "relmse", "relmse_trimmed", "relmse_input")}
                     for d in r["denoised"]],
    }


def traffic_for(workload: str) -> dict:
    """The extend kernel's measured traffic record for the workload (bytes
    per launch below L2, L2 hit rate, shader counters), or an empty dict when
    there is none or it was measured on other kernel sources."""
    try:
        rec = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return {}
    r = rec.get(workload) or {}
    if r.get("kernel_sources_sha16") != kernel_sources_sha():
        return {"stale": r.get("kernel_sources_sha16")} if r else {}
    return r


def roofline_record(measured: dict, traffic, hbm_gbs, algorithmic_gbs: float, alg_bytes: float,
                    avg_launch_ms: float) -> dict:
    """The extend kernel against the HBM roofline. achieved = HBM bytes the
    kernel moves per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-
    calibrated, measured on these kernel sources) over its live launch time.
    The algorithmic bytes of the traversal (ray in, hit out, every node and
    primitive fetch counted as if it came from memory) are reported beside
    it: where the scene fits in the caches they are served by L2 and exceed
    what HBM sees. Without a current traffic record the algorithmic figure
    stands in, and `basis` says so."""
    rec = {"bound": "hbm"}
    if hbm_gbs is not None:
        rec.update(achieved=round(hbm_gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s", frac=round(hbm_gbs / HBM_PEAK_GBS, 4),
                   traffic=traffic, basis="measured HBM traffic per launch / live launch time")
    else:
        rec.update(achieved=round(algorithmic_gbs, 1), peak=HBM_PEAK_GBS, unit="GB/s",
                   frac=round(algorithmic_gbs / HBM_PEAK_GBS, 4), traffic=None,
                   basis="algorithmic bytes (no cache reuse) / live launch time: no traffic record measured on "
                         "these kernel sources" + (f" (record is for {measured['stale']})" if "stale" in measured
                                                   else ""))
    rec.update({
        "kernel": "k_extend (closest-hit traversal)",
        "kernel_sources_sha16": kernel_sources_sha(),
        "avg_launch_ms": round(avg_launch_ms, 4),
        "algorithmic_bytes_per_launch": round(alg_bytes),
        "algorithmic_gbs": round(algorithmic_gbs, 1),
        "l2_hit_rate": measured.get("l2_hit_rate"),
        # shader-sequencer view of the same kernel (same rocprofv3 runs): share of wave cycles spent
        # waiting on memory, VALU issue share, active-lane share of VALU instructions
        "wave_wait_frac": measured.get("wave_wait_frac"),
        "valu_busy_frac": measured.get("valu_busy_frac"),
        "valu_lane_util": measured.get("valu_lane_util"),
        "counters_source": measured.get("source"),
    })
    rec["limiter"] = limiter(rec)
    return rec


def limiter(rec: dict) -> str | None:
    """What the counters say bounds the kernel: HBM bandwidth when the
    measured traffic runs near peak; otherwise, when waves spend most of
    their cycles waiting with little VALU issue, the latency of dependent
    memory fetches; None without counters."""
    wait, busy = rec.get("wave_wait_frac"), rec.get("valu_busy_frac")
    if rec.get("traffic") is None or wait is None or busy is None:
        return None
    if rec["frac"] >= 0.7:
        return "HBM bandwidth"
    if wait >= 0.5 and busy < 0.25:
        return (f"latency of dependent memory fetches: HBM at {rec['frac']:.2f} of peak, waves waiting "
                f"{100 * wait:.0f}% of their cycles, VALU issuing {100 * busy:.0f}%")
    return f"mixed: HBM at {rec['frac']:.2f} of peak, waves waiting {100 * wait:.0f}%, VALU {100 * busy:.0f}%"


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--workload", default="ggx_env_1080p", choices=WORKLOADS)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--block", type=int, default=0)
    ap.add_argument("--bvh-width", type=int, default=0)
    ap.add_argument("--max-paths", type=int, default=0)
    ap.add_argument("--stack-lds", type=int, default=0, help="traversal-stack entries per lane in LDS (0: default)")
    ap.add_argument("--graph", choices=["auto", "on", "off"], default="auto",
                    help="time the render as a captured hipGraph replay (on), eagerly (off), or both with the "
                         "library's default as the value (auto)")
    ap.add_argument("--adaptive", nargs=3, type=float, metavar=("MIN", "MAX", "THRESHOLD"),
                    help="single-frame workloads: render with adaptive sampling (camera.rs:2937-2982)")
    ap.add_argument("--progressive", type=int, default=0, metavar="TILE",
                    help="single-frame workloads: also time gt_render_progressive with tiles of TILE pixels")
    ap.add_argument("--device-build", action="store_true",
                    help="compile the scenes on the GPU (gt_scene_build_device) instead of the host SAH build")
    ap.add_argument("--autotune", action="store_true", help="pick the tuning with gt_autotune (batch64: always)")
    ap.add_argument("--no-denoise-eval", action="store_true",
                    help="skip the time-to-relMSE comparison of plain and denoised rendering (ggx_env_1080p)")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="run only a plain device copy of MIB MiB (counter calibration)")
    ap.add_argument("--summarize-counters", nargs="+", metavar="CSV",
                    help="print per-kernel counter means from rocprofv3 CSV files and exit")
    ap.add_argument("--record-profile", metavar="OUT",
                    help="turn the workload's rocprofv3 passes under OUT into its traffic record and counter "
                         "summary under profiles/, and exit")
    args = ap.parse_args()
    if args.summarize_counters:
        return summarize_counters(args.summarize_counters)
    if args.record_profile:
        return record_profile(Path(args.record_profile), args.workload)
    if args.calibrate_copy:
        return calibrate_copy(args.calibrate_copy, max(args.steps, 1))
    if args.gpus != 1:
        print(f"bench.py measures one GPU only; --gpus {args.gpus} is not supported (use --gpus 1)", file=sys.stderr)
        return 2

    from gartus_amd import abi, api, scenes
    t0 = time.perf_counter()
    frames, desc = build_workload(args.workload)
    log(f"workload {args.workload}: {desc}; generated in {time.perf_counter() - t0:.1f} s")
    ctx = api.Context(0)
    kw = {}
    if args.block:
        kw["block_size"] = args.block
    if args.bvh_width:
        kw["bvh_width"] = args.bvh_width
    if args.max_paths:
        kw["max_paths"] = args.max_paths
    if args.stack_lds:
        kw["stack_lds"] = args.stack_lds
    if kw:
        ctx.set_tuning(**kw)
    t0 = time.perf_counter()
    n = len(frames)
    lib = ctx.lib
    scenes_h = (C.c_void_p * n)()
    lights_h = (C.c_void_p * n)()
    envs_h = (C.c_void_p * n)()
    cams = (abi.GtCamera * n)()
    outs = (C.c_void_p * n)()
    npaths = 0
    prims = nodes = 0
    if args.device_build:  # one context stream: one scene after another
        built = [api.RayScene(fr["data"], ctx).with_device_bvh() for fr in frames]
    else:  # the host SAH builds run in parallel: ctypes drops the GIL
        with ThreadPoolExecutor(max_workers=min(16, os.cpu_count() or 1)) as ex:
            built = list(ex.map(lambda fr: api.RayScene(fr["data"], ctx).with_bvh(), frames))
    build_wall = time.perf_counter() - t0
    build_device_ms = sum(rs.info().build_ms for rs in built)
    for i, fr in enumerate(frames):
        rs = built[i]
        fr["scene"] = rs
        info = rs.info()
        prims += info.primitive_count
        nodes += info.bvh_node_count
        scenes_h[i] = rs._compile()
        lights_h[i] = fr["lights"]._compile(ctx) if fr["lights"] is not None else None
        envs_h[i] = fr["env"]._compile(ctx) if fr["env"] is not None else None
        cams[i] = scenes.camera_record(fr["cam"])
        if args.adaptive:  # with_adaptive_sampling (camera.rs:1244-1264): spp becomes max
            mn = max(int(args.adaptive[0]), 1)
            mx = max(int(args.adaptive[1]), mn)
            cams[i].adaptive_min_samples, cams[i].adaptive_max_samples = mn, mx
            cams[i].adaptive_error_threshold = args.adaptive[2]
            cams[i].samples_per_pixel = mx
            cams[i].pixel_sample_mode = abi.GT_PIXEL_RANDOM
        W, H = cams[i].image_width, lib.gt_camera_image_height(C.byref(cams[i]))
        spp = lib.gt_camera_effective_spp(C.byref(cams[i]))
        npaths += W * H * spp
        fr["wh"] = (W, H)
        d = C.c_void_p()
        abi.check(lib.gt_device_malloc(ctx.handle, W * H * 3 * 4, C.byref(d)))
        outs[i] = d
    log(f"{n} scene(s) compiled in {build_wall:.2f} s ({'device LBVH, ' + format(build_device_ms, '.1f') + ' ms on the GPU' if args.device_build else 'host SAH'}): "
        f"{prims:,} prims, {nodes:,} BVH nodes")
    tuned = None
    explicit = bool(args.block or args.bvh_width or args.max_paths or args.stack_lds)
    if args.autotune or (args.workload == "batch64" and not explicit):
        best = abi.GtTuning()
        ms = C.c_double()
        t0 = time.perf_counter()
        abi.check(lib.gt_autotune(ctx.handle, n, scenes_h, cams, lights_h, envs_h, 0, C.byref(best), C.byref(ms)))
        tuned = {"block_size": best.block_size, "bvh_width": best.bvh_width, "max_paths": best.max_paths,
                 "stack_lds": best.stack_lds}
        log(f"autotune ({time.perf_counter() - t0:.1f} s): {tuned}, {ms.value:.2f} ms per render")
    default_graph = bool(ctx.tuning.use_graph)

    def render(stats: bool = True, count: bool = False, graph: bool = False):
        ctx.set_tuning(count_traversal=1 if count else 0, use_graph=1 if graph and not count else 0)
        st = abi.GtRenderStats()
        sp = C.byref(st) if stats else None
        if n == 1:
            abi.check(lib.gt_render_hdr_device(ctx.handle, scenes_h[0], C.byref(cams[0]), lights_h[0], envs_h[0],
                                               outs[0], 0, sp))
        else:
            abi.check(lib.gt_render_batch_device(ctx.handle, n, scenes_h, cams, lights_h, envs_h, outs, 0, sp))
        return st

    def timed(graph: bool):
        """warmup, then the K timed steps between two synchronisations; eager
        steps are instrumented (per-kernel HIP events), graph replays are not
        (a render that asks for statistics is never captured). The library
        captures a launch sequence the second time it sees it, so graph mode
        warms up with one more render: capture and instantiation stay out of
        the timed steps."""
        for i in range(args.warmup + (1 if graph else 0)):
            st = render(stats=not graph, graph=graph)
            if not graph:
                log(f"warmup {i}: {st.device_ms:.1f} ms device")
        ctx.synchronize()
        t0 = time.perf_counter()
        sts = [render(stats=not graph, graph=graph) for _ in range(args.steps)]
        ctx.synchronize()
        return time.perf_counter() - t0, sts

    # one instrumented render gives the ray counts and per-kernel times (the
    # counts are deterministic)
    ref = render(stats=True)
    modes = {"off": [False], "on": [True], "auto": [False, True]}[args.graph]
    walls = {}
    stats = [ref]
    for graph in modes:
        w, sts = timed(graph)
        walls[graph] = w
        if not graph:
            stats = sts
            for i, st in enumerate(sts):
                log(f"step {i}: {st.device_ms:.1f} ms device, extend {st.extend_ms:.1f} ms, "
                    f"shade {st.shade_ms:.1f} ms, connect {st.connect_ms:.1f} ms")
        log(f"{'graph replay' if graph else 'eager'}: {w / args.steps * 1e3:.3f} ms per step")
    use_graph = {"off": False, "on": True, "auto": default_graph}[args.graph]
    ctx.set_tuning(use_graph=1 if default_graph else 0)
    wall = walls[use_graph]
    rays_step = ref.extension_rays + ref.shadow_rays
    rays = rays_step * args.steps
    ext_rays = ref.extension_rays
    value = rays / wall / 1e6
    cst = render(count=True)
    assert cst.extension_rays == ext_rays, "renders are not deterministic"
    launches = ref.extend_launches
    pb = prim_bytes(frames)
    ext_bytes = ext_rays * (32 + 16) + 64 * cst.ext_node_fetches + pb * cst.ext_prim_tests
    avg_launch_ms = sum(s.extend_ms for s in stats) / sum(s.extend_launches for s in stats)
    algorithmic_gbs = (ext_bytes / launches) / (avg_launch_ms * 1e-3) / 1e9
    measured = traffic_for(args.workload)
    traffic = measured.get("bytes_per_launch")
    hbm_gbs = None if traffic is None else traffic / (avg_launch_ms * 1e-3) / 1e9
    if args.adaptive:  # paths traced, not W * H * max
        npaths = ref.camera_paths
    progressive = None
    if args.progressive and n == 1:
        progressive = time_progressive(ctx, scenes_h[0], cams[0], lights_h[0], envs_h[0], frames[0]["wh"],
                                       args.progressive, max(args.steps, 1))
        progressive["plain_ms_per_frame"] = round(wall / args.steps * 1e3, 3)
        log(f"progressive, tiles of {args.progressive}: {progressive['ms_per_frame']:.2f} ms per frame "
            f"({progressive['tiles']} tile callbacks), plain {progressive['plain_ms_per_frame']:.2f} ms")
    W0, H0 = frames[0]["wh"]
    out = np.zeros(W0 * H0 * 3, np.float32)
    abi.check(lib.gt_memcpy_to_host(ctx.handle, out.ctypes.data_as(C.c_void_p), outs[0], out.nbytes))
    log(f"frame 0 mean {out.reshape(-1, 3).mean(0)}; samples/s {npaths * args.steps / wall:.4g}; "
        f"ext nodes/ray {cst.ext_node_fetches / ext_rays:.2f} prims/ray {cst.ext_prim_tests / ext_rays:.2f}")
    for i in range(n):
        lib.gt_device_free(ctx.handle, outs[i])
    cam0 = cams[0]
    config = {"workload": desc, "frames": n, "primitives": int(prims), "paths_per_step": int(npaths),
              "rays_per_step": int(rays_step), "samples_per_s": round(npaths * args.steps / wall, 1),
              "hipgraph": use_graph,
              "ms_per_step_eager": round(walls[False] / args.steps * 1e3, 3) if False in walls else None,
              "ms_per_step_graph": round(walls[True] / args.steps * 1e3, 3) if True in walls else None,
              "scene_compile": {"builder": "device LBVH" if args.device_build else "host SAH",
                                "wall_s": round(build_wall, 3),
                                "device_ms": round(build_device_ms, 2) if args.device_build else None}}
    if n == 1:
        config.update(width=W0, height=H0, spp=int(lib.gt_camera_effective_spp(C.byref(cam0))),
                      max_depth=int(cam0.max_depth))
    if tuned:
        config["autotuned"] = tuned
    if args.adaptive:
        config["adaptive"] = {"min_samples": int(cam0.adaptive_min_samples),
                              "max_samples": int(cam0.adaptive_max_samples),
                              "threshold": float(cam0.adaptive_error_threshold),
                              "camera_paths": int(ref.camera_paths), "accepted_samples": int(ref.accepted_samples)}
    if progressive:
        config["progressive"] = progressive
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
        "data": "synthetic (seeded procedural scenes and sky)",
        "config": config,
        "roofline": roofline_record(measured, traffic, hbm_gbs, algorithmic_gbs, ext_bytes / launches,
                                    avg_launch_ms),
    }
    from gartus_amd.denoise_eval import REFERENCE_FILE
    if (args.workload == "ggx_env_1080p" and not args.no_denoise_eval and DENOISER_WEIGHTS.exists()
            and REFERENCE_FILE.exists()):
        log("denoiser: time to relMSE, plain against denoised (crop reference, 2 x 2^20 spp) ...")
        result["denoise"] = denoise_summary(ctx)
    if not args.no_cpu_baseline:
        log("CPU baseline ...")
        result["cpu_baseline"] = cpu_baseline(frames, args.cpu_seconds)
    else:
        result["cpu_baseline"] = None
    print(json.dumps(result), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
