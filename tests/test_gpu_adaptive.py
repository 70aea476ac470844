"""Adaptive sampling on the GPU (render_world_pixel_adaptive,
camera.rs:2937-2982) against the CPU oracle in counter-stream mode.

The GPU renders adaptive frames in rounds over the pixels still sampling and
runs the Welford recurrence over each round's samples in sample order, so a
pixel stops at the same sample as the sequential loop whenever its sample
values are the same. Scenes whose samples are exactly 0 or 1 (an emitter seen
directly, depth 1) make the values identical on both sides, and there the
images must agree bit for bit except where FP32 and FP64 disagree on whether
a ray grazing the emitter's edge hits it (at most 0.5% of the pixels), and
the number of samples the pixels used must agree within 0.5%. (The GPU
traces whole rounds, so it traces more samples than it uses.) A lit Cornell
box checks the image mean within 3 standard errors and the sample count
within 10%.
"""
import sys

import numpy as np
import pytest

import cpu_ref
from gartus_amd import abi, api, scenes
from gartus_amd.scene_data import DIFFUSE_LIGHT, LAMBERTIAN, QUAD, SPHERE, Material, SceneData

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def camera(width, lookfrom, lookat, vfov, depth, background, min_s, max_s, thr, nee=False):
    c = abi.default_camera()
    c.image_width = width
    c.aspect_ratio = 1.0
    c.max_depth = depth
    c.vertical_fov = vfov
    c.lookfrom[:] = lookfrom
    c.lookat[:] = lookat
    c.background[0:3] = background
    c.adaptive_min_samples = min_s
    c.adaptive_max_samples = max_s
    c.adaptive_error_threshold = thr
    c.samples_per_pixel = max_s
    if nee:
        c.direct_lighting_mode = abi.GT_LIGHTING_NEE
        c.light_pdf_weight = 0.0
    return c


def edge_scene():
    """An emitter of radiance 1 with slanted edges in front of a black sky."""
    d = SceneData()
    m = d.add_material(Material(DIFFUSE_LIGHT, (1.0, 1.0, 1.0)))
    d.add(QUAD, [-1.0, -0.7, 0.0, 1.7, 0.2, 0.0, 0.3, 1.5, 0.0], m)
    return d.as_float32_exact()


def gpu_render(ctx, data, cam, lights=None):
    tracer = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    tracer.camera.raw = cam
    img, st = tracer._render(api.RayScene(data, ctx), tracer.camera, lights, None, stats=True)
    return img.pixels.astype(np.float64), st


def test_zero_variance_pixels_stop_at_min_samples(gpu_context):
    """camera.rs:4975-4991 at image scale: a constant background gives every
    pixel exactly min samples and exactly the background."""
    say("adaptive: constant background")
    d = SceneData()
    m = d.add_material(Material(LAMBERTIAN, (0.5, 0.5, 0.5)))
    d.add(SPHERE, [0.0, 0.0, 50.0, 0.5], m)  # behind the camera: nothing is hit
    cam = camera(32, (0, 0, 0), (0, 0, -1), 60.0, 4, (0.25, 0.5, 0.75), 4, 32, 1e-3)
    img, st = gpu_render(gpu_context, d, cam)
    assert np.all(img == np.array([0.25, 0.5, 0.75]))
    assert st["camera_paths"] == 32 * 32 * 4
    assert st["accepted_samples"] == 32 * 32 * 4


def test_exact_samples_stop_where_the_oracle_stops(gpu_context):
    say("adaptive: emitter edges, depth 1")
    data = edge_scene()
    cam = camera(48, (0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 40.0, 1, (0.0, 0.0, 0.0), 4, 64, 0.08)
    gpu, st = gpu_render(gpu_context, data, cam)
    ref, ctr = cpu_ref.Scene(data).render(cam, None, None, threads=0)
    # the GPU writes its FP64 running mean as float32
    same = np.all(gpu == ref.astype(np.float32).astype(np.float64), axis=2)
    say(f"pixels equal {same.mean():.5f}, samples used gpu {st['accepted_samples']} oracle {ctr['accepted']}, "
        f"traced {st['camera_paths']}")
    assert same.mean() >= 0.995
    assert abs(st["accepted_samples"] - ctr["accepted"]) <= 0.005 * ctr["accepted"]
    # edge pixels kept sampling, the rest stopped at 4
    assert 48 * 48 * 4 < st["accepted_samples"] < 48 * 48 * 64
    assert 0.02 < ((gpu > 0) & (gpu < 1)).any(axis=2).mean() < 0.5


def test_rounds_and_batches_do_not_change_the_image(gpu_context):
    """A queue of 5,000 paths splits every round into several batches; the
    image must equal the default render bit for bit, and repeat exactly."""
    say("adaptive: small queue")
    data = edge_scene()
    cam = camera(48, (0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 40.0, 1, (0.0, 0.0, 0.0), 4, 64, 0.08)
    a, sa = gpu_render(gpu_context, data, cam)
    b, _ = gpu_render(gpu_context, data, cam)
    np.testing.assert_array_equal(a, b)
    old = gpu_context.tuning.max_paths
    try:
        gpu_context.set_tuning(max_paths=5000)
        c, sc = gpu_render(gpu_context, data, cam)
    finally:
        gpu_context.set_tuning(max_paths=old)
    np.testing.assert_array_equal(a, c)
    assert sa["camera_paths"] == sc["camera_paths"] and sa["accepted_samples"] == sc["accepted_samples"]


def test_lit_cornell_box_matches_oracle(gpu_context):
    say("adaptive: cornell 48x48 nee, min 8 max 64")
    data, lights, spec = scenes.cornell()
    data = data.as_float32_exact()
    cam = camera(48, spec["lookfrom"], spec["lookat"], spec["vfov"], 8, (0.0, 0.0, 0.0), 8, 64, 0.05, nee=True)
    L = api.SamplingTargetList()
    for kind, q, w in lights:
        L._add(kind, q, w)
    gpu, st = gpu_render(gpu_context, data, cam, L)
    ref, ctr = cpu_ref.Scene(data).render(cam, cpu_ref.Lights(lights), None, threads=0)
    g, r = gpu.reshape(-1, 3), ref.reshape(-1, 3)
    se = r.std(axis=0) / np.sqrt(len(r)) + 1e-12
    diff = np.abs(g.mean(0) - r.mean(0))
    say(f"mean gpu {g.mean(0)} oracle {r.mean(0)} diff/se {diff / se}; samples used {st['accepted_samples']} "
        f"{ctr['accepted']}, traced {st['camera_paths']}")
    assert np.all(diff <= 3.0 * se)
    assert abs(st["accepted_samples"] - ctr["accepted"]) <= 0.1 * ctr["accepted"]
    assert st["camera_paths"] < 48 * 48 * 64


def test_adaptive_settings_are_validated_and_batches_refuse_them(gpu_context):
    data = edge_scene()
    cam = camera(16, (0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 40.0, 1, (0.0, 0.0, 0.0), 4, 8, float("nan"))
    with pytest.raises(abi.InvalidArgument):
        gpu_render(gpu_context, data, cam)
    with pytest.raises(abi.InvalidArgument):
        api.RayCamera(16, 1.0).with_adaptive_sampling(4, 8, 0.0)
    c = api.RayCamera(16, 1.0).with_adaptive_sampling(0, 0, 0.5)
    assert c.adaptive_sampling() == (1, 1, 0.5) and c.samples_per_pixel() == 1  # camera.rs:4958-4965
    assert c.without_adaptive_sampling().adaptive_sampling() is None
    ok = camera(16, (0.0, 0.0, 4.0), (0.0, 0.0, 0.0), 40.0, 1, (0.0, 0.0, 0.0), 4, 8, 0.1)
    with pytest.raises(abi.GartusError):
        api.render_batch([(api.RayScene(data, gpu_context), _cam_obj(ok), None, None)])


def _cam_obj(raw):
    c = api.RayCamera.__new__(api.RayCamera)
    c.raw = raw
    return c
