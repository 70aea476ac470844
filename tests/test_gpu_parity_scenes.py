"""GPU images against the CPU oracle on the scene families of the measured
workloads, at sizes the oracle renders in seconds, and on the camera features
the workloads leave off.

Both sides draw the same per-(pixel, sample, bounce) random numbers
(GT_RNG_COUNTER) over FP32-exact geometry, so a pixel differs only where FP32
rounding flips a branch. As in test_gpu_parity: the image mean per channel
must agree within 3 standard errors and most pixels within 1e-3 relative (the
share is stated per test; glass and metal make more pixels sensitive to
rounding than the Lambertian Cornell box).
"""
import sys

import numpy as np
import pytest

import cpu_ref
from gartus_amd import abi, api, scenes
from gartus_amd.scene_data import (DIELECTRIC, DIFFUSE_LIGHT, GGX, LAMBERTIAN, METAL, MOVING_SPHERE, QUAD, SPHERE,
                                   TRIANGLE, Material, SceneData)

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def compare(gpu_img, ref_img, label):
    g = gpu_img.reshape(-1, 3).astype(np.float64)
    r = ref_img.reshape(-1, 3)
    se = r.std(axis=0) / np.sqrt(r.shape[0]) + 1e-12
    diff = np.abs(g.mean(0) - r.mean(0))
    close = np.all(np.abs(g - r) <= 1e-3 * (1.0 + np.abs(r)), axis=1).mean()
    say(f"{label}: gpu mean {g.mean(0)} oracle mean {r.mean(0)} diff/se {diff / se}; pixels within 1e-3 {close:.4f}")
    assert np.all(diff <= 3.0 * se), (diff, se)
    return close


def render(data, cam, lights=None):
    tracer = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    tracer.camera.raw = cam
    world = api.RayScene(data).with_bvh()
    if lights is None:
        return tracer.render_ray_scene_hdr_image(world).pixels
    L = api.SamplingTargetList()
    for kind, q, w in lights:
        L._add(kind, q, w)
    return tracer.render_with_lights_hdr_image(world, L).pixels


def test_terrain_matches_oracle(gpu_context):
    """The 720p terrain's family: a Perlin height map of triangles in 8
    Lambertian bands with Lambertian, metal and glass spheres, random
    sampling, depth 6, constant sky."""
    say("terrain 33x33 grid (2,048 triangles) + 80 spheres, 64x36, 8 spp, depth 6")
    data, spec = scenes.terrain(grid=32, n_spheres=80)
    data = data.as_float32_exact()
    cam = scenes.camera_record(dict(spec, width=64, spp=8))
    gpu = render(data, cam)
    ref, _ = cpu_ref.Scene(data).render(cam, None, None, threads=0)
    assert compare(gpu, ref, "terrain") > 0.99


@pytest.mark.parametrize("seed", [1000, 1010])
def test_batch_scene_matches_oracle(gpu_context, seed):
    """Scenes of the 64-scene batch (torus / crystal field / height map of
    triangles, Lambertian, GGX and glass, 1-8 quad lights, NEE) at 48 pixels
    wide instead of their own resolution."""
    data, lights, spec = scenes.batch_scene(seed)
    say(f"batch scene {seed}: {len(data)} primitives, {len(lights)} lights, 48 px, 8 spp, depth 8, nee")
    data = data.as_float32_exact()
    cam = scenes.camera_record(dict(spec, width=48))
    gpu = render(data, cam, lights)
    ref, _ = cpu_ref.Scene(data).render(cam, cpu_ref.Lights(lights), None, threads=0)
    assert compare(gpu, ref, f"batch scene {seed}") > 0.99


def features_scene():
    """Moving spheres (shutter interval), a quad light, triangles, metal,
    glass and GGX under a vertical-gradient sky."""
    d = SceneData()
    ground = d.add_material(Material(LAMBERTIAN, (0.5, 0.55, 0.5)))
    red = d.add_material(Material(LAMBERTIAN, (0.7, 0.2, 0.15)))
    metal = d.add_material(Material(METAL, (0.8, 0.7, 0.6), 0.15))
    glass = d.add_material(Material(DIELECTRIC, (1.0, 1.0, 1.0), 1.5))
    ggx = d.add_material(Material(GGX, (0.9, 0.6, 0.3), 0.3))
    lamp = d.add_material(Material(DIFFUSE_LIGHT, (6.0, 6.0, 5.0)))
    d.add(SPHERE, [0.0, -1000.0, 0.0, 1000.0], ground)
    d.add(MOVING_SPHERE, [-1.5, 0.6, 0.0, -1.5, 1.1, 0.0, 0.6], red)
    d.add(MOVING_SPHERE, [0.0, 0.7, -0.5, 0.4, 0.7, -0.5, 0.7], glass)
    d.add(SPHERE, [1.6, 0.6, 0.2, 0.6], metal)
    d.add(SPHERE, [0.8, 0.35, 1.3, 0.35], ggx)
    d.add(TRIANGLE, [-2.5, 0.0, -2.0, -0.5, 0.0, -2.5, -1.5, 1.8, -2.2], red)
    light_q = np.array([-1.0, 3.0, -1.0, 2.0, 0.0, 0.0, 0.0, 0.0, 2.0])
    d.add(QUAD, light_q, lamp)
    lights = [(QUAD, light_q, 1.0)]
    return d.as_float32_exact(), lights


def features_camera(**kw):
    c = abi.default_camera()
    c.image_width = 48
    c.aspect_ratio = 1.5
    c.samples_per_pixel = 16
    c.max_depth = 6
    c.vertical_fov = 40.0
    c.lookfrom[:] = (0.5, 1.6, 6.0)
    c.lookat[:] = (0.0, 0.6, 0.0)
    c.defocus_angle = 1.5
    c.focus_distance = 6.0
    c.shutter_start, c.shutter_end = 0.0, 1.0
    c.background_kind = abi.GT_BACKGROUND_VERTICAL_GRADIENT
    c.background[0:9] = (0.2, 0.15, 0.1, 0.6, 0.7, 0.8, 0.3, 0.45, 0.9)
    c.direct_lighting_mode = abi.GT_LIGHTING_CONTINUATION
    c.light_pdf_weight = 0.3
    c.russian_roulette_min_depth = -1
    c.rng_seed = 17
    for k, v in kw.items():
        setattr(c, k, v)
    return c


@pytest.mark.parametrize("variant", ["mixture", "nee_rr2", "stratified"])
def test_camera_features_match_oracle(gpu_context, variant):
    """Defocus disk, shutter time with moving spheres, gradient background,
    the mixture continuation with light weight 0.3 and Russian roulette off;
    then NEE with roulette from bounce 2; then the stratified jitter."""
    data, lights = features_scene()
    if variant == "mixture":
        cam = features_camera()
    elif variant == "nee_rr2":
        cam = features_camera(direct_lighting_mode=abi.GT_LIGHTING_NEE, light_pdf_weight=0.0,
                              russian_roulette_min_depth=2)
    else:
        cam = features_camera(pixel_sample_mode=abi.GT_PIXEL_STRATIFIED_GRID, stratified_grid_width=4)
    say(f"camera features ({variant}): 48x32, 16 spp, depth 6")
    gpu = render(data, cam, lights)
    ref, _ = cpu_ref.Scene(data).render(cam, cpu_ref.Lights(lights), None, threads=0)
    assert compare(gpu, ref, f"features {variant}") > 0.99
