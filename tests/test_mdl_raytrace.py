"""The MDL runtime's raytrace call (src/mdl/runtime.rs:797-837) on this
library: its conversions on the CPU, and its render against the oracle."""
import math
import sys

import numpy as np
import pytest

from gartus_amd import abi, mdl_raytrace as mr
from gartus_amd.scene_data import LAMBERTIAN, METAL, QUAD, SPHERE, Material, SceneData


def test_conversions():
    assert mr.rgb_from_vec3((12.5, -3.0, 300.0)) == (13, 0, 255)  # f64::round: halves away from zero
    assert mr.rgb_from_vec3((0.49, 254.5, 2.5)) == (0, 255, 3)
    np.testing.assert_allclose(mr.linear_from_rgb_srgb((255, 128, 0)), (1.0, (128 / 255) ** 2, 0.0))
    np.testing.assert_allclose(mr.linear_from_rgb_linear_units((255, 51, 0)), (1.0, 0.2, 0.0))
    assert mr.vertical_fov_from_focal(100, 50.0) == pytest.approx(90.0)
    assert mr.vertical_fov_from_focal(100, 0.0) is None
    assert mr.vertical_fov_from_focal(100, math.inf) is None
    assert mr.lookat_or_default_forward((1.0, 2.0, 3.0), (1.0, 2.0, 3.0)) == (1.0, 2.0, 2.0)
    assert mr.lookat_or_default_forward((0.0, 0.0, 0.0), (0.0, 0.0, 5.0)) == (0.0, 0.0, 5.0)


def _scene():
    s = SceneData()
    grey = s.add_material(Material(LAMBERTIAN, (0.6, 0.6, 0.6)))
    red = s.add_material(Material(LAMBERTIAN, (0.7, 0.2, 0.15)))
    steel = s.add_material(Material(METAL, (0.8, 0.8, 0.85), 0.1))
    s.add(QUAD, [-200.0, 0.0, -200.0, 400.0, 0.0, 0.0, 0.0, 0.0, 400.0], grey)
    s.add(SPHERE, [[-60.0, 50.0, 0.0, 50.0]], red)
    s.add(SPHERE, [[70.0, 40.0, 20.0, 40.0]], steel)
    return s


LIGHTS = [mr.MdlLight((-100.0, 300.0, 150.0), (255.0, 240.0, 220.0)), mr.MdlLight((150.0, 200.0, -50.0),
                                                                                 (40.0, 60.5, 200.0))]
CAMERA = mr.MdlCamera(eye=(0.0, 150.0, 500.0), aim=(0.0, 40.0, 0.0), focal=110.0)


def test_raytrace_scene_builds_the_reference_camera_and_lights():
    tracer, world, targets = mr.raytrace_scene(_scene(), 96, 64, LIGHTS, CAMERA, background_rgb=(30, 60, 90),
                                               samples_per_pixel=12, max_depth=6, light_radius=10.0)
    raw = tracer.camera.raw
    assert (raw.image_width, raw.samples_per_pixel, raw.max_depth) == (96, 12, 6)
    assert raw.aspect_ratio == 96 / 64
    assert raw.vertical_fov == pytest.approx(math.degrees(2 * math.atan(32 / 110.0)))
    np.testing.assert_allclose(raw.background[0:3], [(30 / 255) ** 2, (60 / 255) ** 2, (90 / 255) ** 2])
    np.testing.assert_allclose([raw.lookfrom[i] for i in range(3)], CAMERA.eye)
    np.testing.assert_allclose([raw.lookat[i] for i in range(3)], CAMERA.aim)
    # one DiffuseLight sphere per light (12x the byte colour / 255) and the same spheres as targets
    assert len(targets) == 2 and len(world) == 3 + 2
    emits = [m.color for m in world.data.materials[3:]]
    np.testing.assert_allclose(emits[0], [12.0, 12.0 * 240 / 255, 12.0 * 220 / 255])
    np.testing.assert_allclose(emits[1], [12.0 * 40 / 255, 12.0 * 61 / 255, 12.0 * 200 / 255])  # 60.5 -> 61
    kinds = [k for k, _, _ in world.data.primitives()]
    assert kinds[-2:] == [SPHERE, SPHERE]


@pytest.mark.gpu
def test_raytrace_matches_oracle(gpu_context):
    import cpu_ref
    print("MDL raytrace pattern: 96x64, 32 spp, depth 8, two sphere lights", flush=True)
    tracer, world, targets = mr.raytrace_scene(_scene(), 96, 64, LIGHTS, CAMERA, background_rgb=(30, 60, 90),
                                               samples_per_pixel=32, max_depth=8, context=gpu_context)
    world.build_bvh()
    hdr = tracer.render_with_lights_hdr_image(world, targets).pixels.astype(np.float64)
    lights = [(SPHERE, np.array([*l.position, 10.0]), 1.0) for l in LIGHTS]
    ref, _ = cpu_ref.Scene(world.data).render(tracer.camera.raw, cpu_ref.Lights(lights), None,
                                              threads=max(1, min(16, cpu_ref.cpu_count())))
    d = (hdr - ref).reshape(-1, 3)
    se = d.std(0) / np.sqrt(len(d))
    close = np.all(np.abs(d) <= 1e-3 * (1.0 + np.abs(ref.reshape(-1, 3))), axis=1).mean()
    print(f"  gpu mean {hdr.reshape(-1, 3).mean(0)} oracle {ref.reshape(-1, 3).mean(0)} 3se {3 * se}; "
          f"within 1e-3 {close:.4f}", flush=True)
    assert np.all(np.abs(d.mean(0)) <= 3 * se + 1e-5 * np.abs(ref).reshape(-1, 3).mean(0))
    assert close > 0.98
    # the canvas is the tone-mapped image of that render
    canvas = mr.raytrace_canvas(_scene(), 96, 64, LIGHTS, CAMERA, background_rgb=(30, 60, 90), samples_per_pixel=32,
                                max_depth=8, context=gpu_context)
    assert canvas.dtype == np.uint8 and canvas.shape == (64, 96, 3)
    np.testing.assert_array_equal(canvas, tracer.render_with_lights(world, targets))
