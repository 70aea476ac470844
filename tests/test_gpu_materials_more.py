"""The GPU's Isotropic and Henyey-Greenstein materials and mesh-triangle
shading normals, through the C ABI, against closed forms and the oracle.

  - furnace: convex spheres of either phase material with albedo 1 under a
    constant dyadic environment return it exactly per sample, continuation
    only (scattering_pdf / pdf = 1, every path leaves the sphere);
  - a room with Isotropic and Henyey-Greenstein spheres, a smooth
    tessellated sphere (vertex normals) of each surface material and a quad
    light, with NEE and with the mixture continuation: same random streams
    as the oracle, the mean per-pixel difference within 3 sigma of the
    differences' spread (plus 1e-5 relative) and most pixels within 1e-3;
  - the normal AOV of the smooth sphere is its interpolated shading normal,
    equal to the oracle's.
"""
import sys

import numpy as np
import pytest

import cpu_ref
from gartus_amd import abi, api, scenes
from gartus_amd.scene_data import (DIELECTRIC, DIFFUSE_LIGHT, GGX, HENYEY_GREENSTEIN, ISOTROPIC, LAMBERTIAN,
                                   MESH_TRIANGLE, METAL, QUAD, SPHERE, Material, SceneData)

pytestmark = pytest.mark.gpu

THREADS = max(1, min(16, cpu_ref.cpu_count()))


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def tracer_for(cam_raw):
    t = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    t.camera.raw = cam_raw
    return t


@pytest.mark.parametrize("material", [Material(ISOTROPIC, (1.0, 1.0, 1.0)),
                                      Material(HENYEY_GREENSTEIN, (1.0, 1.0, 1.0), 0.7),
                                      Material(HENYEY_GREENSTEIN, (1.0, 1.0, 1.0), -0.4)])
def test_phase_material_furnace_is_exact(gpu_context, material):
    say(f"phase furnace kind {material.kind} g {material.param}")
    s = SceneData()
    m = s.add_material(material)
    s.add(SPHERE, [0.0, 0.0, -3.0, 1.0], m)
    c = scenes.camera_record(dict(width=24, aspect=1.0, spp=32, depth=64, vfov=30.0, lookfrom=(0, 0, 0),
                                  lookat=(0, 0, -1), rr=-1))
    t = tracer_for(c)  # continuation only: the camera's own strategy, not the forced NEE of render_with_environment
    img = t._render(api.RayScene(s), t.camera, None, api.EnvironmentLight.constant((0.75, 1.25, 2.5))).pixels
    ys, xs = np.mgrid[0:24, 0:24]
    inner = img[np.hypot(xs - 11.5, ys - 11.5) < 5.0]
    np.testing.assert_array_equal(inner, np.broadcast_to(np.float32([0.75, 1.25, 2.5]), inner.shape))


def _room():
    s = SceneData()
    white = s.add_material(Material(LAMBERTIAN, (0.73, 0.73, 0.73)))
    green = s.add_material(Material(LAMBERTIAN, (0.12, 0.45, 0.15)))
    s.add(QUAD, [[0, 0, 0, 10, 0, 0, 0, 0, 10], [0, 10, 10, 10, 0, 0, 0, 0, -10], [0, 0, 10, 10, 0, 0, 0, 10, 0],
                 [10, 0, 10, 0, 0, -10, 0, 10, 0]], white)
    s.add(QUAD, [[0, 0, 0, 0, 0, 10, 0, 10, 0]], green)
    iso = s.add_material(Material(ISOTROPIC, (0.8, 0.7, 0.6)))
    hg = s.add_material(Material(HENYEY_GREENSTEIN, (0.6, 0.8, 0.9), 0.6))
    s.add(SPHERE, [[2.5, 1.5, 6.0, 1.2]], iso)
    s.add(SPHERE, [[7.5, 1.5, 6.5, 1.2]], hg)
    for center, mat in (((2.5, 4.5, 3.5), Material(LAMBERTIAN, (0.9, 0.5, 0.3))),
                        ((5.0, 2.0, 4.0), Material(GGX, (0.95, 0.64, 0.54), 0.3)),
                        ((7.5, 4.5, 3.0), Material(METAL, (0.9, 0.9, 0.9), 0.1))):
        mid = s.add_material(mat)
        pos, nrm = scenes.smooth_sphere_mesh(center, 1.1, 10, 20)
        s.add(MESH_TRIANGLE, np.concatenate([pos, nrm], axis=1), mid)
    lm = s.add_material(Material(DIFFUSE_LIGHT, (15.0, 15.0, 15.0)))
    light = [3.5, 9.95, 3.5, 3.0, 0.0, 0.0, 0.0, 0.0, 3.0]
    s.add(QUAD, [light], lm)
    return s, [(QUAD, np.array(light), 1.0)]


@pytest.mark.parametrize("mode", ["nee", "mixture"])
def test_phase_materials_and_smooth_meshes_match_oracle(gpu_context, mode):
    say(f"room with phase materials and smooth meshes, {mode}, 96x96, 32 spp, depth 6")
    data, lights = _room()
    data = data.as_float32_exact()
    lights = [(k, q.astype(np.float32).astype(np.float64), w) for k, q, w in lights]
    c = scenes.camera_record(dict(width=96, aspect=1.0, spp=32, depth=6, vfov=50.0, lookfrom=(5.0, 5.0, -8.0),
                                  lookat=(5.0, 5.0, 5.0), background=(0.0, 0.0, 0.0), nee=(mode == "nee"), seed=7))
    if mode == "mixture":
        c.light_pdf_weight = 0.5
    L = api.SamplingTargetList()
    for k, q, w in lights:
        L._add(k, q, w)
    gpu = tracer_for(c).render_with_lights_hdr_image(api.RayScene(data).with_bvh(), L).pixels
    ref, _ = cpu_ref.Scene(data).render(c, cpu_ref.Lights(lights), None, threads=THREADS)
    g = gpu.astype(np.float64).reshape(-1, 3)
    r = ref.reshape(-1, 3)
    d = g - r
    se = d.std(0) / np.sqrt(len(d))
    close = np.all(np.abs(d) <= 1e-3 * (1.0 + np.abs(r)), axis=1).mean()
    say(f"room {mode}: gpu mean {g.mean(0)} oracle mean {r.mean(0)} |diff| {np.abs(d.mean(0))} 3se {3 * se}; "
        f"within 1e-3 {close:.4f}")
    assert np.all(np.abs(d.mean(0)) <= 3 * se + 1e-5 * np.abs(r).mean(0))
    assert close > 0.97
    assert r.mean() > 0.05


def test_smooth_mesh_normal_aov_matches_oracle(gpu_context):
    say("normal AOV of smooth meshes, 64x64, 4 spp")
    data, _ = _room()
    data = data.as_float32_exact()
    c = scenes.camera_record(dict(width=64, aspect=1.0, spp=4, depth=4, vfov=50.0, lookfrom=(5.0, 5.0, -8.0),
                                  lookat=(5.0, 5.0, 5.0), seed=3))
    aov = tracer_for(c).render_denoising_aovs(api.RayScene(data).with_bvh())
    b, a, n = cpu_ref.Scene(data).render_aovs(c, None, None, threads=THREADS)
    np.testing.assert_allclose(aov.albedo_linear, a, atol=1e-5)
    close = np.all(np.abs(aov.normal_linear - n) <= 1e-4, axis=-1)
    say(f"normal AOV within 1e-4: {close.mean():.4f}")
    assert close.mean() > 0.99
