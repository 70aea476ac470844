"""The scene compile on the device (gt_scene_build_device) against the host
SAH compile and the oracle's brute force.

Closest hits do not depend on the tree: for every ray the device-built scene
must hit the primitive the brute force finds (up to ties within 1e-5 in t),
and where it hits the same primitive as the host-built scene, t, u, v and
the normal must be bit-identical (the primitive records are packed by the
same formulas). Leaves of 1, 4 and 8 primitives, one- to five-primitive
scenes (the root itself a leaf), both node widths, a 526k-primitive terrain
and a rendered frame are covered.
"""
import sys

import numpy as np
import pytest

import cpu_ref
from gartus_amd import api, scenes
from gartus_amd.scene_data import LAMBERTIAN, QUAD, SPHERE, TRIANGLE, Material, SceneData

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def rays(rng, n, lo, hi, spread=1.0):
    o = rng.uniform(lo, hi, (n, 3))
    t = rng.uniform(lo, hi, (n, 3)) * spread
    d = t - o
    return o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)


def check_against(gpu, ref, host=None):
    same = gpu["prim"] == ref["prim"]
    both = (gpu["prim"] >= 0) & (ref["prim"] >= 0)
    rel = np.abs(gpu["t"] - ref["t"]) / np.maximum(np.abs(ref["t"]), 1e-30)
    tie = (~same) & both & (rel <= 1e-5)
    bad = ~(same | tie)
    assert bad.sum() == 0, f"{bad.sum()} closest-hit mismatches: {np.nonzero(bad)[0][:8]}"
    if host is not None:
        eq = (gpu["prim"] == host["prim"]) & (gpu["prim"] >= 0)
        for k in ("t", "u", "v"):
            np.testing.assert_array_equal(gpu[k][eq], host[k][eq])
        np.testing.assert_array_equal(gpu["normal"][eq], host["normal"][eq])
    return same.mean()


@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_randomized_scenes_hit_what_brute_force_hits(gpu_context, seed, leaf):
    say(f"device build: randomized scene seed {seed} leaf {leaf}")
    data = scenes.randomized_ray_scene(seed).as_float32_exact()
    dev = api.RayScene(data, gpu_context).with_device_bvh(leaf)
    host = api.RayScene(data, gpu_context)
    host.build_bvh_with_options(leaf)
    rng = np.random.default_rng(seed)
    o, d = rays(rng, 20000, (-3.5, -2.5, -8.0), (3.5, 2.5, 1.0))
    g = dev.intersect(o, d)
    h = host.intersect(o, d)
    ref = cpu_ref.Scene(data).intersect(o, d, use_bvh=False)
    assert (g["prim"] >= 0).mean() > 0.05
    check_against(g, ref, h)
    info = dev.info()
    assert info.built_on_device == 1 and info.build_ms > 0.0
    assert info.primitive_count == len(data)


@pytest.mark.parametrize("count", [1, 2, 3, 5])
def test_tiny_scenes(gpu_context, count):
    say(f"device build: {count} primitive(s)")
    d = SceneData()
    m = d.add_material(Material(LAMBERTIAN, (0.5, 0.5, 0.5)))
    rng = np.random.default_rng(count)
    for i in range(count):
        if i % 3 == 0:
            d.add(SPHERE, [*rng.uniform(-2, 2, 3), 0.6], m)
        elif i % 3 == 1:
            d.add(TRIANGLE, rng.uniform(-2, 2, 9), m)
        else:
            d.add(QUAD, [*rng.uniform(-2, 2, 3), 1.0, 0.2, 0.1, 0.1, 0.9, 0.3], m)
    data = d.as_float32_exact()
    o, dd = rays(np.random.default_rng(7), 5000, (-3, -3, -3), (3, 3, 3))
    ref = cpu_ref.Scene(data).intersect(o, dd, use_bvh=False)
    for leaf in (1, 4):
        g = api.RayScene(data, gpu_context).with_device_bvh(leaf).intersect(o, dd)
        check_against(g, ref)


def test_both_node_widths(gpu_context):
    say("device build: BVH2 and BVH4 on final_scene")
    data = scenes.final_scene().as_float32_exact()
    rng = np.random.default_rng(3)
    n = 20000
    o = np.stack([rng.uniform(-12, 12, n), rng.uniform(0.5, 3, n), rng.uniform(-12, 12, n)], 1)
    d = rng.normal(size=(n, 3))
    o, d = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
    ref = cpu_ref.Scene(data).intersect(o, d, use_bvh=False)
    old = gpu_context.tuning.bvh_width
    try:
        for width in (2, 4):
            gpu_context.set_tuning(bvh_width=width)
            g = api.RayScene(data, gpu_context).with_device_bvh().intersect(o, d)
            check_against(g, ref)
    finally:
        gpu_context.set_tuning(bvh_width=old)


def test_terrain_matches_the_host_tree(gpu_context):
    """526k primitives: the device tree answers like the host SAH tree."""
    say("device build: terrain 526k")
    data, _ = scenes.terrain()
    data = data.as_float32_exact()
    dev = api.RayScene(data, gpu_context).with_device_bvh()
    host = api.RayScene(data, gpu_context).with_bvh()
    rng = np.random.default_rng(11)
    n = 100000
    o = np.stack([rng.uniform(-40, 40, n), rng.uniform(5, 30, n), rng.uniform(-40, 40, n)], 1)
    d = np.stack([rng.uniform(-1, 1, n), rng.uniform(-1, -0.05, n), rng.uniform(-1, 1, n)], 1)
    o, d = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
    g = dev.intersect(o, d)
    h = host.intersect(o, d)
    frac = check_against(g, h, h)
    say(f"terrain: same primitive for {frac:.6f} of rays; device build {dev.info().build_ms:.2f} ms")
    assert frac > 0.999


def test_rendered_frame_matches_the_host_tree(gpu_context):
    """Cornell box with light connections: the same closest hits give the
    same image, pixel for pixel, except where a tie picks another face."""
    say("device build: cornell render")
    data, lights, cam = scenes.cornell()
    data = data.as_float32_exact()
    cam.update(width=64, grid=4)
    c = scenes.camera_record(cam)
    L = api.SamplingTargetList()
    for kind, q, w in lights:
        L._add(kind, q, w)
    tracer = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    tracer.camera.raw = c
    a = tracer.render_with_lights_hdr_image(api.RayScene(data, gpu_context).with_bvh(), L).pixels
    b = tracer.render_with_lights_hdr_image(api.RayScene(data, gpu_context).with_device_bvh(), L).pixels
    same = np.all(a == b, axis=2).mean()
    say(f"pixels identical: {same:.5f}")
    assert same > 0.99
