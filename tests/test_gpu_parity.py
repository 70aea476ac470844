"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Scenes are made FP32-exact first (SceneData.as_float32_exact) so both sides
see the same geometry. Tolerances:
  - closest hit: primitive ids equal the oracle's brute force except where
    the two best candidates lie within 1e-5 relative in t (a tie); t within
    1e-5 relative;
  - images: both sides draw the same per-(pixel, sample, bounce) random
    numbers (GT_RNG_COUNTER), so a pixel differs only where FP32 rounding
    flips a branch; the image mean per channel must agree within 3 sigma of
    the Monte Carlo standard error and most pixels within 1e-3 relative.
"""
import sys

import numpy as np
import pytest

import cpu_ref
from gartus_amd import abi, api, scenes
from gartus_amd.scene_data import SceneData

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def random_rays(rng, n, lo=(-3.5, -2.5, 0.0), hi=(3.5, 2.5, 1.0), target_z=(-8.5, -0.5)):
    o = rng.uniform(lo, hi, (n, 3))
    tgt = np.stack([rng.uniform(-3.5, 3.5, n), rng.uniform(-2.5, 2.5, n), rng.uniform(*target_z, n)], 1)
    d = tgt - o
    return o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)


def classify(gpu, ref_scene, o, d):
    """Mismatched ids whose t agree within 1e-5 relative are ties."""
    ref = ref_scene.intersect(o, d, use_bvh=False)
    same = gpu["prim"] == ref["prim"]
    both = (gpu["prim"] >= 0) & (ref["prim"] >= 0)
    rel = np.abs(gpu["t"] - ref["t"]) / np.maximum(np.abs(ref["t"]), 1e-30)
    tie = (~same) & both & (rel <= 1e-5)
    return ref, same, tie, rel


@pytest.mark.parametrize("seed", range(16))
def test_intersect_randomized_scene_matches_bruteforce(gpu_context, seed):
    """raytracing.rs:1340-1370 at scale: BVH closest hit vs brute force."""
    say(f"intersect randomized_ray_scene seed={seed}")
    data = scenes.randomized_ray_scene(seed).as_float32_exact()
    rs = api.RayScene(data).with_bvh()
    ref_scene = cpu_ref.Scene(data)
    rng = np.random.default_rng(seed)
    o, d = random_rays(rng, 20000)
    gpu = rs.intersect(o, d)
    ref, same, tie, rel = classify(gpu, ref_scene, o, d)
    other = ~(same | tie)
    assert other.sum() == 0, f"{other.sum()} mismatches: {np.nonzero(other)[0][:10]}"
    hit = gpu["prim"] >= 0
    assert hit.mean() > 0.05
    assert np.max(rel[hit & same]) <= 1e-5
    np.testing.assert_array_equal(gpu["front"][hit], ref["front"][hit])
    # sphere normals are (p - c) / r: the FP32 error of t (1e-6 relative at
    # |t d| ~ 10) divided by radii down to 0.1 allows ~1e-4
    np.testing.assert_allclose(gpu["normal"][hit], ref["normal"][hit], atol=3e-4)


def test_intersect_traversal_stats_match_reference_bvh(gpu_context):
    """The GPU walks the reference's tree in the reference's order, so the
    BvhTraversalStats counters (bvh.rs:9-52) agree with the oracle's."""
    say("intersect stats final_scene")
    data = scenes.final_scene().as_float32_exact()
    rs = api.RayScene(data).with_bvh()
    ref_scene = cpu_ref.Scene(data)
    rng = np.random.default_rng(5)
    n = 20000
    o = np.stack([rng.uniform(-12, 12, n), rng.uniform(0.5, 3, n), rng.uniform(-12, 12, n)], 1)
    d = rng.normal(size=(n, 3))
    o, d = o.astype(np.float32).astype(np.float64), d.astype(np.float32).astype(np.float64)
    gpu = rs.intersect(o, d, with_stats=True)
    ref = ref_scene.intersect(o, d)
    assert rs.bvh_node_count() == ref_scene.node_count()
    gs, rs_ = gpu["stats"], ref["stats"]
    assert gs["rays"] == rs_["rays"] == n
    for k in ("node_bounds_tests", "node_bounds_hits", "leaf_visits", "primitive_candidates", "leaf_hit_results"):
        assert abs(gs[k] - rs_[k]) <= 1e-3 * rs_[k] + 5, (k, gs[k], rs_[k])
    assert gs["max_stack_depth"] == rs_["max_stack_depth"]
    assert (gpu["prim"] == ref["prim"]).mean() > 0.999


def _compare_images(gpu_img, ref_img, n_samples, label):
    g = gpu_img.reshape(-1, 3).astype(np.float64)
    r = ref_img.reshape(-1, 3)
    # standard error of the image mean from the per-pixel spread
    se = r.std(axis=0) / np.sqrt(r.shape[0]) + 1e-12
    diff = np.abs(g.mean(0) - r.mean(0))
    say(f"{label}: gpu mean {g.mean(0)} ref mean {r.mean(0)} diff/se {diff / se}")
    assert np.all(diff <= 3.0 * se), (diff, se)
    close = np.all(np.abs(g - r) <= 1e-3 * (1.0 + np.abs(r)), axis=1)
    say(f"{label}: pixels within 1e-3: {close.mean():.4f}")
    return close.mean()


def test_render_cornell_counter_mode_matches_oracle(gpu_context):
    say("render cornell 64x64 grid 4 depth 8 nee")
    data, lights, cam = scenes.cornell()
    data = data.as_float32_exact()
    cam.update(width=64, grid=4)
    c = scenes.camera_record(cam)
    rs = api.RayScene(data).with_bvh()
    L = api.SamplingTargetList()
    for kind, q, w in lights:
        L._add(kind, q, w)
    tracer = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    tracer.camera.raw = c
    gpu = tracer.render_with_lights_hdr_image(rs, L).pixels
    ref, ctr = cpu_ref.Scene(data).render(c, cpu_ref.Lights(lights), None, threads=0)
    frac = _compare_images(gpu, ref, 16, "cornell")
    assert frac > 0.97


def test_render_bench_scene_continuation_mode_matches_oracle(gpu_context):
    """benches/render_performance.rs scene and camera (no lights, sky)."""
    say("render bench_scene 96x54 spp 2 depth 5")
    data, cam = scenes.bench_scene()
    data = data.as_float32_exact()
    cam["spp"] = 16
    c = scenes.camera_record(cam)
    rs = api.RayScene(data).with_bvh()
    tracer = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    tracer.camera.raw = c
    gpu = tracer.render_ray_scene_hdr_image(rs).pixels
    ref, _ = cpu_ref.Scene(data).render(c, None, None, threads=0)
    assert _compare_images(gpu, ref, 16, "bench") > 0.97


def test_render_ggx_environment_matches_oracle(gpu_context):
    say("render ggx env 96x54 spp 8 depth 10")
    data, env, cam = scenes.ggx_environment(n_random=150, env_size=(256, 128))
    data = data.as_float32_exact()
    cam.update(width=96, spp=8)
    c = scenes.camera_record(cam)
    rs = api.RayScene(data).with_bvh()
    tracer = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    tracer.camera.raw = c
    gpu = tracer.render_with_environment_hdr_image(rs, api.EnvironmentLight.from_rgb32f(env)).pixels
    ref, _ = cpu_ref.Scene(data).render(c, None, cpu_ref.Env(rgb32f=env), threads=0)
    # the sun (5e4) makes a few pixels very sensitive to FP32 rounding
    assert _compare_images(np.minimum(gpu, 50.0), np.minimum(ref, 50.0), 8, "ggx-env") > 0.99


def test_render_ggx_environment_with_rejection_flag_matches_oracle(gpu_context):
    """GT_FLAG_GGX_REJECT_BELOW_HORIZON on both sides: same streams, same
    per-pixel agreement; without the compounding fallback no pixel needs
    clipping."""
    say("render ggx env 96x54 spp 8 depth 10, reject below-horizon GGX samples")
    data, env, cam = scenes.ggx_environment(n_random=150, env_size=(256, 128))
    data = data.as_float32_exact()
    cam.update(width=96, spp=8)
    c = scenes.camera_record(cam)
    c.flags = abi.GT_FLAG_GGX_REJECT_BELOW_HORIZON
    rs = api.RayScene(data).with_bvh()
    tracer = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    tracer.camera.raw = c
    gpu = tracer.render_with_environment_hdr_image(rs, api.EnvironmentLight.from_rgb32f(env)).pixels
    ref, _ = cpu_ref.Scene(data).render(c, None, cpu_ref.Env(rgb32f=env), threads=0)
    assert _compare_images(gpu, ref, 8, "ggx-env-reject") > 0.995


def test_c3_fallback_firefly_and_rejection_on_gpu(gpu_context):
    """The 1080p GGX/environment frame at 1 spp, seed 5: pixel (760, 717) is
    the compounding-fallback sample the oracle reproduces (2.9e8, see
    test_oracle_render); the flag makes it an ordinary sample."""
    say("render ggx env 1920x1080 spp 1 seed 5, with and without rejection")
    data, env, cam = scenes.ggx_environment()
    rs = api.RayScene(data).with_bvh()
    envl = api.EnvironmentLight.from_rgb32f(env)
    out = {}
    for flags in (0, abi.GT_FLAG_GGX_REJECT_BELOW_HORIZON):
        c = scenes.camera_record(cam)
        c.rng_seed, c.samples_per_pixel, c.flags = 5, 1, flags
        tracer = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
        tracer.camera.raw = c
        out[flags] = tracer.render_with_environment_hdr_image(rs, envl).pixels
    assert out[0][717, 760, 0] > 1e8
    fixed = out[abi.GT_FLAG_GGX_REJECT_BELOW_HORIZON]
    assert fixed[717, 760].max() < 10.0
    assert np.percentile(fixed.max(axis=2), 99.99) < np.percentile(out[0].max(axis=2), 99.99)


def test_render_aovs_match_oracle(gpu_context):
    say("render aovs final_scene 64x36 spp 4")
    data = scenes.final_scene().as_float32_exact()
    c = scenes.camera_record(dict(width=64, aspect=16 / 9, spp=4, depth=4, vfov=20.0, lookfrom=(13.0, 2.0, 3.0),
                                  lookat=(0.0, 0.0, 0.0), seed=9))
    rs = api.RayScene(data).with_bvh()
    tracer = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    tracer.camera.raw = c
    aov = tracer.render_denoising_aovs(rs)
    b, a, n = cpu_ref.Scene(data).render_aovs(c, None, None, threads=0)
    np.testing.assert_allclose(aov.albedo_linear, a, atol=1e-5)
    close = np.all(np.abs(aov.normal_linear - n) <= 1e-4, axis=-1)
    assert close.mean() > 0.995
    _compare_images(aov.beauty_linear, b, 4, "aov-beauty")
