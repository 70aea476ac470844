"""Progressive rendering (render_world*_tiled_progressive, camera.rs:3172-3294)
through gt_render_progressive.

As in the reference's own test (camera.rs:5528-5555): the progressive canvas
equals the default render's display canvas, an 8x8 image in tiles of 3 gives
9 updates, and the last update is complete. Beyond it: tiles are reported in
row-major tile order with counters 1..N, every reported tile already holds
its final pixels, a small queue (several bands) changes nothing, a callback
error stops the render and reaches the caller (GT_ECANCELLED at the C ABI),
and adaptive cameras render progressively to the same canvas.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from gartus_amd import abi, api, scenes

pytestmark = pytest.mark.gpu


def say(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def tracer_for(raw):
    t = api.PathTracer(api.RayCamera.__new__(api.RayCamera))
    t.camera.raw = raw
    return t


def bench(width, spp=2, depth=3, seed=654):
    data, spec = scenes.bench_scene()
    spec = dict(spec, width=width, spp=spp, depth=depth, seed=seed, aspect=1.0)
    return data, scenes.camera_record(spec)


def test_progressive_matches_default_render(gpu_context):
    say("progressive: 8x8 tiles of 3")
    data, cam = bench(8)
    world = api.RayScene(data, gpu_context)
    t = tracer_for(cam)
    expected = t.render(world)
    updates = []

    def progress(u):
        assert u.image_width() == 8 and u.image_height() == 8
        assert u.pixels().shape == expected.shape
        updates.append(u.progress())

    got = t.render_progressive(world, progress, tile_size=3)
    np.testing.assert_array_equal(got, expected)
    assert len(updates) == 9 and updates[-1].is_complete()


def test_tiles_come_in_order_with_final_pixels(gpu_context):
    say("progressive: 96x96 tiles of 16, small queue")
    data, cam = bench(96, spp=4)
    world = api.RayScene(data, gpu_context)
    t = tracer_for(cam)
    final = t.render(world)
    seen = []

    def progress(u):
        p = u.progress()
        seen.append((p.tile.x, p.tile.y, p.tile.width, p.tile.height, p.completed_tiles, p.total_tiles))
        tile = u.pixels()[p.tile.y:p.tile.y_end(), p.tile.x:p.tile.x_end()]
        np.testing.assert_array_equal(tile, final[p.tile.y:p.tile.y_end(), p.tile.x:p.tile.x_end()])

    old = gpu_context.tuning.max_paths
    try:
        gpu_context.set_tuning(max_paths=3 * 96 * 16 * 4)  # a band of one tile row at a time
        got = t.render_progressive(world, progress, tile_size=16)
    finally:
        gpu_context.set_tuning(max_paths=old)
    np.testing.assert_array_equal(got, final)
    want = [(x, y, 16, 16) for y in range(0, 96, 16) for x in range(0, 96, 16)]
    assert [s[:4] for s in seen] == want
    assert [s[4] for s in seen] == list(range(1, 37)) and all(s[5] == 36 for s in seen)


def test_callback_error_stops_the_render(gpu_context):
    say("progressive: cancel")
    data, cam = bench(64, spp=2)
    world = api.RayScene(data, gpu_context)
    t = tracer_for(cam)
    calls = []

    class Stop(Exception):
        pass

    def progress(u):
        calls.append(u.progress().completed_tiles)
        if len(calls) == 3:
            raise Stop()

    with pytest.raises(Stop):
        t.render_progressive(world, progress, tile_size=16)
    assert calls == [1, 2, 3]
    # the C ABI reports the stop as GT_ECANCELLED
    ctx = gpu_context
    fn = abi.PROGRESS_FN(lambda user, up: 7)
    canvas = np.zeros((64, 64, 3), np.uint8)
    rc = ctx.lib.gt_render_progressive(ctx.handle, world._compile(), C.byref(cam), None, None, 16, fn, None,
                                       canvas.ctypes.data_as(C.POINTER(C.c_uint8)))
    assert rc == abi.GT_ECANCELLED
    assert b"stopped" in ctx.lib.gt_last_error()


def test_adaptive_camera_renders_progressively(gpu_context):
    say("progressive: adaptive cornell")
    data, lights, spec = scenes.cornell()
    cam = scenes.camera_record(dict(spec, width=40, grid=None, spp=16))
    cam.pixel_sample_mode = abi.GT_PIXEL_RANDOM
    cam.adaptive_min_samples, cam.adaptive_max_samples, cam.adaptive_error_threshold = 4, 32, 0.05
    world = api.RayScene(data, gpu_context)
    L = api.SamplingTargetList()
    for kind, q, w in lights:
        L._add(kind, q, w)
    t = tracer_for(cam)
    final = t.render_with_lights(world, L)
    got = t.render_with_lights_progressive(world, L, None, tile_size=8)
    np.testing.assert_array_equal(got, final)
