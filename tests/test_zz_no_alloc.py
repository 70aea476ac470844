"""gt_intersect_device_noalloc, the closest-hit entry point that allocates
nothing, against gt_intersect (the entry point it is a variant of): hit
records and traversal statistics must be identical, bit for bit, on two
calls in a row, the second with a smaller input, on a scene of spheres and
on a triangle terrain whose tree spills the LDS stack into the scratch.
"""
import ctypes as C
import sys

import numpy as np
import pytest

from gartus_amd import abi, api, scenes

pytestmark = pytest.mark.gpu


def say(msg: str) -> None:
    print(msg, flush=True)
    sys.stdout.flush()


def _rays(rng, n, lo, hi):
    o = rng.uniform(lo, hi, (n, 3))
    tgt = rng.uniform(lo, hi, (n, 3))
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = o
    rays[:, 3] = rng.uniform(0, 1, n)
    rays[:, 4:7] = tgt - o
    rays[:, 7] = np.inf
    return rays


def _host_intersect(ctx, scene_handle, rays):
    n = len(rays)
    hits = np.zeros((n, 8), np.float32)
    st = abi.GtTraversalStats()
    abi.check(ctx.lib.gt_intersect(ctx.handle, scene_handle, rays.ctypes.data_as(C.POINTER(abi.GtRay)), n,
                                   C.c_float(0.001), hits.ctypes.data_as(C.POINTER(abi.GtHit)), C.byref(st)))
    return hits, np.array([getattr(st, k) for k, _ in abi.GtTraversalStats._fields_], np.uint64)


class DeviceBuffers:
    """Caller-owned device memory for the rays, hits and stats."""

    def __init__(self, ctx, n):
        self.ctx = ctx
        self.ptrs = []
        self.rays = self._alloc(32 * n)
        self.hits = self._alloc(32 * n)
        self.stats = self._alloc(8 * 7)

    def _alloc(self, nbytes):
        p = C.c_void_p()
        abi.check(self.ctx.lib.gt_device_malloc(self.ctx.handle, nbytes, C.byref(p)))
        self.ptrs.append(p)
        return p

    def free(self):
        for p in self.ptrs:
            self.ctx.lib.gt_device_free(self.ctx.handle, p)
        self.ptrs = []


def _noalloc_intersect(ctx, scene_handle, rays, buf):
    n = len(rays)
    lib = ctx.lib
    abi.check(lib.gt_memcpy_to_device(ctx.handle, buf.rays, rays.ctypes.data_as(C.c_void_p), rays.nbytes))
    abi.check(lib.gt_intersect_device_noalloc(ctx.handle, scene_handle, buf.rays, n, C.c_float(0.001), buf.hits,
                                              buf.stats))
    abi.check(lib.gt_context_synchronize(ctx.handle))
    hits = np.zeros((n, 8), np.float32)
    stats = np.zeros(7, np.uint64)
    abi.check(lib.gt_memcpy_to_host(ctx.handle, hits.ctypes.data_as(C.c_void_p), buf.hits, hits.nbytes))
    abi.check(lib.gt_memcpy_to_host(ctx.handle, stats.ctypes.data_as(C.c_void_p), buf.stats, stats.nbytes))
    return hits, stats


def _scene_cases():
    spheres = scenes.randomized_ray_scene(3).as_float32_exact()
    terrain = scenes.terrain(grid=96, n_spheres=200, seed=5)[0].as_float32_exact()
    return [("randomized spheres", spheres, (-4, -3, -9), (4, 3, 1)),
            ("terrain 96x96", terrain, (-52, -1, -52), (52, 9, 52))]


@pytest.mark.parametrize("case", range(2))
def test_noalloc_intersect_matches_gt_intersect(gpu_context, case):
    name, data, lo, hi = _scene_cases()[case]
    rs = api.RayScene(data).with_bvh()
    h = rs._compile()
    ctx = rs.context
    rng = np.random.default_rng(11 + case)
    sizes = (200_000, 70_001)  # the second call is smaller
    buf = DeviceBuffers(ctx, max(sizes))
    try:
        for k, n in enumerate(sizes):
            say(f"noalloc intersect: {name}, call {k}, {n} rays")
            rays = _rays(rng, n, lo, hi)
            ref_hits, ref_stats = _host_intersect(ctx, h, rays)
            hits, stats = _noalloc_intersect(ctx, h, rays, buf)
            assert (ref_hits[:, 3].view(np.int32) >= 0).mean() > 0.05
            np.testing.assert_array_equal(hits.view(np.uint32), ref_hits.view(np.uint32))
            np.testing.assert_array_equal(stats, ref_stats)
            assert stats[0] == n
    finally:
        buf.free()


def test_noalloc_intersect_requires_a_built_scene(gpu_context):
    say("noalloc intersect: unbuilt scene")
    ctx = gpu_context
    lib = ctx.lib
    h = C.c_void_p()
    abi.check(lib.gt_scene_create(ctx.handle, C.byref(h)))
    try:
        mid = C.c_uint32()
        abi.check(lib.gt_scene_add_lambertian(h, 0.5, 0.5, 0.5, C.byref(mid)))
        sph = np.array([0.0, 0.0, -1.0, 0.5])
        abi.check(lib.gt_scene_add_spheres(h, sph.ctypes.data_as(abi.D), 1, C.byref(mid)))
        rc = lib.gt_intersect_device_noalloc(ctx.handle, h, None, 0, C.c_float(0.001), None, None)
        assert rc == abi.GT_ESTATE
        assert b"not built" in lib.gt_last_error()
    finally:
        lib.gt_scene_destroy(h)
