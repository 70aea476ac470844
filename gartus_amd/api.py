"""Python mirror of gartus's path-tracing API over the HIP library.

Names, argument meaning and error behaviour follow the reference's
``PathTracer`` / ``RayCamera`` / ``RayScene`` / ``SamplingTargetList`` /
``EnvironmentLight`` / ``HdrImage`` / ``ToneMap``
(src/graphics/raytracing/renderer.rs, src/graphics/camera.rs,
src/graphics/raytracing/scene.rs, environment.rs, src/graphics/display.rs).
Where the reference panics on an invalid argument, this mirror raises
``InvalidArgument`` (a ValueError).

Every render runs on the GPU through ``libgartus_hip.so``; a missing library
or device raises, there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import math
import weakref
from dataclasses import dataclass, replace
from typing import Optional

import numpy as np

from . import abi
from .scene_data import (DIELECTRIC, DIFFUSE_LIGHT, GGX, HENYEY_GREENSTEIN, ISOTROPIC, LAMBERTIAN, MESH_TRIANGLE, METAL,
                         MOVING_SPHERE, QUAD, SPHERE, TRIANGLE, Material, SceneData)

# ---------------------------------------------------------------- context ---

_CONTEXT = None


class Context:
    """One HIP device and stream (gt_context). Scenes, light lists and
    environments compiled on it belong to it: closing the context releases
    them first (a native object must not outlive its context)."""

    def __init__(self, device: int = 0):
        self.lib = abi.load()
        h = C.c_void_p()
        abi.check(self.lib.gt_context_create(device, C.byref(h)))
        self.handle = h
        self._owned = weakref.WeakSet()

    def _adopt(self, obj) -> None:
        self._owned.add(obj)

    @property
    def tuning(self) -> abi.GtTuning:
        t = abi.GtTuning()
        abi.check(self.lib.gt_context_get_tuning(self.handle, C.byref(t)))
        return t

    def set_tuning(self, **kw) -> None:
        t = self.tuning
        for k, v in kw.items():
            setattr(t, k, v)
        abi.check(self.lib.gt_context_set_tuning(self.handle, C.byref(t)))

    def stream(self) -> int:
        return self.lib.gt_context_stream(self.handle) or 0

    def synchronize(self) -> None:
        abi.check(self.lib.gt_context_synchronize(self.handle))

    def close(self) -> None:
        if self.handle:
            for obj in list(self._owned):
                obj._release()
            self.lib.gt_context_destroy(self.handle)
            self.handle = None


def default_context() -> Context:
    global _CONTEXT
    if _CONTEXT is None:
        _CONTEXT = Context(0)
    return _CONTEXT


def _dp(a: np.ndarray):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def _v3(v) -> tuple:
    t = tuple(float(x) for x in v)
    if len(t) != 3:
        raise abi.InvalidArgument(abi.GT_EINVAL, "expected three components")
    return t


# -------------------------------------------------------------- materials ---

class RayMaterial:
    """RayMaterial constructors (material.rs:2355-2582) for the five
    RayScene BSDFs of the GPU path."""

    @staticmethod
    def lambertian(albedo) -> Material:
        c = _v3(albedo)
        if not all(math.isfinite(x) for x in c):
            raise abi.InvalidArgument(abi.GT_EINVAL, "Lambertian albedo must be finite")
        return Material(LAMBERTIAN, c)

    @staticmethod
    def diffuse_light(emit) -> Material:
        c = _v3(emit)
        if not all(math.isfinite(x) for x in c):
            raise abi.InvalidArgument(abi.GT_EINVAL, "diffuse light color must be finite")
        return Material(DIFFUSE_LIGHT, c)

    @staticmethod
    def ggx_microfacet(specular_color, roughness: float) -> Material:
        c = _v3(specular_color)
        if not (all(math.isfinite(x) for x in c) and math.isfinite(roughness)):
            raise abi.InvalidArgument(abi.GT_EINVAL, "GGX microfacet material values must be finite")
        return Material(GGX, c, min(max(float(roughness), 0.0), 1.0))

    @staticmethod
    def metal(albedo, fuzz: float) -> Material:
        c = _v3(albedo)
        if not (all(math.isfinite(x) for x in c) and math.isfinite(fuzz)):
            raise abi.InvalidArgument(abi.GT_EINVAL, "metal material values must be finite")
        return Material(METAL, c, min(max(float(fuzz), 0.0), 1.0))

    @staticmethod
    def isotropic(albedo) -> Material:
        """Isotropic::new (material.rs:1061-1066)."""
        c = _v3(albedo)
        if not all(math.isfinite(x) for x in c):
            raise abi.InvalidArgument(abi.GT_EINVAL, "isotropic albedo must be finite")
        return Material(ISOTROPIC, c)

    @staticmethod
    def henyey_greenstein(albedo, g: float) -> Material:
        """HenyeyGreenstein::new (material.rs:1194-1204)."""
        c = _v3(albedo)
        if not all(math.isfinite(x) for x in c):
            raise abi.InvalidArgument(abi.GT_EINVAL, "Henyey-Greenstein albedo must be finite")
        if not (math.isfinite(g) and abs(g) < 1.0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "Henyey-Greenstein anisotropy must be finite and in (-1, 1)")
        return Material(HENYEY_GREENSTEIN, c, float(g))

    @staticmethod
    def dielectric(refraction_index: float) -> Material:
        return Material(DIELECTRIC, (1.0, 1.0, 1.0), float(refraction_index))


# ----------------------------------------------------------------- scene ---

class RayScene:
    """RayScene (scene.rs:1102-1578) compiled for the GPU on first use."""

    def __init__(self, data: Optional[SceneData] = None, context: Optional[Context] = None):
        self.data = data if data is not None else SceneData()
        self._ctx = context
        self._handle = None
        self._leaf_size = None   # None: not built (lazy default build on first hit)
        self._on_device = False  # build with gt_scene_build_device
        self._dirty = True

    # construction ------------------------------------------------------
    def add_material(self, material: Material) -> int:
        self._dirty = True
        return self.data.add_material(material)

    def _add(self, kind, data, material) -> None:
        try:
            self.data.add(kind, data, material)
        except ValueError as e:
            raise abi.InvalidArgument(abi.GT_EINVAL, str(e)) from None
        self._dirty = True
        self._leaf_size = None   # invalidate_bvh (scene.rs:1575-1577)

    def add_sphere(self, center, radius, material: int) -> None:
        self._add(SPHERE, [*_v3(center), float(radius)], material)

    def add_spheres(self, material: int, spheres) -> None:
        self._add(SPHERE, spheres, material)

    def add_moving_sphere(self, center_start, center_end, radius, material: int) -> None:
        self._add(MOVING_SPHERE, [*_v3(center_start), *_v3(center_end), float(radius)], material)

    def add_triangle(self, p0, p1, p2, material: int) -> None:
        self._add(TRIANGLE, [*_v3(p0), *_v3(p1), *_v3(p2)], material)

    def add_triangles(self, material, triangles) -> None:
        self._add(TRIANGLE, triangles, material)

    def add_mesh_triangle(self, p0, p1, p2, n0, n1, n2, material: int) -> None:
        """A TriangleMesh triangle with vertex normals (MeshTriangle::with_vertex_normals,
        mesh.rs:40-50): shaded with the interpolated normal (mesh.rs:105-119)."""
        self._add(MESH_TRIANGLE, [*_v3(p0), *_v3(p1), *_v3(p2), *_v3(n0), *_v3(n1), *_v3(n2)], material)

    def add_mesh_triangles(self, material, positions, normals) -> None:
        """Many mesh triangles: positions (n, 9) and vertex normals (n, 9)."""
        pos = np.asarray(positions, np.float64).reshape(-1, 9)
        nrm = np.asarray(normals, np.float64).reshape(-1, 9)
        if len(pos) != len(nrm):
            raise abi.InvalidArgument(abi.GT_EINVAL, "one set of vertex normals per triangle")
        self._add(MESH_TRIANGLE, np.concatenate([pos, nrm], axis=1), material)

    def add_quad(self, corner, u, v, material: int) -> None:
        self._add(QUAD, [*_v3(corner), *_v3(u), *_v3(v)], material)

    def add_quads(self, material, quads) -> None:
        self._add(QUAD, quads, material)

    def add_sphere_with_material(self, center, radius, material: Material) -> int:
        mid = self.add_material(material)
        self.add_sphere(center, radius, mid)
        return mid

    def add_quad_with_material(self, corner, u, v, material: Material) -> int:
        mid = self.add_material(material)
        self.add_quad(corner, u, v, mid)
        return mid

    def build_bvh(self) -> None:
        self.build_bvh_with_options(4)

    def build_bvh_with_options(self, leaf_size: int = 4) -> None:
        self._leaf_size = max(int(leaf_size), 1)
        self._on_device = False
        self._compile(force_build=True)

    def with_bvh(self) -> "RayScene":
        self.build_bvh()
        return self

    def build_bvh_on_device(self, leaf_size: int = 4) -> None:
        """gt_scene_build_device: the scene compile on the GPU (a linear BVH
        over Morton codes; no reference counterpart). Closest hits are those
        of build_bvh; the tree, and so the traversal counts, differ."""
        self._leaf_size = max(int(leaf_size), 1)
        self._on_device = True
        self._compile(force_build=True)

    def with_device_bvh(self, leaf_size: int = 4) -> "RayScene":
        self.build_bvh_on_device(leaf_size)
        return self

    # queries -------------------------------------------------------------
    def __len__(self) -> int:
        return len(self.data)

    def material_count(self) -> int:
        return len(self.data.materials)

    def emissive_targets(self) -> "SamplingTargetList":
        t = SamplingTargetList()
        for kind, q, _ in self.data.emissive_targets():
            t._add(kind, q, 1.0)
        return t

    @property
    def context(self) -> Context:
        if self._ctx is None:
            self._ctx = default_context()
        return self._ctx

    def _compile(self, force_build: bool = False):
        ctx = self.context
        lib = ctx.lib
        if self._handle is None or self._dirty:
            if self._handle is not None:
                lib.gt_scene_destroy(self._handle)
            h = C.c_void_p()
            abi.check(lib.gt_scene_create(ctx.handle, C.byref(h)))
            self._handle = h
            ctx._adopt(self)
            mid = C.c_uint32()
            for m in self.data.materials:
                r, g, b = m.color
                if m.kind == LAMBERTIAN:
                    rc = lib.gt_scene_add_lambertian(h, r, g, b, C.byref(mid))
                elif m.kind == DIFFUSE_LIGHT:
                    rc = lib.gt_scene_add_diffuse_light(h, r, g, b, C.byref(mid))
                elif m.kind == GGX:
                    rc = lib.gt_scene_add_ggx(h, r, g, b, m.param, C.byref(mid))
                elif m.kind == METAL:
                    rc = lib.gt_scene_add_metal(h, r, g, b, m.param, C.byref(mid))
                elif m.kind == ISOTROPIC:
                    rc = lib.gt_scene_add_isotropic(h, r, g, b, C.byref(mid))
                elif m.kind == HENYEY_GREENSTEIN:
                    rc = lib.gt_scene_add_henyey_greenstein(h, r, g, b, m.param, C.byref(mid))
                else:
                    rc = lib.gt_scene_add_dielectric(h, m.param, C.byref(mid))
                abi.check(rc)
            adders = {SPHERE: lib.gt_scene_add_spheres, MOVING_SPHERE: lib.gt_scene_add_moving_spheres,
                      TRIANGLE: lib.gt_scene_add_triangles, QUAD: lib.gt_scene_add_quads}
            for kind, data, mids in self.data.chunks:
                if kind == MESH_TRIANGLE:
                    pos = np.ascontiguousarray(data[:, :9])
                    nrm = np.ascontiguousarray(data[:, 9:])
                    abi.check(lib.gt_scene_add_mesh_triangles(h, _dp(pos), _dp(nrm), len(data),
                                                              mids.ctypes.data_as(abi.U32)))
                    continue
                abi.check(adders[kind](h, _dp(data), len(data), mids.ctypes.data_as(abi.U32)))
            self._dirty = False
            if self._leaf_size is not None:
                build = lib.gt_scene_build_device if self._on_device else lib.gt_scene_build
                abi.check(build(h, self._leaf_size))
        elif force_build:
            build = lib.gt_scene_build_device if self._on_device else lib.gt_scene_build
            abi.check(build(self._handle, self._leaf_size))
        return self._handle

    def info(self) -> abi.GtSceneInfo:
        h = self._compile()
        info = abi.GtSceneInfo()
        abi.check(self.context.lib.gt_scene_get_info(h, C.byref(info)))
        return info

    def bvh_node_count(self) -> Optional[int]:
        if self._leaf_size is None:
            return None
        return int(self.info().bvh_node_count)

    def intersect(self, origins, directions, times=None, t_min: float = 0.001, t_max=np.inf,
                  with_stats: bool = False):
        """Closest hits of many rays (RayScene::hit, scene.rs:1639-1657) with
        the primitive id; returns a dict of numpy arrays (and stats)."""
        o = np.asarray(origins, np.float32).reshape(-1, 3)
        d = np.asarray(directions, np.float32).reshape(-1, 3)
        n = len(o)
        rays = np.zeros((n, 8), np.float32)
        rays[:, 0:3] = o
        rays[:, 3] = 0.0 if times is None else np.asarray(times, np.float32)
        rays[:, 4:7] = d
        rays[:, 7] = np.broadcast_to(np.asarray(t_max, np.float32), (n,))
        hits = np.zeros((n, 8), np.float32)
        h = self._compile()
        st = abi.GtTraversalStats()
        abi.check(self.context.lib.gt_intersect(
            self.context.handle, h, rays.ctypes.data_as(C.POINTER(abi.GtRay)), n, float(t_min),
            hits.ctypes.data_as(C.POINTER(abi.GtHit)), C.byref(st) if with_stats else None))
        out = {
            "t": hits[:, 0].copy(), "u": hits[:, 1].copy(), "v": hits[:, 2].copy(),
            "prim": hits[:, 3].view(np.int32).copy(), "normal": hits[:, 4:7].copy(),
            "front": hits[:, 7].view(np.uint32).astype(bool),
        }
        if with_stats:
            out["stats"] = {k: int(getattr(st, k)) for k, _ in abi.GtTraversalStats._fields_}
        return out

    def close(self) -> None:
        self._release()

    def _release(self) -> None:
        """Destroys the native scene while its context is alive (a closed
        context has released it already)."""
        if self._handle is not None and self._ctx is not None and self._ctx.handle:
            self._ctx.lib.gt_scene_destroy(self._handle)
        self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class WeightedSamplingTargetList:
    """WeightedSamplingTargetList (scene.rs:458-609)."""

    def __init__(self):
        self.targets = []       # (kind, params, weight)
        self.total_weight = 0.0
        self._handle = None
        self._ctx = None

    def _add(self, kind, q, weight: float) -> None:
        weight = float(weight)
        if not (math.isfinite(weight) and weight > 0.0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "sampling target weight must be finite and positive")
        if not math.isfinite(self.total_weight + weight):
            raise abi.InvalidArgument(abi.GT_EINVAL, "sampling target total weight must remain finite")
        self.total_weight += weight
        self.targets.append((kind, np.asarray(q, np.float64), weight))
        self._drop()

    def add_sphere_weighted(self, center, radius, weight) -> None:
        self._add(SPHERE, [*_v3(center), max(float(radius), 0.0)], weight)

    def add_moving_sphere_weighted(self, c0, c1, radius, weight) -> None:
        self._add(MOVING_SPHERE, [*_v3(c0), *_v3(c1), max(float(radius), 0.0)], weight)

    def add_quad_weighted(self, corner, u, v, weight) -> None:
        self._add(QUAD, [*_v3(corner), *_v3(u), *_v3(v)], weight)

    def add_triangle_weighted(self, p0, p1, p2, weight) -> None:
        self._add(TRIANGLE, [*_v3(p0), *_v3(p1), *_v3(p2)], weight)

    def __len__(self) -> int:
        return len(self.targets)

    def is_empty(self) -> bool:
        return not self.targets

    def clear(self) -> None:
        self.targets = []
        self.total_weight = 0.0
        self._drop()

    def _drop(self) -> None:
        if self._handle is not None and self._ctx is not None and self._ctx.handle:
            self._ctx.lib.gt_lights_destroy(self._handle)
        self._handle = None

    _release = _drop

    def _compile(self, ctx: Context):
        if self._handle is not None and self._ctx is ctx:
            return self._handle
        self._drop()
        h = C.c_void_p()
        abi.check(ctx.lib.gt_lights_create(ctx.handle, C.byref(h)))
        for kind, q, w in self.targets:
            arr = (C.c_double * 9)(*list(q) + [0.0] * (9 - len(q)))
            a = C.cast(arr, abi.D)
            b = C.cast(C.byref(arr, 24), abi.D)
            c = C.cast(C.byref(arr, 48), abi.D)
            if kind == SPHERE:
                rc = ctx.lib.gt_lights_add_sphere(h, a, q[3], w)
            elif kind == MOVING_SPHERE:
                rc = ctx.lib.gt_lights_add_moving_sphere(h, a, b, q[6], w)
            elif kind == QUAD:
                rc = ctx.lib.gt_lights_add_quad(h, a, b, c, w)
            else:
                rc = ctx.lib.gt_lights_add_triangle(h, a, b, c, w)
            abi.check(rc)
        self._handle, self._ctx = h, ctx
        ctx._adopt(self)
        return h

    def __del__(self):
        try:
            self._drop()
        except Exception:
            pass


class SamplingTargetList(WeightedSamplingTargetList):
    """SamplingTargetList (scene.rs:363-450): every target has weight 1."""

    def add_sphere(self, center, radius) -> None:
        self.add_sphere_weighted(center, radius, 1.0)

    def add_moving_sphere(self, c0, c1, radius) -> None:
        self.add_moving_sphere_weighted(c0, c1, radius, 1.0)

    def add_quad(self, corner, u, v) -> None:
        self.add_quad_weighted(corner, u, v, 1.0)

    def add_triangle(self, p0, p1, p2) -> None:
        self.add_triangle_weighted(p0, p1, p2, 1.0)


class EnvironmentLight:
    """EnvironmentLight (environment.rs:22-195): a constant colour, or the
    FP32 HDR lat-long form (rows from +y down to -y, linear RGB)."""

    def __init__(self, rgb: Optional[np.ndarray] = None, color=None):
        self.rgb = None if rgb is None else np.ascontiguousarray(np.asarray(rgb, np.float32))
        self.color = None if color is None else _v3(color)
        if self.rgb is not None and (self.rgb.ndim != 3 or self.rgb.shape[2] != 3 or self.rgb.size == 0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "environment map must be an (H, W, 3) array")
        self._handle = None
        self._ctx = None

    @staticmethod
    def constant(color) -> "EnvironmentLight":
        return EnvironmentLight(color=color)

    @staticmethod
    def from_rgb32f(rgb) -> "EnvironmentLight":
        return EnvironmentLight(rgb=rgb)

    def _compile(self, ctx: Context):
        if self._handle is not None and self._ctx is ctx:
            return self._handle
        h = C.c_void_p()
        if self.rgb is None:
            abi.check(ctx.lib.gt_env_create_constant(ctx.handle, *self.color, C.byref(h)))
        else:
            hh, ww, _ = self.rgb.shape
            abi.check(ctx.lib.gt_env_create_rgb32f(ctx.handle, self.rgb.ctypes.data_as(abi.F), ww, hh, C.byref(h)))
        self._release()
        self._handle, self._ctx = h, ctx
        ctx._adopt(self)
        return h

    def _release(self) -> None:
        if self._handle is not None and self._ctx is not None and self._ctx.handle:
            self._ctx.lib.gt_env_destroy(self._handle)
        self._handle = None

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


# ---------------------------------------------------------------- camera ---

class PixelSampleMode:
    RANDOM = abi.GT_PIXEL_RANDOM
    STRATIFIED = abi.GT_PIXEL_STRATIFIED


class DirectLightingMode:
    CURRENT_PATH_CONTINUATION = abi.GT_LIGHTING_CONTINUATION
    NEXT_EVENT_ESTIMATION = abi.GT_LIGHTING_NEE


class RayCamera:
    """RayCamera builder (camera.rs:1034-1481). Builders return a new camera;
    invalid parameters raise InvalidArgument where the reference panics."""

    def __init__(self, image_width: int = 100, aspect_ratio: float = 1.0):
        if int(image_width) <= 0:
            raise abi.InvalidArgument(abi.GT_EINVAL, "image width must be positive")
        if not (math.isfinite(aspect_ratio) and aspect_ratio > 0.0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "aspect ratio must be positive and finite")
        self.raw = abi.default_camera()
        self.raw.image_width = int(image_width)
        self.raw.aspect_ratio = float(aspect_ratio)
        self._validate()

    @classmethod
    def new(cls, image_width: int, aspect_ratio: float) -> "RayCamera":
        return cls(image_width, aspect_ratio)

    def _copy(self) -> "RayCamera":
        c = RayCamera.__new__(RayCamera)
        c.raw = abi.GtCamera.from_buffer_copy(self.raw)
        return c

    def _validate(self) -> None:  # validate_view (camera.rs:1161-1201)
        r = self.raw
        if not (math.isfinite(r.vertical_fov) and 0.0 < r.vertical_fov < 180.0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "vertical field of view must be finite and in 0..180 degrees")
        w = np.array(r.lookfrom[:]) - np.array(r.lookat[:])
        if not float(w @ w) > np.finfo(float).eps:
            raise abi.InvalidArgument(abi.GT_EINVAL, "lookfrom and lookat must be distinct")
        up = np.array(r.view_up[:])
        if not float(up @ up) > np.finfo(float).eps:
            raise abi.InvalidArgument(abi.GT_EINVAL, "view-up vector must be nonzero")
        fwd = -w / np.linalg.norm(w)
        right = np.cross(fwd, up)
        if np.linalg.norm(right) < np.finfo(float).eps:
            raise abi.InvalidArgument(abi.GT_EINVAL, "view-up vector must not be parallel to the viewing direction")
        if not (math.isfinite(r.defocus_angle) and 0.0 <= r.defocus_angle < 180.0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "defocus angle must be finite and in 0..180 degrees")
        if not (math.isfinite(r.focus_distance) and r.focus_distance > 0.0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "focus distance must be positive and finite")
        if not (math.isfinite(r.shutter_start) and math.isfinite(r.shutter_end) and r.shutter_start <= r.shutter_end):
            raise abi.InvalidArgument(abi.GT_EINVAL, "shutter interval must be finite and ordered")
        nbg = 9 if r.background_kind == abi.GT_BACKGROUND_VERTICAL_GRADIENT else 3
        if not all(math.isfinite(x) for x in r.background[:nbg]):
            raise abi.InvalidArgument(abi.GT_EINVAL, "background color components must be finite")

    def _with(self, validate=True, **kw) -> "RayCamera":
        c = self._copy()
        for k, v in kw.items():
            if k in ("lookfrom", "lookat", "view_up"):
                getattr(c.raw, k)[:] = _v3(v)
            else:
                setattr(c.raw, k, v)
        if validate:
            c._validate()
        return c

    def with_image_width(self, w: int):
        if int(w) <= 0:
            raise abi.InvalidArgument(abi.GT_EINVAL, "image width must be positive")
        return self._with(image_width=int(w))

    def with_aspect_ratio(self, a: float):
        if not (math.isfinite(a) and a > 0.0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "aspect ratio must be positive and finite")
        return self._with(aspect_ratio=float(a))

    def with_samples_per_pixel(self, n: int):
        return self._with(samples_per_pixel=max(int(n), 1))

    def with_pixel_sample_mode(self, mode: int):
        return self._with(pixel_sample_mode=int(mode))

    def with_stratified_sampling(self):
        return self.with_pixel_sample_mode(abi.GT_PIXEL_STRATIFIED)

    def with_stratified_grid_width(self, g: int):
        g = max(int(g), 1)
        return self._with(samples_per_pixel=g * g, pixel_sample_mode=abi.GT_PIXEL_STRATIFIED_GRID,
                          stratified_grid_width=g)

    def with_max_depth(self, d: int):
        return self._with(max_depth=int(d))

    def with_russian_roulette_min_depth(self, d: int):
        return self._with(russian_roulette_min_depth=int(d))

    def without_russian_roulette(self):
        return self._with(russian_roulette_min_depth=-1)

    def with_rng_seed(self, seed: int):
        return self._with(rng_seed=int(seed) & (2 ** 64 - 1))

    def with_vertical_fov(self, fov: float):
        return self._with(vertical_fov=float(fov))

    def with_look_at(self, lookfrom, lookat):
        return self._with(lookfrom=lookfrom, lookat=lookat)

    def with_view_up(self, up):
        return self._with(view_up=up)

    def with_defocus_angle(self, a: float):
        return self._with(defocus_angle=float(a))

    def with_focus_distance(self, d: float):
        return self._with(focus_distance=float(d))

    def with_shutter_interval(self, start: float, end: float):
        if not (math.isfinite(start) and math.isfinite(end) and start <= end):
            raise abi.InvalidArgument(abi.GT_EINVAL, "shutter interval must be finite and ordered")
        return self._with(shutter_start=float(start), shutter_end=float(end))

    def with_background(self, color):
        c = self._with(validate=False, background_kind=abi.GT_BACKGROUND_CONSTANT)
        c.raw.background[0:3] = _v3(color)
        c._validate()
        return c

    def with_vertical_gradient_background(self, nadir, horizon, zenith):
        c = self._with(validate=False, background_kind=abi.GT_BACKGROUND_VERTICAL_GRADIENT)
        c.raw.background[0:9] = (*_v3(nadir), *_v3(horizon), *_v3(zenith))
        c._validate()
        return c

    def with_direct_lighting_mode(self, mode: int):  # SamplingStrategy::from_direct_lighting_mode
        w = 0.0 if mode == abi.GT_LIGHTING_NEE else 0.5
        return self._with(direct_lighting_mode=int(mode), light_pdf_weight=w)

    def with_light_pdf_weight(self, w: float):  # SamplingStrategy::with_light_pdf_weight
        if not math.isfinite(w):
            raise abi.InvalidArgument(abi.GT_EINVAL, "light PDF weight must be finite")
        return self._with(direct_lighting_mode=abi.GT_LIGHTING_CONTINUATION, light_pdf_weight=min(max(w, 0.0), 1.0))

    def with_adaptive_sampling(self, min_samples: int, max_samples: int, error_threshold: float):
        """with_adaptive_sampling (camera.rs:1244-1264): AdaptiveSampling::new
        (:377-393) raises min to 1 and max to min; spp becomes max."""
        if not (math.isfinite(error_threshold) and error_threshold > 0.0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "adaptive sampling error threshold must be positive and finite")
        mn = max(int(min_samples), 1)
        mx = max(int(max_samples), mn)
        return self._with(adaptive_min_samples=mn, adaptive_max_samples=mx,
                          adaptive_error_threshold=float(error_threshold), samples_per_pixel=mx)

    def without_adaptive_sampling(self):  # camera.rs:1266-1271
        return self._with(adaptive_min_samples=0, adaptive_max_samples=0, adaptive_error_threshold=0.0)

    def adaptive_sampling(self):
        """(min_samples, max_samples, error_threshold), or None (camera.rs:1513-1517)."""
        r = self.raw
        if r.adaptive_error_threshold == 0.0:
            return None
        return int(r.adaptive_min_samples), int(r.adaptive_max_samples), float(r.adaptive_error_threshold)

    def image_width(self) -> int:
        return int(self.raw.image_width)

    def image_height(self) -> int:  # camera.rs:1071
        return max(int(self.raw.image_width / self.raw.aspect_ratio), 1)

    def samples_per_pixel(self) -> int:
        return int(self.raw.samples_per_pixel)

    def effective_samples_per_pixel(self) -> int:  # camera.rs:1547-1557
        r = self.raw
        s = max(int(r.samples_per_pixel), 1)
        if r.pixel_sample_mode == abi.GT_PIXEL_RANDOM:
            return s
        if r.pixel_sample_mode == abi.GT_PIXEL_STRATIFIED:
            g = max(int(math.sqrt(s)), 1)
            return g * g
        g = max(int(r.stratified_grid_width), 1)
        return g * g

    def max_depth(self) -> int:
        return int(self.raw.max_depth)


# ----------------------------------------------------------------- images ---

class ToneMappingOperator:
    NONE, REINHARD, ACES = 0, 1, 2


def _encode(pixels: np.ndarray, width: int, height: int, fmt: int, tone_map: Optional["ToneMap"] = None,
            context: Optional[Context] = None) -> np.ndarray:
    """gt_encode_image on a host image: float32 or float64 [h][w][3] in, the
    format's body out (uint8 [h][w][3|4] or float32 [h][w][3])."""
    ctx = context or default_context()
    src = np.ascontiguousarray(pixels)
    if src.dtype not in (np.float32, np.float64):
        src = src.astype(np.float64)
    if src.size != width * height * 3:
        raise abi.InvalidArgument(abi.GT_EINVAL, "HDR image pixel count must match dimensions")
    chans = {abi.GT_ENCODE_RGBE: 4}.get(fmt, 3)
    out = np.zeros((height, width, chans), np.float32 if fmt == abi.GT_ENCODE_PFM else np.uint8)
    tm = tone_map._raw() if tone_map is not None else None
    abi.check(ctx.lib.gt_encode_image(ctx.handle, src.ctypes.data_as(C.c_void_p), int(src.dtype == np.float64), 0,
                                      width, height, fmt, C.byref(tm) if tm is not None else None,
                                      out.ctypes.data_as(C.c_void_p), 0))
    return out


def encoded_header(fmt: int, width: int, height: int) -> bytes:
    """The file header the reference writes before an encoded body."""
    lib = abi.load()
    n = C.c_size_t()
    buf = C.create_string_buffer(128)
    abi.check(lib.gt_encoded_header(fmt, width, height, buf, 128, C.byref(n)))
    return buf.raw[:n.value]


@dataclass(frozen=True)
class ToneMap:
    """ToneMap (display.rs:191-278)."""
    exposure: float = 1.0
    operator: int = ToneMappingOperator.NONE
    gamma: float = 2.0

    def with_exposure(self, e: float) -> "ToneMap":
        if not (math.isfinite(e) and e >= 0.0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "tone-map exposure must be finite and non-negative")
        return replace(self, exposure=float(e))

    def with_operator(self, op: int) -> "ToneMap":
        return replace(self, operator=int(op))

    def with_gamma(self, g: float) -> "ToneMap":
        if not (math.isfinite(g) and g > 0.0):
            raise abi.InvalidArgument(abi.GT_EINVAL, "tone-map gamma must be positive and finite")
        return replace(self, gamma=float(g))

    def _raw(self) -> abi.GtToneMap:
        return abi.GtToneMap(self.exposure, self.gamma, self.operator, 0)

    def map(self, rgb: np.ndarray, context: Optional[Context] = None) -> np.ndarray:
        """ToneMap::map_color per channel (display.rs:264-277) -> uint8, on
        the device; rgb is any array of 3-channel pixels."""
        a = np.asarray(rgb)
        if a.size % 3:
            raise abi.InvalidArgument(abi.GT_EINVAL, "pixels must have 3 channels")
        n = a.size // 3
        return _encode(a.reshape(-1, 3), n, 1, abi.GT_ENCODE_CANVAS, self, context).reshape(a.shape)


class HdrImage:
    """HdrImage (display.rs:282-441): linear RGB, row-major, top row first.
    Conversions and file bodies are computed by the library on the device of
    `context` (the rendering context, else the default one)."""

    def __init__(self, width: int, height: int, pixels: np.ndarray, context: Optional[Context] = None):
        pixels = np.asarray(pixels)
        if pixels.size != width * height * 3:
            raise abi.InvalidArgument(abi.GT_EINVAL, "HDR image pixel count must match dimensions")
        self.width, self.height = int(width), int(height)
        self.pixels = pixels.reshape(height, width, 3)
        self.context = context

    def _enc(self, fmt: int, tone_map: Optional[ToneMap] = None) -> np.ndarray:
        return _encode(self.pixels, self.width, self.height, fmt, tone_map, self.context)

    def to_canvas(self) -> np.ndarray:
        return self.to_canvas_tone_mapped(ToneMap())

    def to_canvas_tone_mapped(self, tone_map: ToneMap) -> np.ndarray:
        return self._enc(abi.GT_ENCODE_CANVAS, tone_map)

    def pfm_bytes(self) -> bytes:
        """write_pfm (display.rs:365-387): little-endian f32, bottom row first,
        non-finite values written as 0."""
        return encoded_header(abi.GT_ENCODE_PFM, self.width, self.height) + \
            self._enc(abi.GT_ENCODE_PFM).astype("<f4").tobytes()

    def radiance_hdr_bytes(self) -> bytes:
        """write_radiance_hdr (display.rs:397-409): flat RGBE, top row first."""
        return encoded_header(abi.GT_ENCODE_RGBE, self.width, self.height) + self._enc(abi.GT_ENCODE_RGBE).tobytes()

    def save_pfm(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(self.pfm_bytes())

    def save_radiance_hdr(self, path: str) -> None:
        with open(path, "wb") as f:
            f.write(self.radiance_hdr_bytes())

    def save_extension(self, path: str) -> None:
        """save_extension (display.rs:412-431): .pfm and .hdr/.rgbe keep HDR
        data; .ppm writes the default tone-mapped canvas (binary P6). The
        reference's other raster encoders (PNG, JPEG, ...) belong to its
        Canvas and are not part of this library."""
        ext = path.rsplit(".", 1)[-1].lower() if "." in path else ""
        if ext == "pfm":
            self.save_pfm(path)
        elif ext in ("hdr", "rgbe"):
            self.save_radiance_hdr(path)
        elif ext == "ppm":
            self.save_tone_mapped(path, ToneMap())
        else:
            raise abi.InvalidArgument(abi.GT_EINVAL, f"unsupported image extension: {path}")

    def save_tone_mapped(self, path: str, tone_map: ToneMap) -> None:
        canvas = self.to_canvas_tone_mapped(tone_map)
        with open(path, "wb") as f:
            f.write(f"P6\n{self.width} {self.height}\n255\n".encode() + canvas.tobytes())


@dataclass
class DenoisingAovs:
    """DenoisingAovs (camera.rs:491-526) with the canvases of
    camera.rs:4212-4243, encoded on the device."""
    beauty_linear: np.ndarray
    albedo_linear: np.ndarray
    normal_linear: np.ndarray
    context: Optional[Context] = None

    def _canvas(self, a: np.ndarray, fmt: int) -> np.ndarray:
        h, w = a.shape[0], a.shape[1]
        return _encode(a, w, h, fmt, None, self.context)

    @property
    def beauty(self) -> np.ndarray:  # Rgb::from_linear_color (gamma 2)
        return self._canvas(self.beauty_linear, abi.GT_ENCODE_CANVAS)

    @property
    def albedo(self) -> np.ndarray:  # Rgb::from_raw_linear_color
        return self._canvas(self.albedo_linear, abi.GT_ENCODE_RAW)

    @property
    def normal(self) -> np.ndarray:
        return self._canvas(self.normal_linear, abi.GT_ENCODE_RAW)


# ------------------------------------------------------------------ render ---

@dataclass(frozen=True)
class RenderTile:
    """RenderTile (camera.rs:398-424)."""
    x: int
    y: int
    width: int
    height: int

    def x_end(self) -> int:
        return self.x + self.width

    def y_end(self) -> int:
        return self.y + self.height


@dataclass(frozen=True)
class RenderProgress:
    """RenderProgress (camera.rs:430-447)."""
    tile: RenderTile
    completed_tiles: int
    total_tiles: int

    def is_complete(self) -> bool:
        return self.completed_tiles >= self.total_tiles


class ProgressiveRenderUpdate:
    """ProgressiveRenderUpdate (camera.rs:451-490): the tile counters and the
    canvas so far (a read-only view, valid during the callback)."""

    def __init__(self, raw: abi.GtRenderProgress, canvas: np.ndarray):
        t = raw.tile
        self._progress = RenderProgress(RenderTile(t.x, t.y, t.width, t.height), int(raw.completed_tiles),
                                        int(raw.total_tiles))
        self._w, self._h = int(raw.image_width), int(raw.image_height)
        self._pixels = canvas.view()
        self._pixels.flags.writeable = False

    def progress(self) -> RenderProgress:
        return self._progress

    def image_width(self) -> int:
        return self._w

    def image_height(self) -> int:
        return self._h

    def pixels(self) -> np.ndarray:
        return self._pixels


@dataclass(frozen=True)
class RenderOptions:
    """RenderOptions (renderer.rs:20-45). Tile size is accepted for API
    parity; the GPU path schedules whole-image wavefronts and its result does
    not depend on it, as the reference's does not."""
    tile_size: Optional[int] = None

    def with_tile_size(self, t: int) -> "RenderOptions":
        return RenderOptions(max(int(t), 1))


class PathTracer:
    """PathTracer (renderer.rs:47-887) over RayScene worlds."""

    def __init__(self, camera: RayCamera, options: RenderOptions = RenderOptions()):
        self.camera = camera
        self.options = options

    @classmethod
    def new(cls, camera: RayCamera) -> "PathTracer":
        return cls(camera)

    def with_camera(self, camera: RayCamera) -> "PathTracer":
        return PathTracer(camera, self.options)

    def with_options(self, options: RenderOptions) -> "PathTracer":
        return PathTracer(self.camera, options)

    # core ---------------------------------------------------------------
    def _render(self, world: RayScene, camera: RayCamera, lights=None, env=None, stats: bool = False):
        if not isinstance(world, RayScene):
            raise TypeError("the GPU path renders RayScene worlds")
        ctx = world.context
        sh = world._compile()
        lh = lights._compile(ctx) if lights is not None else None
        eh = env._compile(ctx) if env is not None else None
        w, h = camera.image_width(), camera.image_height()
        out = np.zeros((h, w, 3), np.float32)
        st = abi.GtRenderStats()
        abi.check(ctx.lib.gt_render_hdr(ctx.handle, sh, C.byref(camera.raw), lh, eh, out.ctypes.data_as(abi.F),
                                        C.byref(st) if stats else None))
        img = HdrImage(w, h, out, ctx)
        if stats:
            return img, {k: getattr(st, k) for k, _ in abi.GtRenderStats._fields_}
        return img

    def _render_aovs(self, world: RayScene, camera: RayCamera, lights=None, env=None) -> DenoisingAovs:
        ctx = world.context
        sh = world._compile()
        lh = lights._compile(ctx) if lights is not None else None
        eh = env._compile(ctx) if env is not None else None
        w, h = camera.image_width(), camera.image_height()
        b = np.zeros((h, w, 3), np.float32)
        a = np.zeros_like(b)
        n = np.zeros_like(b)
        abi.check(ctx.lib.gt_render_aovs(ctx.handle, sh, C.byref(camera.raw), lh, eh, b.ctypes.data_as(abi.F),
                                         a.ctypes.data_as(abi.F), n.ctypes.data_as(abi.F), None))
        return DenoisingAovs(b, a, n, ctx)

    @staticmethod
    def _nee(camera: RayCamera) -> RayCamera:  # light_connection_camera / environment_importance_camera
        return camera.with_direct_lighting_mode(abi.GT_LIGHTING_NEE)

    def _progressive(self, world: RayScene, camera: RayCamera, lights, env, progress, tile_size) -> np.ndarray:
        if not isinstance(world, RayScene):
            raise TypeError("the GPU path renders RayScene worlds")
        ctx = world.context
        sh = world._compile()
        lh = lights._compile(ctx) if lights is not None else None
        eh = env._compile(ctx) if env is not None else None
        w, h = camera.image_width(), camera.image_height()
        canvas = np.zeros((h, w, 3), np.uint8)
        failure = []

        def cb(_user, up):
            try:
                if progress is not None:
                    progress(ProgressiveRenderUpdate(up.contents, canvas))
                return 0
            except BaseException as e:  # noqa: BLE001 (handed back to the caller below)
                failure.append(e)
                return 1

        fn = abi.PROGRESS_FN(cb)
        ts = tile_size if tile_size is not None else (self.options.tile_size or 0)
        rc = ctx.lib.gt_render_progressive(ctx.handle, sh, C.byref(camera.raw), lh, eh, int(ts), fn, None,
                                           canvas.ctypes.data_as(C.POINTER(C.c_uint8)))
        if failure:
            raise failure[0]
        abi.check(rc)
        return canvas

    def render_progressive(self, world, progress, tile_size: Optional[int] = None) -> np.ndarray:
        """render_progressive (renderer.rs:130-144): the display canvas, with
        progress(update) called on this thread after each tile; an exception
        raised by progress stops the render and propagates (the reference's
        Err)."""
        return self._progressive(world, self.camera, None, None, progress, tile_size)

    def render_with_lights_progressive(self, world, lights, progress, tile_size: Optional[int] = None) -> np.ndarray:
        return self._progressive(world, self.camera, lights, None, progress, tile_size)

    def render_with_light_connections_progressive(self, world, lights, progress,
                                                  tile_size: Optional[int] = None) -> np.ndarray:
        return self._progressive(world, self._nee(self.camera), lights, None, progress, tile_size)

    # renderer.rs entry points ----------------------------------------------
    def render(self, world) -> np.ndarray:
        return self.render_hdr_image(world).to_canvas()

    def render_tone_mapped(self, world, tone_map: ToneMap) -> np.ndarray:
        return self.render_hdr_image(world).to_canvas_tone_mapped(tone_map)

    def render_hdr_image(self, world) -> HdrImage:
        return self._render(world, self.camera)

    def render_denoising_aovs(self, world) -> DenoisingAovs:
        return self._render_aovs(world, self.camera)

    def render_with_lights(self, world, lights) -> np.ndarray:
        return self.render_with_lights_hdr_image(world, lights).to_canvas()

    def render_with_lights_hdr_image(self, world, lights) -> HdrImage:
        return self._render(world, self.camera, lights)

    def render_with_lights_denoising_aovs(self, world, lights) -> DenoisingAovs:
        return self._render_aovs(world, self.camera, lights)

    def render_with_light_connections(self, world, lights) -> np.ndarray:
        return self._render(world, self._nee(self.camera), lights).to_canvas()

    def render_with_environment(self, world, environment) -> np.ndarray:
        return self.render_with_environment_hdr_image(world, environment).to_canvas()

    def render_with_environment_hdr_image(self, world, environment) -> HdrImage:
        return self._render(world, self._nee(self.camera), None, environment)

    def render_with_lights_and_environment(self, world, lights, environment) -> np.ndarray:
        return self.render_with_lights_and_environment_hdr_image(world, lights, environment).to_canvas()

    def render_with_lights_and_environment_hdr_image(self, world, lights, environment) -> HdrImage:
        return self._render(world, self._nee(self.camera), lights, environment)

    def render_ray_scene(self, scene: RayScene) -> np.ndarray:
        return self.render_ray_scene_hdr_image(scene).to_canvas()

    def render_ray_scene_hdr_image(self, scene: RayScene) -> HdrImage:  # renderer.rs:193-201
        lights = scene.emissive_targets()
        if lights.is_empty():
            return self.render_hdr_image(scene)
        return self.render_with_lights_hdr_image(scene, lights)

    def render_ray_scene_with_light_connections(self, scene: RayScene) -> np.ndarray:
        lights = scene.emissive_targets()
        if lights.is_empty():
            return self.render(scene)
        return self.render_with_light_connections(scene, lights)

    def render_ray_scene_denoising_aovs(self, scene: RayScene) -> DenoisingAovs:
        lights = scene.emissive_targets()
        if lights.is_empty():
            return self.render_denoising_aovs(scene)
        return self.render_with_lights_denoising_aovs(scene, lights)

    def render_scene(self, scene: "SurfaceScene") -> np.ndarray:
        return self.render_scene_hdr_image(scene).to_canvas()

    def render_scene_hdr_image(self, scene: "SurfaceScene") -> HdrImage:  # renderer.rs:283-286
        return self.render_hdr_image(scene.to_ray_scene())

    def render_scene_denoising_aovs(self, scene: "SurfaceScene") -> DenoisingAovs:
        return self.render_denoising_aovs(scene.to_ray_scene())

    def render_scene_with_lights(self, scene: "SurfaceScene", lights) -> np.ndarray:
        return self.render_with_lights(scene.to_ray_scene(), lights)


# ------------------------------------------------------------ SurfaceScene ---

@dataclass
class SurfaceMaterial:
    """The part of SurfaceMaterial (src/graphics/material.rs) that
    SurfaceRayMaterialMode::Lambertian reads: the base colour."""
    base_color: tuple = (0.5, 0.5, 0.5)


class SurfaceScene:
    """SurfaceScene (src/graphics/scene.rs:80-145): triangle meshes with
    materials; to_ray_scene compiles one Lambertian material per mesh and one
    triangle per polygon (scene.rs:1517-1547)."""

    def __init__(self):
        self.meshes = []   # (triangles (n, 9), SurfaceMaterial)

    def add_mesh(self, triangles, material: SurfaceMaterial = SurfaceMaterial()) -> None:
        self.meshes.append((np.asarray(triangles, np.float64).reshape(-1, 9), material))

    def __len__(self) -> int:
        return len(self.meshes)

    def to_ray_scene(self, leaf_size: int = 4) -> RayScene:
        rs = RayScene()
        for tris, mat in self.meshes:
            mid = rs.add_material(RayMaterial.lambertian(mat.base_color))
            rs.add_triangles(mid, tris)
        rs.build_bvh_with_options(leaf_size)
        return rs


def render_batch(items, stats: bool = False):
    """Renders many (scene, camera, lights, environment) frames in one wavefront
    (gt_render_batch): every launch mixes the frames' paths. `items` is a list
    of (RayScene, RayCamera, SamplingTargetList | None, EnvironmentLight | None);
    all scenes must share one Context. Returns a list of HdrImage (and the
    summed render statistics when `stats`). Each image equals the frame's own
    single render."""
    items = list(items)
    if not items:
        return ([], {}) if stats else []
    ctx = items[0][0].context
    n = len(items)
    scenes = (C.c_void_p * n)()
    lights = (C.c_void_p * n)()
    envs = (C.c_void_p * n)()
    cams = (abi.GtCamera * n)()
    outs = (abi.F * n)()
    images = []
    for i, (world, camera, L, env) in enumerate(items):
        if not isinstance(world, RayScene) or world.context is not ctx:
            raise abi.InvalidArgument(abi.GT_EINVAL, "batched scenes must be RayScenes of one context")
        scenes[i] = world._compile()
        lights[i] = L._compile(ctx) if L is not None else None
        envs[i] = env._compile(ctx) if env is not None else None
        cams[i] = camera.raw
        w, h = camera.image_width(), camera.image_height()
        img = np.zeros((h, w, 3), np.float32)
        images.append(img)
        outs[i] = img.ctypes.data_as(abi.F)
    st = abi.GtRenderStats()
    abi.check(ctx.lib.gt_render_batch(ctx.handle, n, scenes, cams, lights, envs, outs,
                                      C.byref(st) if stats else None))
    result = [HdrImage(im.shape[1], im.shape[0], im, ctx) for im in images]
    if stats:
        return result, {k: getattr(st, k) for k, _ in abi.GtRenderStats._fields_}
    return result


def render_pixels(world: RayScene, camera: RayCamera, pixels, lights: Optional["SamplingTargetList"] = None,
                  environment: Optional["EnvironmentLight"] = None, double: bool = True,
                  stats: bool = False):
    """gt_render_pixels_device: only the listed pixels (indices y * W + x) of
    the camera's image, each equal to the same pixel of the full render.
    Returns an (n, 3) array in list order (float64 by default), and the
    render statistics when `stats`. The camera is used as given: a caller
    that wants the PathTracer's forced next-event estimation for lights or
    an environment sets it on the camera."""
    ctx = world.context
    sh = world._compile()
    lh = lights._compile(ctx) if lights is not None else None
    eh = environment._compile(ctx) if environment is not None else None
    px = np.ascontiguousarray(np.asarray(pixels, dtype=np.uint32).reshape(-1))
    w, h = camera.image_width(), camera.image_height()
    item = 8 if double else 4
    buf = C.c_void_p()
    abi.check(ctx.lib.gt_device_malloc(ctx.handle, w * h * 3 * item, C.byref(buf)))
    try:
        st = abi.GtRenderStats()
        abi.check(ctx.lib.gt_render_pixels_device(ctx.handle, sh, C.byref(camera.raw), lh, eh,
                                                  px.ctypes.data_as(abi.U32), len(px), buf, 1 if double else 0,
                                                  C.byref(st) if stats else None))
        img = np.zeros((h * w, 3), np.float64 if double else np.float32)
        abi.check(ctx.lib.gt_memcpy_to_host(ctx.handle, img.ctypes.data_as(C.c_void_p), buf, img.nbytes))
    finally:
        ctx.lib.gt_device_free(ctx.handle, buf)
    out = img[px.astype(np.int64)]
    if stats:
        return out, {k: getattr(st, k) for k, _ in abi.GtRenderStats._fields_}
    return out
