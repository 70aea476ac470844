"""The MDL runtime's `shading raytrace` call of the path tracer, over this
library (src/mdl/runtime.rs:797-837, raytrace_canvas).

The reference's MDL runtime renders a scene through the path tracer like
this: a camera of the canvas's size with the output's samples per pixel,
depth and an sRGB background; the scene camera's eye and aim, its focal
length giving the vertical field of view; the surface scene as a RayScene;
every MDL light as a DiffuseLight sphere of radius `light_radius` with 12x
its colour, and the same spheres as the NEE target list; then render, or
render_with_lights when there are lights, to an 8-bit canvas. This module is
that caller, with the same conversions, on the GPU path tracer; the MDL
parser and the rest of its runtime stay with the reference.
"""
from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Optional, Sequence

import numpy as np

from . import api
from .scene_data import SceneData

F64_EPSILON = 2.220446049250313e-16
# runtime.rs:44-46
DEFAULT_SAMPLES_PER_PIXEL = 16
DEFAULT_MAX_DEPTH = 8
DEFAULT_LIGHT_RADIUS = 10.0
LIGHT_EMIT_SCALE = 12.0  # runtime.rs:823


@dataclass
class MdlCamera:
    """The MDL scene camera: eye, aim point and focal length in pixels."""
    eye: Sequence[float]
    aim: Sequence[float]
    focal: float


@dataclass
class MdlLight:
    """An MDL point light: position and colour in 0..255 units."""
    position: Sequence[float]
    color: Sequence[float]


def rgb_from_vec3(color) -> tuple:
    """runtime.rs:1019-1025: each component rounded and clamped to a byte."""
    return tuple(int(min(max(round_half_away(c), 0.0), 255.0)) for c in color)


def round_half_away(x: float) -> float:
    """f64::round: halves away from zero (Python's round() goes to even)."""
    return math.copysign(math.floor(abs(x) + 0.5), x)


def linear_from_rgb_linear_units(rgb) -> tuple:
    """LinearRgb::from_rgb_linear_units (colors.rs:68-74): byte / 255."""
    return tuple(c / 255.0 for c in rgb)


def linear_from_rgb_srgb(rgb) -> tuple:
    """LinearRgb::from_rgb_srgb (colors.rs:78-84): the gamma-2 decode (byte / 255)^2."""
    return tuple((c / 255.0) * (c / 255.0) for c in rgb)


def lookat_or_default_forward(eye, aim) -> tuple:
    """runtime.rs:1035-1041: the aim point, or one unit down -z when it is the eye."""
    d = [a - e for a, e in zip(aim, eye)]
    if d[0] * d[0] + d[1] * d[1] + d[2] * d[2] > F64_EPSILON:
        return tuple(aim)
    return (eye[0], eye[1], eye[2] - 1.0)


def vertical_fov_from_focal(height: int, focal: float) -> Optional[float]:
    """runtime.rs:806-814: 2 atan(height / 2 / focal) in degrees, used only
    when the focal length is finite and positive and the angle in (0, 180)."""
    if not (math.isfinite(focal) and focal > F64_EPSILON):
        return None
    fov = math.degrees(2.0 * math.atan(height * 0.5 / focal))
    return fov if math.isfinite(fov) and 0.0 < fov < 180.0 else None


def raytrace_scene(scene: SceneData, width: int, height: int, lights: Sequence[MdlLight] = (),
                   camera: Optional[MdlCamera] = None, background_rgb=(0, 0, 0),
                   samples_per_pixel: int = DEFAULT_SAMPLES_PER_PIXEL, max_depth: int = DEFAULT_MAX_DEPTH,
                   light_radius: float = DEFAULT_LIGHT_RADIUS, context: Optional[api.Context] = None):
    """The tracer, RayScene (its BVH not built yet) and target list that
    raytrace_canvas builds; the scene is copied, and the light spheres are
    added to the copy."""
    aspect = float(width) / float(max(height, 1))
    cam = (api.RayCamera.new(max(width, 1), aspect).with_samples_per_pixel(samples_per_pixel)
           .with_max_depth(max_depth).with_background(linear_from_rgb_srgb(background_rgb)))
    if camera is not None:
        eye = tuple(float(v) for v in camera.eye)
        cam = cam.with_look_at(eye, lookat_or_default_forward(eye, tuple(float(v) for v in camera.aim)))
        fov = vertical_fov_from_focal(height, float(camera.focal))
        if fov is not None:
            cam = cam.with_vertical_fov(fov)
    data = SceneData(materials=list(scene.materials), chunks=list(scene.chunks))
    world = api.RayScene(data, context)
    targets = api.SamplingTargetList()
    for light in lights:
        emit = tuple(LIGHT_EMIT_SCALE * c for c in linear_from_rgb_linear_units(rgb_from_vec3(light.color)))
        m = world.add_material(api.RayMaterial.diffuse_light(emit))
        world.add_sphere(light.position, light_radius, m)
        targets.add_sphere(light.position, light_radius)
    return api.PathTracer.new(cam), world, targets


def raytrace_canvas(scene: SceneData, width: int, height: int, lights: Sequence[MdlLight] = (),
                    camera: Optional[MdlCamera] = None, background_rgb=(0, 0, 0),
                    samples_per_pixel: int = DEFAULT_SAMPLES_PER_PIXEL, max_depth: int = DEFAULT_MAX_DEPTH,
                    light_radius: float = DEFAULT_LIGHT_RADIUS,
                    context: Optional[api.Context] = None) -> np.ndarray:
    """runtime.rs:797-837: the 8-bit canvas of the scene (render, or
    render_with_lights when there are lights)."""
    tracer, world, targets = raytrace_scene(scene, width, height, lights, camera, background_rgb,
                                            samples_per_pixel, max_depth, light_radius, context)
    world.build_bvh()
    if len(targets) == 0:
        return tracer.render(world)
    return tracer.render_with_lights(world, targets)
