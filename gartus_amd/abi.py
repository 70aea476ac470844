"""ctypes mirror of ``include/gartus_hip.h`` and the library loader.

Only structure layouts and function prototypes live here; loading the
library never touches the GPU. ``load()`` raises when the built library is
missing: there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

LIB_PATH = Path(__file__).resolve().parent / "libgartus_hip.so"

GT_OK, GT_EINVAL, GT_ENOMEM, GT_EDEVICE, GT_ESTATE, GT_EUNSUPPORTED = range(6)
GT_PIXEL_RANDOM, GT_PIXEL_STRATIFIED, GT_PIXEL_STRATIFIED_GRID = range(3)
GT_LIGHTING_CONTINUATION, GT_LIGHTING_NEE = range(2)
GT_BACKGROUND_CONSTANT, GT_BACKGROUND_VERTICAL_GRADIENT = range(2)
GT_FLAG_GGX_REJECT_BELOW_HORIZON = 1
GT_RNG_COUNTER, GT_RNG_PIXEL_STREAM = range(2)


class GtCamera(C.Structure):
    _fields_ = [
        ("image_width", C.c_uint32),
        ("samples_per_pixel", C.c_uint32),
        ("aspect_ratio", C.c_double),
        ("pixel_sample_mode", C.c_uint32),
        ("stratified_grid_width", C.c_uint32),
        ("max_depth", C.c_uint32),
        ("russian_roulette_min_depth", C.c_int32),
        ("rng_seed", C.c_uint64),
        ("vertical_fov", C.c_double),
        ("lookfrom", C.c_double * 3),
        ("lookat", C.c_double * 3),
        ("view_up", C.c_double * 3),
        ("defocus_angle", C.c_double),
        ("focus_distance", C.c_double),
        ("shutter_start", C.c_double),
        ("shutter_end", C.c_double),
        ("background_kind", C.c_uint32),
        ("direct_lighting_mode", C.c_uint32),
        ("background", C.c_double * 9),
        ("light_pdf_weight", C.c_double),
        ("rng_mode", C.c_uint32),
        ("flags", C.c_uint32),
        ("adaptive_min_samples", C.c_uint32),
        ("adaptive_max_samples", C.c_uint32),
        ("adaptive_error_threshold", C.c_double),
    ]


class GtTuning(C.Structure):
    _fields_ = [
        ("block_size", C.c_uint32),
        ("blocks_per_cu", C.c_uint32),
        ("max_paths", C.c_uint64),
        ("bvh_width", C.c_uint32),
        ("use_graph", C.c_uint32),
        ("count_traversal", C.c_uint32),
        ("stack_lds", C.c_uint32),
    ]


class GtSceneInfo(C.Structure):
    _fields_ = [
        ("primitive_count", C.c_uint64),
        ("material_count", C.c_uint64),
        ("bvh_node_count", C.c_uint64),
        ("bvh_depth", C.c_uint64),
        ("device_bytes", C.c_uint64),
        ("bounds_min", C.c_double * 3),
        ("bounds_max", C.c_double * 3),
        ("build_ms", C.c_double),
        ("built_on_device", C.c_uint32),
        ("reserved", C.c_uint32),
    ]


class GtRay(C.Structure):
    _fields_ = [
        ("origin", C.c_float * 3),
        ("time", C.c_float),
        ("direction", C.c_float * 3),
        ("t_max", C.c_float),
    ]


class GtHit(C.Structure):
    _fields_ = [
        ("t", C.c_float),
        ("u", C.c_float),
        ("v", C.c_float),
        ("prim_id", C.c_int32),
        ("normal", C.c_float * 3),
        ("front_face", C.c_uint32),
    ]


class GtTraversalStats(C.Structure):
    _fields_ = [
        ("rays", C.c_uint64),
        ("node_bounds_tests", C.c_uint64),
        ("node_bounds_hits", C.c_uint64),
        ("leaf_visits", C.c_uint64),
        ("primitive_candidates", C.c_uint64),
        ("leaf_hit_results", C.c_uint64),
        ("max_stack_depth", C.c_uint64),
    ]


class GtRenderStats(C.Structure):
    _fields_ = [
        ("camera_paths", C.c_uint64),
        ("extension_rays", C.c_uint64),
        ("shadow_rays", C.c_uint64),
        ("accepted_samples", C.c_uint64),
        ("device_ms", C.c_double),
        ("extend_ms", C.c_double),
        ("extend_launches", C.c_uint64),
        ("shade_ms", C.c_double),
        ("connect_ms", C.c_double),
        ("ext_node_fetches", C.c_uint64),
        ("ext_prim_tests", C.c_uint64),
        ("shadow_node_fetches", C.c_uint64),
        ("shadow_prim_tests", C.c_uint64),
    ]


P = C.c_void_p
PP = C.POINTER(C.c_void_p)
D = C.POINTER(C.c_double)
F = C.POINTER(C.c_float)
U32 = C.POINTER(C.c_uint32)
I = C.c_int

# name -> (restype, argtypes); every exported symbol of include/gartus_hip.h
GT_ECANCELLED = 6


class GtRenderTile(C.Structure):
    _fields_ = [("x", C.c_uint32), ("y", C.c_uint32), ("width", C.c_uint32), ("height", C.c_uint32)]


class GtRenderProgress(C.Structure):
    _fields_ = [("tile", GtRenderTile), ("completed_tiles", C.c_uint64), ("total_tiles", C.c_uint64),
                ("image_width", C.c_uint32), ("image_height", C.c_uint32), ("pixels", C.POINTER(C.c_uint8))]


PROGRESS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(GtRenderProgress))

GT_TONE_NONE, GT_TONE_REINHARD, GT_TONE_ACES = 0, 1, 2
GT_ENCODE_CANVAS, GT_ENCODE_RAW, GT_ENCODE_RGBE, GT_ENCODE_PFM = 0, 1, 2, 3


class GtToneMap(C.Structure):
    _fields_ = [("exposure", C.c_double), ("gamma", C.c_double), ("op", C.c_uint32), ("reserved", C.c_uint32)]


GT_DENOISE_LOSS_REL_L2, GT_DENOISE_LOSS_L1_TONEMAP = 0, 1


class GtDenoiserConfig(C.Structure):
    _fields_ = [
        ("levels", C.c_uint32),
        ("channels", C.c_uint32 * 6),
        ("seed", C.c_uint64),
        ("loss", C.c_uint32),
        ("beauty_halves", C.c_uint32),
    ]


PROTOTYPES = {
    "gt_camera_default": (None, [C.POINTER(GtCamera)]),
    "gt_camera_image_height": (C.c_uint32, [C.POINTER(GtCamera)]),
    "gt_camera_effective_spp": (C.c_uint32, [C.POINTER(GtCamera)]),
    "gt_context_create": (I, [I, PP]),
    "gt_context_destroy": (I, [P]),
    "gt_context_synchronize": (I, [P]),
    "gt_context_stream": (P, [P]),
    "gt_device_malloc": (I, [P, C.c_size_t, PP]),
    "gt_device_free": (I, [P, P]),
    "gt_memcpy_to_host": (I, [P, P, P, C.c_size_t]),
    "gt_memcpy_to_device": (I, [P, P, P, C.c_size_t]),
    "gt_last_error": (C.c_char_p, []),
    "gt_version": (C.c_char_p, []),
    "gt_context_get_tuning": (I, [P, C.POINTER(GtTuning)]),
    "gt_context_set_tuning": (I, [P, C.POINTER(GtTuning)]),
    "gt_scene_create": (I, [P, PP]),
    "gt_scene_destroy": (I, [P]),
    "gt_scene_add_lambertian": (I, [P, C.c_double, C.c_double, C.c_double, U32]),
    "gt_scene_add_diffuse_light": (I, [P, C.c_double, C.c_double, C.c_double, U32]),
    "gt_scene_add_ggx": (I, [P, C.c_double, C.c_double, C.c_double, C.c_double, U32]),
    "gt_scene_add_metal": (I, [P, C.c_double, C.c_double, C.c_double, C.c_double, U32]),
    "gt_scene_add_dielectric": (I, [P, C.c_double, U32]),
    "gt_scene_add_isotropic": (I, [P, C.c_double, C.c_double, C.c_double, U32]),
    "gt_scene_add_henyey_greenstein": (I, [P, C.c_double, C.c_double, C.c_double, C.c_double, U32]),
    "gt_scene_add_spheres": (I, [P, D, C.c_size_t, U32]),
    "gt_scene_add_moving_spheres": (I, [P, D, C.c_size_t, U32]),
    "gt_scene_add_triangles": (I, [P, D, C.c_size_t, U32]),
    "gt_scene_add_quads": (I, [P, D, C.c_size_t, U32]),
    "gt_scene_add_mesh_triangles": (I, [P, D, D, C.c_size_t, U32]),
    "gt_scene_build": (I, [P, C.c_uint32]),
    "gt_scene_build_device": (I, [P, C.c_uint32]),
    "gt_scene_get_info": (I, [P, C.POINTER(GtSceneInfo)]),
    "gt_intersect": (I, [P, P, C.POINTER(GtRay), C.c_size_t, C.c_float, C.POINTER(GtHit),
                         C.POINTER(GtTraversalStats)]),
    "gt_intersect_device_noalloc": (I, [P, P, P, C.c_size_t, C.c_float, P, P]),
    "gt_lights_create": (I, [P, PP]),
    "gt_lights_destroy": (I, [P]),
    "gt_lights_add_sphere": (I, [P, D, C.c_double, C.c_double]),
    "gt_lights_add_moving_sphere": (I, [P, D, D, C.c_double, C.c_double]),
    "gt_lights_add_quad": (I, [P, D, D, D, C.c_double]),
    "gt_lights_add_triangle": (I, [P, D, D, D, C.c_double]),
    "gt_lights_len": (C.c_size_t, [P]),
    "gt_scene_emissive_targets": (I, [P, PP]),
    "gt_env_create_constant": (I, [P, C.c_double, C.c_double, C.c_double, PP]),
    "gt_env_create_rgb32f": (I, [P, F, C.c_uint32, C.c_uint32, PP]),
    "gt_env_destroy": (I, [P]),
    "gt_render_hdr": (I, [P, P, C.POINTER(GtCamera), P, P, F, C.POINTER(GtRenderStats)]),
    "gt_render_hdr_device": (I, [P, P, C.POINTER(GtCamera), P, P, P, I, C.POINTER(GtRenderStats)]),
    "gt_render_aovs": (I, [P, P, C.POINTER(GtCamera), P, P, F, F, F, C.POINTER(GtRenderStats)]),
    "gt_render_pixels_device": (I, [P, P, C.POINTER(GtCamera), P, P, U32, C.c_size_t, P, I,
                                    C.POINTER(GtRenderStats)]),
    "gt_render_aovs_device": (I, [P, P, C.POINTER(GtCamera), P, P, F, F, F, C.POINTER(GtRenderStats)]),
    "gt_render_batch": (I, [P, C.c_size_t, PP, C.POINTER(GtCamera), PP, PP, C.POINTER(F),
                            C.POINTER(GtRenderStats)]),
    "gt_render_batch_device": (I, [P, C.c_size_t, PP, C.POINTER(GtCamera), PP, PP, PP, I,
                                   C.POINTER(GtRenderStats)]),
    "gt_autotune": (I, [P, C.c_size_t, PP, C.POINTER(GtCamera), PP, PP, C.c_uint32, C.POINTER(GtTuning),
                        C.POINTER(C.c_double)]),
    # learned denoiser (gartus_denoise.h)
    "gt_denoiser_default_config": (None, [C.POINTER(GtDenoiserConfig)]),
    "gt_denoiser_create": (I, [P, C.POINTER(GtDenoiserConfig), PP]),
    "gt_denoiser_destroy": (I, [P]),
    "gt_denoiser_param_count": (I, [P, C.POINTER(C.c_uint64)]),
    "gt_denoiser_get_params": (I, [P, F]),
    "gt_denoiser_set_params": (I, [P, F]),
    "gt_denoiser_infer": (I, [P, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P, P]),
    "gt_denoiser_loss_and_gradients": (I, [P, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P, P, F, F]),
    "gt_denoiser_train_step": (I, [P, C.c_uint32, C.c_uint32, C.c_uint32, P, P, P, P, C.c_float, F]),
    "gt_render_progressive": (I, [P, P, C.POINTER(GtCamera), P, P, C.c_uint32, PROGRESS_FN, C.c_void_p,
                                  C.POINTER(C.c_uint8)]),
    # output boundary
    "gt_encoded_size": (C.c_uint64, [C.c_int, C.c_uint32, C.c_uint32]),
    "gt_encoded_header": (I, [C.c_int, C.c_uint32, C.c_uint32, C.c_char_p, C.c_size_t, C.POINTER(C.c_size_t)]),
    "gt_encode_image": (I, [P, P, C.c_int, C.c_int, C.c_uint32, C.c_uint32, C.c_int, C.POINTER(GtToneMap),
                            P, C.c_int]),
}

_lib = None


def load() -> C.CDLL:
    """Loads libgartus_hip.so (built by ``gartus_amd.build``). No fallback.

    The HIP runtime is per process. A process that also uses PyTorch on the
    GPU must import torch first: torch ships its own libamdhip64, and the
    library then binds to that copy (same soname) instead of loading
    /opt/rocm's, which could not open the device a second time."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise RuntimeError(
                f"{LIB_PATH} is missing: build it with `python -m gartus_amd.build` "
                "(the HIP path has no CPU fallback)")
        lib = C.CDLL(str(LIB_PATH))
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib


class GartusError(RuntimeError):
    def __init__(self, code: int, message: str):
        super().__init__(f"gartus error {code}: {message}")
        self.code = code


class InvalidArgument(GartusError, ValueError):
    pass


def check(rc: int) -> None:
    if rc != GT_OK:
        msg = load().gt_last_error().decode(errors="replace")
        if rc == GT_EINVAL:
            raise InvalidArgument(rc, msg)
        raise GartusError(rc, msg)


def default_camera() -> GtCamera:
    """RayCamera::new(100, 1.0) defaults (camera.rs:1034-1064), set in Python
    so that code without the library can still build camera records."""
    c = GtCamera()
    c.image_width = 100
    c.aspect_ratio = 1.0
    c.samples_per_pixel = 1
    c.pixel_sample_mode = GT_PIXEL_RANDOM
    c.stratified_grid_width = 1
    c.max_depth = 10
    c.russian_roulette_min_depth = 5
    c.rng_seed = 1
    c.vertical_fov = 90.0
    c.lookfrom[:] = (0.0, 0.0, 0.0)
    c.lookat[:] = (0.0, 0.0, -1.0)
    c.view_up[:] = (0.0, 1.0, 0.0)
    c.defocus_angle = 0.0
    c.focus_distance = 1.0
    c.shutter_start = 0.0
    c.shutter_end = 1.0
    c.background_kind = GT_BACKGROUND_CONSTANT
    c.background[0:3] = (0.70, 0.80, 1.00)
    c.direct_lighting_mode = GT_LIGHTING_CONTINUATION
    c.light_pdf_weight = 0.5
    c.rng_mode = GT_RNG_COUNTER
    return c
