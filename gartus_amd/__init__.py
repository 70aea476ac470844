"""gartus_amd — MI355X (gfx950) path tracer for gartus's PathTracer over a RayScene.

The GPU library is ``libgartus_hip.so`` (built by ``gartus_amd.build``),
reached through the C ABI of ``include/gartus_hip.h``. ``gartus_amd.api``
mirrors the reference's PathTracer / RayCamera / RayScene API on top of it.
"""
from .api import (Context, DenoisingAovs, DirectLightingMode, EnvironmentLight, HdrImage, PathTracer,
                  PixelSampleMode, ProgressiveRenderUpdate, RayCamera, RayMaterial, RayScene, RenderOptions,
                  RenderProgress, RenderTile, SamplingTargetList, SurfaceMaterial, SurfaceScene, ToneMap,
                  ToneMappingOperator, WeightedSamplingTargetList, default_context)
from .scene_data import SceneData

__all__ = [
    "Context", "DenoisingAovs", "DirectLightingMode", "EnvironmentLight", "HdrImage", "PathTracer",
    "PixelSampleMode", "ProgressiveRenderUpdate", "RayCamera", "RayMaterial", "RayScene", "RenderOptions",
    "RenderProgress", "RenderTile", "SamplingTargetList",
    "SceneData", "SurfaceMaterial", "SurfaceScene", "ToneMap", "ToneMappingOperator",
    "WeightedSamplingTargetList", "default_context",
]
