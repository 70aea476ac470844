"""Plain-data scene description shared by the GPU library and its checkers.

A ``SceneData`` records materials and primitives in insertion order, the
order that fixes primitive ids in a RayScene (scene.rs:1148-1222). It holds
no device state; ``gartus_amd.api.RayScene`` compiles it for the GPU.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

# material kinds (layout.h GM_*), primitive kinds (GK_*); a mesh triangle
# carries its three vertex normals after its positions (18 values)
LAMBERTIAN, DIFFUSE_LIGHT, GGX, METAL, DIELECTRIC, ISOTROPIC, HENYEY_GREENSTEIN = range(7)
SPHERE, MOVING_SPHERE, TRIANGLE, QUAD, MESH_TRIANGLE = range(5)
PRIM_STRIDE = {SPHERE: 4, MOVING_SPHERE: 7, TRIANGLE: 9, QUAD: 9, MESH_TRIANGLE: 18}


@dataclass
class Material:
    kind: int
    color: tuple = (0.0, 0.0, 0.0)
    param: float = 0.0   # ggx roughness, metal fuzz, dielectric ior, henyey-greenstein anisotropy

    @property
    def is_emissive(self) -> bool:  # RayMaterial::is_emissive
        return self.kind == DIFFUSE_LIGHT


@dataclass
class SceneData:
    materials: list = field(default_factory=list)
    chunks: list = field(default_factory=list)   # (kind, float64 array (n, stride), uint32 material ids)

    def add_material(self, m: Material) -> int:
        self.materials.append(m)
        return len(self.materials) - 1

    def add(self, kind: int, data, material_ids) -> None:
        data = np.ascontiguousarray(np.asarray(data, dtype=np.float64).reshape(-1, PRIM_STRIDE[kind]))
        mids = np.asarray(material_ids, dtype=np.uint32)
        if mids.ndim == 0:
            mids = np.full(len(data), int(mids), dtype=np.uint32)
        mids = np.ascontiguousarray(mids.reshape(-1))
        if len(mids) != len(data):
            raise ValueError("one material id per primitive")
        if len(mids) and int(mids.max()) >= len(self.materials):
            raise ValueError("ray scene material id out of bounds")
        self.chunks.append((kind, data, mids))

    def __len__(self) -> int:
        return sum(len(d) for _, d, _ in self.chunks)

    def primitives(self):
        """(kind, params, material) per primitive id, for small scenes."""
        for kind, data, mids in self.chunks:
            for row, m in zip(data, mids):
                yield kind, row, int(m)

    def emissive_targets(self):
        """RayScene::emissive_targets (scene.rs:1403-1423): (kind, params, 1.0)."""
        out = []
        for kind, data, mids in self.chunks:
            for row, m in zip(data, mids):
                if self.materials[int(m)].is_emissive:  # a mesh triangle is a triangle target
                    out.append((TRIANGLE, np.array(row[:9]), 1.0) if kind == MESH_TRIANGLE
                               else (kind, np.array(row), 1.0))
        return out

    def as_float32_exact(self) -> "SceneData":
        """A copy whose coordinates are exactly representable in FP32, so a
        float64 checker and the FP32 GPU see the same geometry."""
        s = SceneData(list(self.materials), [])
        for kind, data, mids in self.chunks:
            s.chunks.append((kind, data.astype(np.float32).astype(np.float64), mids.copy()))
        return s
