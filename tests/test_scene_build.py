"""The product's host SAH builder (gartus_amd/csrc/scene_build.cpp) builds
the oracle's tree exactly: same node order and bounds, same leaf ranges,
same index permutation (FlatBvh::build, bvh.rs:233-535). The GPU's
closest-hit tie-breaking depends on this. CPU only: the builder is
compiled with g++ through tests/bvh_shim.cpp."""
import ctypes as C
import subprocess
from pathlib import Path

import numpy as np
import pytest

import cpu_ref
from gartus_amd import scenes
from gartus_amd.scene_data import LAMBERTIAN, SPHERE, TRIANGLE, Material, SceneData

ROOT = Path(__file__).resolve().parent.parent


@pytest.fixture(scope="module")
def shim(tmp_path_factory):
    out = tmp_path_factory.mktemp("shim") / "libbvhshim.so"
    subprocess.run(["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-o", str(out),
                    str(ROOT / "tests" / "bvh_shim.cpp"), str(ROOT / "gartus_amd" / "csrc" / "scene_build.cpp")],
                   check=True)
    lib = C.CDLL(str(out))
    lib.shim_build.restype = C.c_uint64
    lib.shim_build.argtypes = [C.c_size_t, C.POINTER(C.c_int), C.POINTER(C.c_double), C.c_uint32]
    return lib


def product_tree(lib, data: SceneData, leaf: int):
    kinds, q = [], []
    for kind, row, _ in data.primitives():
        kinds.append(kind)
        r = np.zeros(9)
        r[: len(row)] = row
        q.append(r)
    kinds = np.ascontiguousarray(kinds, np.int32)
    q = np.ascontiguousarray(np.array(q), np.float64)
    n = lib.shim_build(len(kinds), kinds.ctypes.data_as(C.POINTER(C.c_int)), q.ctypes.data_as(C.POINTER(C.c_double)),
                       leaf)
    b = np.zeros((n, 6))
    lf = np.zeros(n, np.uint8)
    a = np.zeros(n, np.uint64)
    c = np.zeros(n, np.uint64)
    idx = np.zeros(len(kinds), np.uint64)
    depth, n2, n4 = C.c_uint64(), C.c_uint64(), C.c_uint64()
    P = lambda x, t: x.ctypes.data_as(C.POINTER(t))  # noqa: E731
    lib.shim_dump(P(b, C.c_double), P(lf, C.c_uint8), P(a, C.c_uint64), P(c, C.c_uint64), P(idx, C.c_uint64),
                  C.byref(depth), C.byref(n2), C.byref(n4))
    return b, lf.astype(bool), a, c, idx, depth.value, n2.value, n4.value


def degenerate_scene():
    """All centroids equal on every axis: the midpoint fallback (bvh.rs:508-522)."""
    s = SceneData()
    m = s.add_material(Material(LAMBERTIAN, (0.5, 0.5, 0.5)))
    for i in range(37):
        s.add(SPHERE, [1.0, 2.0, 3.0, 0.1 + 0.01 * i], m)
    return s


def flat_triangles():
    s = SceneData()
    m = s.add_material(Material(LAMBERTIAN, (0.5, 0.5, 0.5)))
    g = np.stack(np.meshgrid(np.arange(9.0), np.arange(9.0)), -1).reshape(-1, 2)
    tris = [[x, 0, z, x + 1, 0, z, x, 0, z + 1] for x, z in g]
    s.add(TRIANGLE, tris, m)
    return s


CASES = {
    "randomized_3": lambda: scenes.randomized_ray_scene(3),
    "final_scene": scenes.final_scene,
    "cornell": lambda: scenes.cornell()[0],
    "bench": lambda: scenes.bench_scene()[0],
    "degenerate": degenerate_scene,
    "flat": flat_triangles,
    "terrain_small": lambda: scenes.terrain(grid=24, n_spheres=40)[0],
}


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("leaf", [1, 4, 8])
def test_product_builder_equals_oracle_tree(shim, name, leaf):
    data = CASES[name]()
    ob, olf, oa, oc, oidx = cpu_ref.Scene(data, leaf_size=leaf).bvh()
    pb, plf, pa, pc, pidx, depth, n2, n4 = product_tree(shim, data, leaf)
    np.testing.assert_array_equal(plf, olf)
    np.testing.assert_array_equal(pa, oa)
    np.testing.assert_array_equal(pc, oc)
    np.testing.assert_array_equal(pidx, oidx)
    np.testing.assert_array_equal(pb, ob)
    assert depth == cpu_ref.Scene(data, leaf_size=leaf).depth()
    assert n2 == int((~olf).sum())   # one GPU record per internal node
    assert n4 >= 1 or n2 == 0
