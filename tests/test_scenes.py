"""Host scene generators: the mesh builders of PolygonMatrix
(polygon_matrix.rs:510-541,680-704,795-864) and the batched workload."""
import math

import numpy as np

from gartus_amd import scenes


def _torus_loop(center, r_small, r_big, steps):
    """add_torus written as the reference's nested loops."""
    def ring(i):
        phi = i / steps * 2.0 * math.pi
        out = []
        for j in range(steps):
            th = j / steps * 2.0 * math.pi
            out.append((math.cos(phi) * (r_small * math.cos(th) + r_big) + center[0],
                        r_small * math.sin(th) + center[1],
                        -math.sin(phi) * (r_small * math.cos(th) + r_big) + center[2]))
        return out
    tris = []
    for i in range(steps):
        cur, nxt = ring(i), ring((i + 1) % steps)
        for j in range(steps):
            p0, p1, p2, p3 = cur[j], cur[(j + 1) % steps], nxt[(j + 1) % steps], nxt[j]
            tris.append([*p0, *p3, *p2])
            tris.append([*p0, *p2, *p1])
    return np.array(tris)


def test_torus_matches_loop_form():
    np.testing.assert_allclose(scenes.torus_triangles((1.0, 2.0, -3.0), 0.5, 2.0, 7),
                               _torus_loop((1.0, 2.0, -3.0), 0.5, 2.0, 7), rtol=0, atol=1e-12)


def test_crystal_field_matches_single_crystals():
    rng = np.random.default_rng(4)
    n, sides = 5, 7
    c = rng.uniform(-3, 3, (n, 3))
    r, h, yaw = rng.uniform(0.1, 1, n), rng.uniform(0.5, 2, n), rng.uniform(0, 3, n)
    field = scenes.crystal_field_triangles(c, sides, r, h, yaw)
    single = np.concatenate([scenes.crystal_triangles(tuple(c[i]), sides, r[i], h[i], yaw[i]) for i in range(n)])
    np.testing.assert_allclose(field, single, rtol=0, atol=1e-12)
    assert len(single) == 2 * sides * n


def test_crystal_is_closed_double_pyramid():
    tris = scenes.crystal_triangles((0.0, 1.0, 0.0), 6, 1.0, 2.0, 0.0).reshape(-1, 3, 3)
    assert tris.shape == (12, 3, 3)
    # every edge is shared by exactly two triangles (closed surface)
    edges = {}
    for t in np.round(tris, 9):
        for a, b in ((0, 1), (1, 2), (2, 0)):
            key = tuple(sorted((tuple(t[a]), tuple(t[b]))))
            edges[key] = edges.get(key, 0) + 1
    assert set(edges.values()) == {2}


def test_batch_scene_is_deterministic_and_sized():
    a, la, ca = scenes.batch_scene(1005)
    b, lb, cb = scenes.batch_scene(1005)
    assert len(a) == len(b) and ca == cb and len(la) == len(lb)
    for (k1, d1, m1), (k2, d2, m2) in zip(a.chunks, b.chunks):
        assert k1 == k2 and np.array_equal(d1, d2) and np.array_equal(m1, m2)
    assert 1 <= len(la) <= 8
    assert (ca["width"], round(ca["width"] / ca["aspect"])) in scenes.BATCH_RESOLUTIONS
    assert ca["spp"] == 8 and ca["depth"] == 8 and ca["nee"]
