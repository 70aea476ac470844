"""The CPU oracle against the reference's own known-answer tests and against
independent restatements (published SplitMix64 vectors, numpy uint64
arithmetic, closed forms). CPU only.

Each test names the reference test it restates (file:line); where a value
comes from mathematics rather than from the reference's tests, the test says
which identity pins it.
"""
import ctypes as C
import math

import numpy as np
import pytest

import cpu_ref
from gartus_amd import abi, scenes
from gartus_amd.scene_data import (DIELECTRIC, DIFFUSE_LIGHT, GGX, LAMBERTIAN, METAL, QUAD, SPHERE, TRIANGLE,
                                   Material, SceneData)

L = cpu_ref.lib()
D = C.POINTER(C.c_double)
PI = math.pi


def arr(*v):
    return np.ascontiguousarray(np.array(v, dtype=np.float64))


def dp(a):
    return a.ctypes.data_as(D)


def close(a, b, tol=1e-10):
    assert abs(a - b) < tol, (a, b)


# ------------------------------------------------------------------ RNG ---

def test_splitmix64_published_vectors():
    """SplitMix64 (Steele, Lea, Flood 2014; Vigna's splitmix64.c): seed
    1234567 yields the sequence used by rand_xoshiro's reference test."""
    out = np.zeros(5, np.uint64)
    L.cr_rng_u64(1234567, 5, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    assert [int(x) for x in out] == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                                     4593380528125082431, 16408922859458223821]


def _np_mix(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def test_random_double_and_pixel_seed_match_numpy_restatement():
    """random.rs:23-29,56-59 and camera.rs:1720-1727 recomputed with numpy
    uint64 wrap-around arithmetic."""
    seed = 0x5EED2026
    out = np.zeros(64, np.uint64)
    L.cr_rng_u64(seed, 64, out.ctypes.data_as(C.POINTER(C.c_uint64)))
    with np.errstate(over="ignore"):
        states = np.uint64(seed) + np.arange(1, 65, dtype=np.uint64) * np.uint64(0x9E3779B97F4A7C15)
    np.testing.assert_array_equal(out, _np_mix(states))
    dbl = np.zeros(64)
    L.cr_rng_double(seed, 64, dbl.ctypes.data_as(D))
    np.testing.assert_array_equal(dbl, (out >> np.uint64(11)).astype(np.float64) * 2.0 ** -53)
    for x, y in [(0, 0), (5, 9), (1919, 1079), (123456, 7)]:
        with np.errstate(over="ignore"):
            z = (np.uint64(1) ^ (np.uint64(x) * np.uint64(0x9E3779B97F4A7C15))
                 ^ (np.uint64(y) * np.uint64(0xD1B54A32D192ED03)))
        assert int(L.cr_pixel_seed(1, x, y)) == int(_np_mix(z))


def test_pixel_seed_lands_on_published_splitmix64_outputs():
    """pixel_seed (camera.rs:1720-1727) is the SplitMix64 output function
    applied to seed ^ x K1 ^ y K2, with K1, K2 read off that source. SplitMix64
    seeded with 1234567 first outputs mix(1234567 + gamma) (its published
    vector), so a seed chosen to make seed ^ x K1 ^ y K2 equal to
    1234567 + gamma must give that published output, and likewise for the
    second state 1234567 + 2 gamma. This pins the finalizer's shifts and
    multipliers and the way x and y enter, without restating them."""
    M = (1 << 64) - 1
    gamma, k1, k2 = 0x9E3779B97F4A7C15, 0x9E3779B97F4A7C15, 0xD1B54A32D192ED03
    published = [6457827717110365317, 3203168211198807973]
    for j, want in enumerate(published):
        state = (1234567 + (j + 1) * gamma) & M
        for x, y in [(0, 0), (1, 0), (0, 1), (1919, 1079), (77, 4)]:
            seed = state ^ ((x * k1) & M) ^ ((y * k2) & M)
            assert int(L.cr_pixel_seed(seed, x, y)) == want, (j, x, y)


def test_counter_streams_are_distinct_and_deterministic():
    key = int(L.cr_pixel_seed(1, 3, 4))
    seeds = {int(L.cr_stream_seed(key, s, d)) for s in range(64) for d in range(11)}
    assert len(seeds) == 64 * 11
    assert int(L.cr_stream_seed(key, 5, 2)) == int(L.cr_stream_seed(key, 5, 2))


@pytest.mark.parametrize("which,seed", [(0, 29), (1, 31), (2, 17)])
def test_rng_vector_samplers(which, seed):
    """random.rs tests :196-253: unit length / hemisphere / unit disk."""
    v = np.zeros((200, 3))
    L.cr_rng_vectors(seed, which, 200, v.ctypes.data_as(D))
    n = np.linalg.norm(v, axis=1)
    if which == 2:
        assert np.all(n < 1.0) and np.all(v[:, 2] == 0.0)
    else:
        assert np.all(np.abs(n - 1.0) < 1e-12)
    if which == 1:
        assert np.all((v[:, 2] >= 0) & (v[:, 2] <= 1))


def test_sphere_samples_average_near_zero_and_cosine_mean():
    """sampling.rs:433-483: sphere samples mean ~0 (+-0.025); cosine samples
    mean cosine 2/3 (+-0.025)."""
    v = np.zeros((4096, 3))
    L.cr_rng_vectors(101, 0, 4096, v.ctypes.data_as(D))
    assert np.all(np.abs(v.mean(0)) < 0.025)
    c = np.zeros((4096, 3))
    L.cr_rng_vectors(103, 1, 4096, c.ctypes.data_as(D))
    assert abs(c[:, 2].mean() - 2.0 / 3.0) < 0.025


# ------------------------------------------------------------ geometry ---

def one_prim_scene(kind, q):
    s = SceneData()
    m = s.add_material(Material(LAMBERTIAN, (0.5, 0.5, 0.5)))
    s.add(kind, q, m)
    return cpu_ref.Scene(s)


def test_sphere_hit_roots():
    """raytracing.rs:336-352: direct hit at t 0.5, far root 1.5 with t_min 0.6."""
    sc = one_prim_scene(SPHERE, [0, 0, -1, 0.5])
    h = sc.intersect([[0, 0, 0]], [[0, 0, -1]], t_min=-np.inf)
    close(h["t"][0], 0.5)
    h = sc.intersect([[0, 0, 0]], [[0, 0, -1]], t_min=0.6)
    close(h["t"][0], 1.5)
    assert sc.intersect([[0, 0, 0]], [[0, 1, -1]], t_min=-np.inf)["prim"][0] == -1
    assert sc.intersect([[0, 0, 0]], [[0, 0, 0]], t_min=-np.inf)["prim"][0] == -1


def test_triangle_barycentrics_and_bounds():
    """geometry.rs:645-680: t 1, (u, v) = (0.25, 0.25); bounds padded 1e-10."""
    sc = one_prim_scene(TRIANGLE, [0, 0, -1, 1, 0, -1, 0, 1, -1])
    h = sc.intersect([[0.25, 0.25, 0]], [[0, 0, -1]], t_min=0.0)
    close(h["t"][0], 1.0)
    close(h["u"][0], 0.25)
    close(h["v"][0], 0.25)
    np.testing.assert_allclose(h["normal"][0], [0, 0, 1])
    b, leaf, *_ = sc.bvh()
    close(b[0, 2], -1.0 - 0.5e-10)
    close(b[0, 5], -1.0 + 0.5e-10)


def test_quad_planar_coordinates_bounds_and_rejections():
    """geometry.rs:664-723: (u, v) = (0.5, 0.75); bounds padded 1e-4;
    outside, parallel and degenerate rays rejected."""
    sc = one_prim_scene(QUAD, [-2, -2, -1, 4, 0, 0, 0, 4, 0])
    h = sc.intersect([[0, 1, 0]], [[0, 0, -1]], t_min=0.0)
    close(h["t"][0], 1.0)
    close(h["u"][0], 0.5)
    close(h["v"][0], 0.75)
    flat = one_prim_scene(QUAD, [-2, -2, 0, 4, 0, 0, 0, 4, 0])
    b, *_ = flat.bvh()
    close(b[0, 2], -0.5e-4)
    close(b[0, 5], 0.5e-4)
    q = one_prim_scene(QUAD, [-1, -1, -1, 2, 0, 0, 0, 2, 0])
    assert q.intersect([[2, 0, 0]], [[0, 0, -1]], t_min=0.0)["prim"][0] == -1
    assert q.intersect([[0, 0, 0]], [[1, 0, 0]], t_min=0.0)["prim"][0] == -1
    deg = one_prim_scene(QUAD, [0, 0, 0, 1, 0, 0, 2, 0, 0])
    assert deg.intersect([[0, 0, 1]], [[0, 0, -1]], t_min=0.0)["prim"][0] == -1


def test_front_face_orientation():
    """object.rs:127-149: front iff d . n_out < 0; normal flipped to face the ray."""
    sc = one_prim_scene(SPHERE, [0, 0, 0, 1])
    out = sc.intersect([[0, 0, 5]], [[0, 0, -1]])
    inn = sc.intersect([[0, 0, 0]], [[0, 0, -1]])
    assert out["front"][0] and not inn["front"][0]
    np.testing.assert_allclose(out["normal"][0], [0, 0, 1])
    np.testing.assert_allclose(inn["normal"][0], [0, 0, 1])


# ---------------------------------------------------------------- BVH ---

def hit_bounds(box, o, d, tmin=0.0, tmax=10.0):
    e = np.zeros(1)
    ok = L.cr_hit_bounds(dp(arr(*box)), dp(arr(*o, 0.0, *d)), tmin, tmax, dp(e))
    return float(e[0]) if ok else None


UNIT = (-1, -1, -1, 1, 1, 1)


def test_slab_axis_parallel_and_boundary_cases():
    """bvh.rs:711-770."""
    for d in [(1, 0, 0), (0, 1, 0), (0, 0, 1), (-1, 0, 0), (0, -1, 0), (0, 0, -1)]:
        assert hit_bounds(UNIT, (0, 0, 0), d) == 0.0
    for o, d in [((0, 2, 0), (1, 0, 0)), ((0, 0, 2), (0, 1, 0)), ((2, 0, 0), (0, 0, 1))]:
        assert hit_bounds(UNIT, o, d) is None
    assert hit_bounds(UNIT, (0.25, -0.25, 0.5), (0, 0, 1)) == 0.0
    assert hit_bounds(UNIT, (1, -2, 0), (0, 1, 0)) == 1.0
    assert hit_bounds(UNIT, (2, 2, 0), (-1, -1, 1)) is None
    assert hit_bounds(UNIT, (0, 0, -2), (0, 0, 1)) == 1.0
    assert hit_bounds(UNIT, (2, 0, -2), (0, 0, 1)) is None


def test_traversal_stats_and_far_child_pruning():
    """bvh.rs:627-709: a root with two leaves at z in [1, 2] and [10, 11];
    a ray along +z visits only the near leaf; stats (1, 3, 3, 1, 1, 1)."""
    s = SceneData()
    m = s.add_material(Material(LAMBERTIAN, (0.5, 0.5, 0.5)))
    s.add(QUAD, [-1, -1, 1.5, 2, 0, 0, 0, 2, 0], m)
    s.add(QUAD, [-1, -1, 10.5, 2, 0, 0, 0, 2, 0], m)
    sc = cpu_ref.Scene(s, leaf_size=1)
    assert sc.node_count() == 3
    h = sc.intersect([[0, 0, 0]], [[0, 0, 1]], t_min=0.0, t_max=20.0)
    assert h["prim"][0] == 0
    close(h["t"][0], 1.5)
    st = h["stats"]
    assert (st["rays"], st["node_bounds_tests"], st["node_bounds_hits"], st["leaf_visits"],
            st["primitive_candidates"], st["leaf_hit_results"]) == (1, 3, 3, 1, 1, 1)
    assert st["max_stack_depth"] >= 1


def test_leaf_size_controls_node_count():
    """scene.rs:1800-1822 and bvh.rs:598-604: leaf size 8 over 8 primitives is
    one node; leaf size 0 clamps to 1."""
    s = SceneData()
    m = s.add_material(Material(LAMBERTIAN, (0.5, 0.5, 0.5)))
    for i in range(8):
        s.add(SPHERE, [i * 2.0, 0, -3, 0.5], m)
    assert cpu_ref.Scene(s, leaf_size=8).node_count() == 1
    assert cpu_ref.Scene(s, leaf_size=0).node_count() == cpu_ref.Scene(s, leaf_size=1).node_count() == 15


@pytest.mark.parametrize("seed", range(16))
def test_bvh_matches_bruteforce_randomized_scene(seed):
    """raytracing.rs:1340-1370: 16 seeds x 64 rays, t/point/normal/front to 1e-10."""
    data = scenes.randomized_ray_scene(seed)
    sc = cpu_ref.Scene(data)
    rng = scenes.SampleRng(seed + 1000)
    o, d = [], []
    for _ in range(64):
        o.append((rng.random_range(-0.5, 0.5), rng.random_range(-0.5, 0.5), 0.0))
        d.append((rng.random_range(-0.6, 0.6), rng.random_range(-0.5, 0.5), -1.0))
    a = sc.intersect(o, d)
    b = sc.intersect(o, d, use_bvh=False)
    np.testing.assert_array_equal(a["prim"], b["prim"])
    hit = a["prim"] >= 0
    np.testing.assert_allclose(a["t"][hit], b["t"][hit], atol=1e-10)
    np.testing.assert_allclose(a["normal"][hit], b["normal"][hit], atol=1e-10)
    np.testing.assert_array_equal(a["front"], b["front"])


def test_bvh_is_a_valid_sah_partition():
    """Invariants of FlatBvh::build (bvh.rs:233-288): pre-order numbering,
    leaves partition the index permutation, leaf size bound, every node's
    box encloses its children."""
    data = scenes.final_scene()
    sc = cpu_ref.Scene(data, leaf_size=4)
    b, leaf, a, c, idx = sc.bvh()
    assert sorted(int(i) for i in idx) == list(range(len(data)))
    covered = np.zeros(len(data), int)
    for i in range(len(b)):
        if leaf[i]:
            assert 1 <= c[i] <= 4
            covered[int(a[i]):int(a[i] + c[i])] += 1
        else:
            assert a[i] == i + 1 and c[i] > a[i]
            for ch in (int(a[i]), int(c[i])):
                assert np.all(b[ch, :3] >= b[i, :3]) and np.all(b[ch, 3:] <= b[i, 3:])
    assert np.all(covered == 1)


# ----------------------------------------------------------- sampling ---

def test_cosine_pdf_values():
    """raytracing.rs:369-375, sampling.rs:454-465: 1/pi along the normal,
    non-unit directions accepted, 0 below the horizon and for zero vectors."""
    n = arr(0, 1, 0)
    close(L.cr_cosine_pdf_value(dp(n), dp(arr(0, 2, 0))), 1 / PI)
    close(L.cr_cosine_pdf_value(dp(n), dp(arr(0, 0, 0))), 0.0)
    close(L.cr_cosine_pdf_value(dp(n), dp(arr(0, -2, 0))), 0.0)
    close(L.cr_cosine_pdf_value(dp(n), dp(arr(1, 0, 0))), 0.0)


def test_ggx_known_values():
    """sampling.rs:518-538: normal-incidence reflection pdf 1/(4 pi) at
    roughness 1; D ordering; Smith G = 1 at normal incidence; Schlick."""
    n = arr(0, 0, 1)
    v = np.zeros(1)
    assert L.cr_ggx_pdf(dp(n), dp(n), 1.0, dp(n), 0, 0, dp(v), dp(np.zeros(3))) == 0
    close(v[0], 1 / (4 * PI))
    assert L.cr_ggx_pdf(dp(n), dp(n), 1.0, dp(arr(0, 0, -1)), 0, 0, dp(v), dp(np.zeros(3))) == 0
    close(v[0], 0.0)
    assert L.cr_ggx_D(1.0, 0.2) > L.cr_ggx_D(1.0, 0.8)
    close(L.cr_smith_g1(1.0, 0.5) * L.cr_smith_g1(1.0, 0.5), 1.0)
    close(L.cr_schlick(1.0, 0.04), 0.04)
    close(L.cr_schlick(0.0, 0.04), 1.0)


def test_ggx_samples_above_surface_and_pdf_normalised():
    """sampling.rs:541-556 (samples above the surface, finite positive pdf),
    and normalisation: generate() returns `outgoing` whenever the reflected
    direction falls below the horizon (sampling.rs:302-306), so the pdf
    integrates over the upper hemisphere to 1 - P(generate == outgoing).
    Both sides are estimated by Monte Carlo; the identity pins D, the
    1/(4 o.h) Jacobian and the roughness-to-alpha map."""
    n = arr(0, 0, 1)
    o = arr(0.3, 0.0, 0.95)
    o = o / np.linalg.norm(o)
    v = np.zeros(1)
    for r in (0.45, 0.8, 1.0):
        g = np.zeros((64, 3))
        assert L.cr_ggx_pdf(dp(n), dp(o), r, dp(n), 109, 64, dp(v), dp(g)) == 0
        assert np.all(np.abs(np.linalg.norm(g, axis=1) - 1) < 1e-12) and np.all(g[:, 2] > 0)
        many = np.zeros((100000, 3))
        L.cr_ggx_pdf(dp(n), dp(o), r, dp(n), 7, 100000, dp(v), dp(many))
        p_outgoing = np.all(many == o, axis=1).mean()
        u = np.zeros((30000, 3))
        L.cr_rng_vectors(5, 0, 30000, u.ctypes.data_as(D))
        vals = np.empty(len(u))
        for i in range(len(u)):
            L.cr_ggx_pdf(dp(n), dp(o), r, dp(np.ascontiguousarray(u[i])), 0, 0, dp(v), dp(np.zeros(3)))
            vals[i] = v[0]
        est = vals.mean() * 4 * PI
        assert abs(est - (1.0 - p_outgoing)) < 0.03, (r, est, p_outgoing)
    # roughness 1 is alpha 1: D = 1/pi, half the reflections fall below the horizon
    many = np.zeros((100000, 3))
    L.cr_ggx_pdf(dp(n), dp(n), 1.0, dp(n), 3, 100000, dp(v), dp(many))
    assert abs(np.all(many == n, axis=1).mean() - 0.5) < 0.01


def test_ggx_brdf_weight_closed_form():
    """The GGX 'scattering_pdf' weight D G F / (4 n.o) (material.rs:1686-1722)
    at normal incidence/exitance: D = 1/(pi a^2), G = 1, F/f0 = 1, so the
    weight is 1/(4 pi a^2) with a = r^2."""
    n = arr(0, 0, 1)
    for r in (0.3, 0.7, 1.0):
        a = max(r * r, 1e-4)
        close(L.cr_ggx_weight(dp(n), dp(n), dp(n), r, 0.5), 1 / (4 * PI * a * a), 1e-9)


def test_power_heuristic_and_russian_roulette():
    """camera.rs:2486-2523 and its tests :5047-5103."""
    close(L.cr_power_heuristic(1.0, 1.0), 0.5)
    close(L.cr_power_heuristic(0.0, 1.0), 0.0)
    close(L.cr_power_heuristic(1.0, 0.0), 1.0)
    att = arr(0, 0, 0)
    assert L.cr_russian_roulette(3, 5, dp(att), 41) == 1
    assert L.cr_russian_roulette(5, -1, dp(att), 41) == 1
    assert L.cr_russian_roulette(5, 5, dp(att), 43) == 0
    for seed in range(100):
        a = arr(2.0, 1.0, 0.5)
        if L.cr_russian_roulette(5, 5, dp(a), seed):
            np.testing.assert_allclose(a, np.array([2.0, 1.0, 0.5]) / 0.95, atol=1e-10)
            break
    else:
        pytest.fail("no seed survived")


def test_light_targets_pdfs():
    """raytracing.rs:929-964,1638-1701: quad pdf 0.25; sphere cone pdf
    1/(2 pi (1 - sqrt(0.75))); weighted pick ratio 0.75; emissive targets."""
    quad = cpu_ref.Lights([(QUAD, [-1, -1, -1, 2, 0, 0, 0, 2, 0], 1.0)])
    close(quad.pdf_value([0, 0, 0], 0.0, [0, 0, -1]), 0.25)
    close(quad.pdf_value([0, 0, 0], 0.0, [0, 1, 0]), 0.0)
    assert np.all(quad.random_directions([0, 0, 0], 0.0, 31, 16)[:, 2] < 0)
    sph = cpu_ref.Lights([(SPHERE, [0, 0, -1, 0.5], 1.0)])
    close(sph.pdf_value([0, 0, 0], 0.0, [0, 0, -1]), 1 / (2 * PI * (1 - math.sqrt(0.75))))
    close(sph.pdf_value([0, 0, 0], 0.0, [0, 1, 0]), 0.0)
    # weighted selection (scene.rs:584-600): two quads, weights 1 and 3
    two = cpu_ref.Lights([(QUAD, [10, -1, -1, 0, 2, 0, 0, 0, 2], 1.0), (QUAD, [-1, 10, -1, 2, 0, 0, 0, 0, 2], 3.0)])
    d = two.random_directions([0, 0, 0], 0.0, 149, 10000)
    assert abs((d[:, 1] > 5).mean() - 0.75) < 0.025
    s = SceneData()
    dif = s.add_material(Material(LAMBERTIAN, (0.2, 0.4, 0.6)))
    light = s.add_material(Material(DIFFUSE_LIGHT, (4.0, 4.0, 4.0)))
    s.add(SPHERE, [0, 0, -1, 0.5], dif)
    s.add(QUAD, [-1, -1, -1, 2, 0, 0, 0, 2, 0], light)
    t = cpu_ref.Scene(s).emissive_targets()
    assert len(t) == 1
    close(t.pdf_value([0, 0, 0], 0.0, [0, 0, -1]), 0.25)


def test_triangle_light_pdf_closed_form():
    """Area-to-solid-angle conversion d^2 / (cos A) (object.rs:1101-1113) for
    a triangle of area 2 at distance 1 on the axis: 1/2."""
    tri = cpu_ref.Lights([(TRIANGLE, [-1, -1, -1, 1, -1, -1, -1, 1, -1], 1.0)])
    close(tri.pdf_value([0, 0, 0], 0.0, [-0.5, -0.5, -1]), (0.25 + 0.25 + 1) / (1 / math.sqrt(1.5) * 2.0) * 1.0, 1e-9)


def test_environment_known_answers():
    """environment.rs:208-269: constant light (radiance, 1/(4 pi) pdf); 8-bit
    map samples bright texels more; a 1x2 white-over-black map samples the
    upper hemisphere with pdf 1/(2 pi) and mean cos 0.5 (+-0.03)."""
    e = cpu_ref.Env(color=(0.25, 0.5, 0.75))
    rad, pdf = e.eval([[1, 0, 0]])
    np.testing.assert_allclose(rad[0], [0.25, 0.5, 0.75])
    close(pdf[0], 1 / (4 * PI), 1e-12)
    hdr = cpu_ref.Env(color=(4.0, 2.0, 1.5))
    np.testing.assert_allclose(hdr.eval([[0, 1, 0]])[0][0], [4.0, 2.0, 1.5])
    two = cpu_ref.Env(rgb8=np.array([[[0, 0, 0], [255, 255, 255]]], np.uint8))
    _, p = two.eval([[-1, 0, 0], [1, 0, 0]])
    assert p[0] > p[1]
    _, ps = two.sample(5, 1)
    assert np.isfinite(ps[0]) and ps[0] > 0
    half = cpu_ref.Env(rgb8=np.array([[[255, 255, 255]], [[0, 0, 0]]], np.uint8))
    dirs, pdfs = half.sample(42, 4096)
    assert np.all(np.abs(pdfs - 1 / (2 * PI)) < 1e-12)
    assert np.all(dirs[:, 1] >= -1e-12)
    assert abs(dirs[:, 1].mean() - 0.5) < 0.03


def test_environment_hdr_pdf_matches_sampling_histogram():
    """For the FP32 HDR map, the solid-angle pdf returned with a sample must
    equal the probability of its texel divided by the texel's solid angle:
    the sampled-texel histogram over many draws matches pdf * solid angle."""
    rng = np.random.default_rng(3)
    rgb = rng.uniform(0, 1, (8, 16, 3)).astype(np.float32)
    rgb[2, 5] = 40.0
    e = cpu_ref.Env(rgb32f=rgb)
    dirs, pdfs = e.sample(11, 200000)
    theta = np.arccos(np.clip(dirs[:, 1], -1, 1))
    phi = np.mod(np.arctan2(dirs[:, 2], dirs[:, 0]), 2 * PI)
    x = np.minimum((phi / (2 * PI) * 16).astype(int), 15)
    y = np.minimum((theta / PI * 8).astype(int), 7)
    hist = np.bincount(y * 16 + x, minlength=128) / len(dirs)
    yy = np.arange(8)
    sa = (2 * PI / 16) * (np.cos(PI * yy / 8) - np.cos(PI * (yy + 1) / 8))
    expect = np.zeros(128)
    for k in range(128):
        sel = (y * 16 + x) == k
        if sel.any():
            expect[k] = pdfs[sel][0] * sa[k // 16]
    assert abs(expect.sum() - 1.0) < 1e-9
    assert np.max(np.abs(hist - expect)) < 0.005


# ------------------------------------------------------------ materials ---

def _material_scene():
    s = SceneData()
    ids = {
        "lam": s.add_material(Material(LAMBERTIAN, (0.5, 0.6, 0.7))),
        "ggx": s.add_material(Material(GGX, (0.9, 0.8, 0.7), 0.3)),
        "metal": s.add_material(Material(METAL, (0.8, 0.8, 0.8), 0.0)),
        "glass": s.add_material(Material(DIELECTRIC, (1, 1, 1), 1.5)),
        "light": s.add_material(Material(DIFFUSE_LIGHT, (4, 4, 4))),
    }
    s.add(SPHERE, [0, 0, -10, 1], ids["lam"])
    return cpu_ref.Scene(s), ids


def test_material_scatter_known_answers():
    """material.rs: Lambertian scatters with the albedo and spdf = cos/pi;
    a smooth metal reflects exactly; glass at normal incidence refracts
    straight through or reflects back; a light does not scatter."""
    sc, ids = _material_scene()
    ray = [0, 0, 1, 0.0, 0, 0, -1]
    hit = [0, 0, 0, 0, 0, 1, 1.0]
    k, d, att, pdf = sc.material_scatter(ids["lam"], ray, hit, 7)
    assert k == 2 and d[2] > 0
    np.testing.assert_allclose(att, [0.5, 0.6, 0.7])
    close(pdf, d[2] / np.linalg.norm(d) / PI)
    close(sc.scattering_pdf(ids["lam"], ray, hit, d), pdf)
    k, d, att, _ = sc.material_scatter(ids["metal"], ray, hit, 7)
    assert k == 1
    np.testing.assert_allclose(d, [0, 0, 1], atol=1e-12)
    outs = {tuple(np.round(sc.material_scatter(ids["glass"], ray, hit, s)[1], 9)) for s in range(64)}
    assert outs <= {(0.0, 0.0, 1.0), (0.0, 0.0, -1.0)} and len(outs) == 2
    assert sc.material_scatter(ids["light"], ray, hit, 7)[0] == 0
    k, d, att, pdf = sc.material_scatter(ids["ggx"], ray, hit, 7)
    assert k == 2 and d[2] > 0 and pdf > 0
    np.testing.assert_allclose(att, [0.9, 0.8, 0.7])


def test_dielectric_reflectance_fraction_matches_schlick():
    """Dielectric (material.rs:2321-2350): at normal incidence the
    reflected fraction is Schlick's r0 = ((1 - 1/1.5)/(1 + 1/1.5))^2 = 0.04."""
    sc, ids = _material_scene()
    ray = [0, 0, 1, 0.0, 0, 0, -1]
    hit = [0, 0, 0, 0, 0, 1, 1.0]
    refl = sum(sc.material_scatter(ids["glass"], ray, hit, s)[1][2] > 0 for s in range(20000))
    assert abs(refl / 20000 - 0.04) < 0.006


# ------------------------------------------------------------ camera ---

def cam(**kw):
    c = abi.default_camera()
    for k, v in kw.items():
        if k in ("lookfrom", "lookat", "view_up"):
            getattr(c, k)[:] = v
        else:
            setattr(c, k, v)
    return c


def frame(c):
    out = np.zeros(15)
    h = C.c_uint32()
    assert L.cr_camera_frame(C.byref(c), dp(out), C.byref(h)) == 0
    return out.reshape(5, 3), h.value


def test_camera_image_height_and_center_ray():
    """camera.rs:4879-4896: 400 x 16/9 gives 225 rows; the centre pixel
    looks down -z from the origin."""
    f, h = frame(cam(image_width=400, aspect_ratio=16 / 9))
    assert h == 225
    center, p00, du, dv, _ = f
    d = p00 + 200 * du + 112 * dv - center
    np.testing.assert_allclose(center, [0, 0, 0])
    assert d[2] < 0 and abs(d[0]) < 0.01 and abs(d[1]) < 0.01


def test_camera_look_at_and_vfov():
    """camera.rs:5266-5290: vfov controls the spread; look_at aims the
    centre ray at the target."""
    fw, _ = frame(cam(image_width=101, vertical_fov=90.0))
    fn, _ = frame(cam(image_width=101, vertical_fov=20.0))
    top = lambda f: (f[1] + 50 * f[2] - f[0]) / np.linalg.norm(f[1] + 50 * f[2] - f[0])  # noqa: E731
    assert abs(top(fw)[1]) > abs(top(fn)[1])
    f, _ = frame(cam(image_width=101, lookfrom=(-2, 2, 1), lookat=(0, 0, -1)))
    d = f[1] + 50 * f[2] + 50 * f[3] - f[0]
    e = np.array([2.0, -2.0, -2.0]) / np.linalg.norm([2.0, -2.0, -2.0])
    close(float(d @ e / np.linalg.norm(d)), 1.0)
    np.testing.assert_allclose(f[0], [-2, 2, 1])


def test_camera_validation_errors():
    """validate_view (camera.rs:1161-1201) rejects bad parameters."""
    for bad in [dict(vertical_fov=0.0), dict(vertical_fov=180.0), dict(lookat=(0, 0, 0)),
                dict(view_up=(0, 0, 0)), dict(view_up=(0, 0, 1)), dict(focus_distance=0.0),
                dict(defocus_angle=180.0), dict(shutter_start=1.0, shutter_end=0.0)]:
        out = np.zeros(15)
        h = C.c_uint32()
        assert L.cr_camera_frame(C.byref(cam(**bad)), dp(out), C.byref(h)) == abi.GT_EINVAL, bad


def test_stratified_offsets_stay_in_pixel():
    """camera.rs:5031-5045: stratified offsets stay inside the pixel square."""
    c = cam(image_width=8, pixel_sample_mode=abi.GT_PIXEL_STRATIFIED_GRID, stratified_grid_width=4,
            samples_per_pixel=16)
    f, _ = frame(c)
    center, p00, du, dv, _ = f
    ray = np.zeros(7)
    for s in range(16):
        assert L.cr_camera_ray(C.byref(c), 3, 2, s, dp(ray)) == 0
        d = ray[4:7]
        rel = d - (p00 + 3 * du + 2 * dv - center)
        ox = rel @ du / (du @ du)
        oy = rel @ dv / (dv @ dv)
        assert -0.5 <= ox < 0.5 and -0.5 <= oy < 0.5
        sx, sy = s % 4, s // 4
        assert sx / 4 - 0.5 <= ox < (sx + 1) / 4 - 0.5 + 1e-12
        assert sy / 4 - 0.5 <= oy < (sy + 1) / 4 - 0.5 + 1e-12


# ------------------------------------------------------------ display ---

def test_tone_map_known_answers():
    """raytracing.rs:553-557,1504-1522: gamma-2 clip (255, 255, 128) and
    Reinhard with gamma 1 (204, 128, 51) for radiance (4, 1, 0.25)."""
    out = np.zeros(3, np.uint8)
    L.cr_tone_map(dp(arr(4.0, 1.0, 0.25)), 1, 1.0, 0, 2.0, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    assert tuple(out) == (255, 255, 128)
    L.cr_tone_map(dp(arr(4.0, 1.0, 0.25)), 1, 1.0, 1, 1.0, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    assert tuple(out) == (204, 128, 51)


def test_rgbe_known_answer_and_exponent_cases():
    """display.rs:2188-2201: (1, 0.5, 0.25) encodes as [128, 64, 32, 129].
    Then what the arithmetic fixes: 0 and values at or below 1e-32 give four
    zero bytes; a power of two 2^k gives mantissa 128 and exponent k + 129;
    non-finite and negative channels count as 0; 2^-e is evaluated like
    f64::powi (the reciprocal of 2^e), so from e = 1024 on it is 0."""
    def enc(*px):
        a = np.ascontiguousarray(np.array(px, np.float64).reshape(-1))
        out = np.zeros(len(a) // 3 * 4, np.uint8)
        L.cr_encode_rgbe(dp(a), len(a) // 3, out.ctypes.data_as(C.POINTER(C.c_uint8)))
        return [tuple(int(v) for v in out[i:i + 4]) for i in range(0, len(out), 4)]
    assert enc(1.0, 0.5, 0.25) == [(128, 64, 32, 129)]
    assert enc(0.0, 0.0, 0.0, 1e-32, 0.0, 0.0, -5.0, np.nan, -np.inf) == [(0, 0, 0, 0)] * 3
    for k in (-100, -3, 0, 1, 7, 60, 126, 127):
        assert enc(2.0 ** k, 0.0, 0.0) == [(128, 0, 0, min(k + 129, 255))]
    assert enc(np.inf, 3.0, 0.0) == [(0, 192, 0, 130)]  # inf -> 0; 3 = 0.75 * 2^2
    assert enc(2.0 ** 1023, 0.0, 0.0)[0][:3] == (0, 0, 0)  # 2^-1024 evaluates to 0
    assert enc(2.0 ** 1022, 0.0, 0.0) == [(128, 0, 0, 255)]  # exponent byte clamps at 255


def test_raw_and_pfm_encodings():
    """Rgb::from_raw_linear_color (colors.rs:412-422): 255.999 * clamp(x, 0, 1)
    truncated, NaN -> 0. write_pfm (display.rs:365-387): rows bottom first,
    non-finite -> 0; the reference's PFM test (display.rs:2172-2186) fixes the
    12 bytes of a 1x1 image."""
    x = np.ascontiguousarray([0.0, 0.5, 1.0, 2.0, -1.0, np.nan, np.inf, 0.999, 1.0 / 255.999], np.float64)
    out = np.zeros(9, np.uint8)
    L.cr_encode_raw(dp(x), 3, out.ctypes.data_as(C.POINTER(C.c_uint8)))
    assert list(out) == [0, 127, 255, 255, 0, 0, 255, 255, 1]
    img = np.ascontiguousarray(np.arange(2 * 3 * 3, dtype=np.float64).reshape(3, 2, 3))
    img[0, 0, 0] = np.nan
    body = np.zeros(18, np.float32)
    L.cr_encode_pfm(dp(img), 2, 3, body.ctypes.data_as(C.POINTER(C.c_float)))
    expect = np.where(np.isfinite(img), img, 0.0)[::-1].reshape(-1).astype(np.float32)
    np.testing.assert_array_equal(body, expect)
    one = np.ascontiguousarray([2.0, 0.5, 0.25])
    b1 = np.zeros(3, np.float32)
    L.cr_encode_pfm(dp(one), 1, 1, b1.ctypes.data_as(C.POINTER(C.c_float)))
    assert b1.astype("<f4").tobytes() == np.array([2.0, 0.5, 0.25], "<f4").tobytes()


def test_encoded_headers_and_sizes():
    """The library's header text (pure host code, no device needed) is the
    reference's: display.rs:366-369 and :398-401, as its tests read them
    (display.rs:2183,2199)."""
    from gartus_amd import api
    assert api.encoded_header(abi.GT_ENCODE_PFM, 1, 1) == b"PF\n1 1\n-1.0\n"
    assert api.encoded_header(abi.GT_ENCODE_RGBE, 1, 1) == b"#?RADIANCE\nFORMAT=32-bit_rle_rgbe\n\n-Y 1 +X 1\n"
    assert api.encoded_header(abi.GT_ENCODE_RGBE, 1920, 1080).endswith(b"-Y 1080 +X 1920\n")
    assert api.encoded_header(abi.GT_ENCODE_CANVAS, 4, 4) == b""
    lib = abi.load()
    assert lib.gt_encoded_size(abi.GT_ENCODE_PFM, 7, 5) == 7 * 5 * 12
    assert lib.gt_encoded_size(abi.GT_ENCODE_RGBE, 7, 5) == 7 * 5 * 4
    assert lib.gt_encoded_size(abi.GT_ENCODE_RAW, 7, 5) == 7 * 5 * 3
    assert lib.gt_encoded_size(9, 7, 5) == 0
    with pytest.raises(abi.InvalidArgument):
        api.encoded_header(9, 1, 1)


def _alias_probabilities(thr, alias):
    """Texel probabilities implied by an alias table: slot j (probability 1/n)
    keeps j with probability thr[j] / 2^32 and yields alias[j] otherwise."""
    n = len(thr)
    keep = thr.astype(np.float64) / 2.0 ** 32
    own = np.arange(n)
    p = np.where(alias == own, 1.0, keep)
    np.add.at(p, alias[alias != own], 1.0 - keep[alias != own])
    return p / n


def test_environment_alias_table_reproduces_the_reference_distribution():
    """The counter-stream (GPU) texel choice samples exactly the reference's
    distribution w_i / W (rebuild_distribution, environment.rs:158-184): the
    alias table's implied probabilities equal it to within the 2^-32
    threshold quantisation."""
    rng = np.random.default_rng(7)
    h, w = 16, 32
    rgb = rng.uniform(0, 1, (h, w, 3)).astype(np.float32)
    rgb[3, 7] = 500.0
    rgb[12, 30] = 0.0
    e = cpu_ref.Env(rgb32f=rgb)
    thr, alias = e.alias_table()
    assert len(thr) == h * w
    theta = PI * (np.arange(h) + 0.5) / h
    lum = (0.2126 * rgb[..., 0].astype(np.float64) + 0.7152 * rgb[..., 1] + 0.0722 * rgb[..., 2])
    wt = np.maximum(lum, 0.0) * np.sin(theta)[:, None]
    expect = (wt / wt.sum()).ravel()
    got = _alias_probabilities(thr, alias)
    assert np.max(np.abs(got - expect)) < 1e-9
    assert got[12 * w + 30] == 0.0          # a black texel is never chosen
    assert abs(got.sum() - 1.0) < 1e-12


def test_environment_alias_sampling_histogram_matches_pdf():
    rng = np.random.default_rng(3)
    rgb = rng.uniform(0, 1, (8, 16, 3)).astype(np.float32)
    rgb[2, 5] = 40.0
    e = cpu_ref.Env(rgb32f=rgb)
    dirs, pdfs = e.sample(11, 200000, alias=True)
    theta = np.arccos(np.clip(dirs[:, 1], -1, 1))
    phi = np.mod(np.arctan2(dirs[:, 2], dirs[:, 0]), 2 * PI)
    x = np.minimum((phi / (2 * PI) * 16).astype(int), 15)
    y = np.minimum((theta / PI * 8).astype(int), 7)
    hist = np.bincount(y * 16 + x, minlength=128) / len(dirs)
    yy = np.arange(8)
    sa = (2 * PI / 16) * (np.cos(PI * yy / 8) - np.cos(PI * (yy + 1) / 8))
    thr, alias = e.alias_table()
    expect = _alias_probabilities(thr, alias)
    assert np.max(np.abs(hist - expect)) < 0.005
    # the returned pdf is the texel probability over the texel's solid angle
    k = y * 16 + x
    np.testing.assert_allclose(pdfs * sa[k // 16], expect[k], rtol=1e-6)


# ------------------------------------------------------------ adaptive sampling ---

def _adaptive(samples, min_samples, threshold):
    s = np.ascontiguousarray(np.asarray(samples, np.float64).reshape(-1, 3))
    mean = np.zeros(3)
    taken, accepted = C.c_uint32(), C.c_uint32()
    L.cr_adaptive_run(dp(s), len(s), min_samples, threshold, dp(mean), C.byref(taken), C.byref(accepted))
    return mean, taken.value, accepted.value


def test_adaptive_error_known_answers():
    """camera.rs:4967-4972: m2 (1, 4, 9) over 4 samples gives sqrt(0.75);
    one sample gives infinity."""
    m2 = np.array([1.0, 4.0, 9.0])
    assert L.cr_adaptive_error(dp(m2), 4) == pytest.approx(math.sqrt(0.75), rel=1e-15)
    assert math.isinf(L.cr_adaptive_error(dp(m2), 1))


def test_adaptive_stops_at_min_for_zero_variance_and_runs_to_max_for_noise():
    """camera.rs:4975-4991 (a constant pixel takes exactly min samples) and
    :5006-5019 (samples alternating 0, 1 never get below 0.001 and take max,
    mean 0.5)."""
    mean, taken, acc = _adaptive(np.full((16, 3), 0.25), 4, 1e-3)
    assert (taken, acc) == (4, 4) and np.all(mean == 0.25)
    alt = np.repeat((np.arange(16) % 2).astype(float)[:, None], 3, axis=1)
    mean, taken, acc = _adaptive(alt, 4, 1e-3)
    assert (taken, acc) == (16, 16)
    assert np.all(np.abs(mean - 0.5) < 1e-15)


def test_adaptive_skips_non_finite_samples_but_counts_them_against_max():
    seq = np.full((10, 3), 1.0)
    seq[[0, 2, 3]] = np.nan
    mean, taken, acc = _adaptive(seq, 4, 1e-3)
    assert acc == 4 and taken == 7 and np.all(mean == 1.0)
    mean, taken, acc = _adaptive(np.full((5, 3), np.inf), 1, 1e-3)
    assert (taken, acc) == (5, 0) and np.all(mean == 0.0)


def test_adaptive_matches_a_brute_force_restatement():
    """The stopping sample and the mean against numpy recomputed from
    scratch at every step (two-pass mean and unbiased variance): the
    Welford recurrence, the min-samples gate and the max-channel rule."""
    rng = np.random.default_rng(3)
    for trial in range(200):
        n = int(rng.integers(1, 200))
        seq = rng.gamma(rng.uniform(0.5, 4.0), 1.0, (n, 3)) * rng.uniform(0.1, 3.0, 3)
        mn = int(rng.integers(1, 12))
        thr = float(rng.uniform(0.02, 0.5))
        mean, taken, acc = _adaptive(seq, mn, thr)
        stop = n
        for k in range(1, n + 1):
            if k >= mn and k > 1:
                se = np.sqrt(np.max(seq[:k].var(axis=0, ddof=1) / k))
                if se < thr * (1 - 1e-9):
                    stop = k
                    break
                if se < thr * (1 + 1e-9):  # too close to call in floating point
                    stop = None
                    break
        if stop is None:
            continue
        assert taken == stop, (trial, taken, stop)
        np.testing.assert_allclose(mean, seq[:stop].mean(axis=0), rtol=1e-12)
