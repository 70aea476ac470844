"""Seeded scene generators for the benchmark workloads and the reference's
own test and bench scenes. Host-side data generation only.

Workloads (sizes as the benchmark configurations state them):
  cornell()        Cornell box, 256x256, 8x8 stratified = 64 spp, depth 8, NEE
  terrain()        Perlin height-map terrain (512x512 grid, 524,288
                   triangles) + 2,000 spheres, 1280x720, 32 spp, depth 6
  ggx_environment() 10x10 GGX spheres + 1,000 random spheres under a
                   2048x1024 FP32 HDR sky with a 5e4 sun, 1920x1080, 16 spp,
                   depth 10
  bench_scene()    benches/render_performance.rs:10-40 (129 spheres)
  final_scene()    scenes.rs:563-624 (final_scene_ray_scene)
"""
from __future__ import annotations

import math

import numpy as np

from .scene_data import (DIELECTRIC, DIFFUSE_LIGHT, GGX, LAMBERTIAN, METAL, QUAD, SPHERE, TRIANGLE,
                         Material, SceneData)

MASK = (1 << 64) - 1


class SampleRng:
    """SplitMix64 SampleRng (src/gmath/random.rs:12-59) for host generation."""

    def __init__(self, seed: int):
        self.state = seed & MASK

    def next_u64(self) -> int:
        self.state = (self.state + 0x9E3779B97F4A7C15) & MASK
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & MASK
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & MASK
        return z ^ (z >> 31)

    def random_double(self) -> float:
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def random_range(self, lo: float, hi: float) -> float:
        return lo + (hi - lo) * self.random_double()

    def random_vector(self):
        return (self.random_double(), self.random_double(), self.random_double())

    def random_vector_range(self, lo, hi):
        return (self.random_range(lo, hi), self.random_range(lo, hi), self.random_range(lo, hi))

    def random_unit_vector(self):  # random_unit_vector_spherical (random.rs:107-115)
        r1, r2 = self.random_double(), self.random_double()
        z = 1.0 - 2.0 * r2
        radius = math.sqrt(1.0 - z * z)
        phi = 2.0 * math.pi * r1
        return (math.cos(phi) * radius, math.sin(phi) * radius, z)


class Perlin:
    """Seeded Perlin noise (src/gmath/perlin.rs:15-125), vectorised."""

    def __init__(self, seed: int):
        rng = SampleRng(seed)
        self.grad = np.array([rng.random_unit_vector() for _ in range(256)])
        self.perm = [self._permutation(rng) for _ in range(3)]

    @staticmethod
    def _permutation(rng: SampleRng) -> np.ndarray:  # perlin.rs:85-101
        p = list(range(256))
        for i in range(255, 0, -1):
            t = int(rng.random_double() * (i + 1.0))
            p[i], p[t] = p[t], p[i]
        return np.array(p, dtype=np.int64)

    def noise(self, pts: np.ndarray) -> np.ndarray:  # perlin.rs:37-60, interp :104-122
        pts = np.asarray(pts, np.float64)
        fl = np.floor(pts)
        frac = pts - fl
        base = fl.astype(np.int64)
        uu, vv, ww = (frac[..., i] * frac[..., i] * (3.0 - 2.0 * frac[..., i]) for i in range(3))
        acc = np.zeros(pts.shape[:-1])
        for ox in (0, 1):
            for oy in (0, 1):
                for oz in (0, 1):
                    gi = (self.perm[0][(base[..., 0] + ox) & 255] ^ self.perm[1][(base[..., 1] + oy) & 255]
                          ^ self.perm[2][(base[..., 2] + oz) & 255])
                    g = self.grad[gi]
                    wx, wy, wz = frac[..., 0] - ox, frac[..., 1] - oy, frac[..., 2] - oz
                    dot = g[..., 0] * wx + g[..., 1] * wy + g[..., 2] * wz
                    bx = uu if ox else 1.0 - uu
                    by = vv if oy else 1.0 - vv
                    bz = ww if oz else 1.0 - ww
                    acc = acc + bx * by * bz * dot
        return acc

    def turbulence(self, pts: np.ndarray, depth: int = 7) -> np.ndarray:  # perlin.rs:63-76
        acc = np.zeros(np.asarray(pts).shape[:-1])
        p = np.asarray(pts, np.float64)
        w = 1.0
        for _ in range(depth):
            acc = acc + w * self.noise(p)
            w *= 0.5
            p = p * 2.0
        return np.abs(acc)


# ------------------------------------------------------------ camera specs ---
# Plain dicts so both the GPU camera (gartus_amd.api.RayCamera) and the oracle
# camera record can be built from them (see camera_record()).

def camera_record(spec: dict):
    """A gt_camera record (gartus_amd.abi.GtCamera) from a camera spec."""
    from . import abi
    c = abi.default_camera()
    c.image_width = spec["width"]
    c.aspect_ratio = spec["aspect"]
    if spec.get("grid"):
        g = spec["grid"]
        c.pixel_sample_mode = abi.GT_PIXEL_STRATIFIED_GRID
        c.stratified_grid_width = g
        c.samples_per_pixel = g * g
    else:
        c.samples_per_pixel = spec["spp"]
    c.max_depth = spec["depth"]
    c.vertical_fov = spec["vfov"]
    c.lookfrom[:] = spec["lookfrom"]
    c.lookat[:] = spec["lookat"]
    c.view_up[:] = spec.get("vup", (0.0, 1.0, 0.0))
    c.background[0:3] = spec.get("background", (0.7, 0.8, 1.0))
    c.rng_seed = spec.get("seed", 1)
    if spec.get("nee"):
        c.direct_lighting_mode = abi.GT_LIGHTING_NEE
        c.light_pdf_weight = 0.0
    if "rr" in spec:
        c.russian_roulette_min_depth = spec["rr"]
    return c


# ----------------------------------------------------------------- scenes ---

def _rotate_y(p, s, c):  # instance.rs:379-385 (rotate_y_point / rotate_y_vector)
    return (c * p[0] + s * p[2], p[1], -s * p[0] + c * p[2])


def _box_quads(a, b):  # box_object (object.rs:930-976): six (corner, u, v)
    mn = tuple(min(x, y) for x, y in zip(a, b))
    mx = tuple(max(x, y) for x, y in zip(a, b))
    dx, dy, dz = (mx[0] - mn[0], 0.0, 0.0), (0.0, mx[1] - mn[1], 0.0), (0.0, 0.0, mx[2] - mn[2])
    neg = lambda v: tuple(-x for x in v)  # noqa: E731
    return [
        ((mn[0], mn[1], mx[2]), dx, dy),
        ((mx[0], mn[1], mx[2]), neg(dz), dy),
        ((mx[0], mn[1], mn[2]), neg(dx), dy),
        ((mn[0], mn[1], mn[2]), dz, dy),
        ((mn[0], mx[1], mx[2]), dx, neg(dz)),
        ((mn[0], mn[1], mn[2]), dx, dz),
    ]


def cornell():
    """Cornell box (scenes.rs:226-294) as a RayScene: the two boxes are baked
    into six quads each with RotateY + Translate applied (instance.rs:379-409).
    Returns (SceneData, light targets, camera spec)."""
    s = SceneData()
    green = s.add_material(Material(LAMBERTIAN, (0.12, 0.45, 0.15)))
    red = s.add_material(Material(LAMBERTIAN, (0.65, 0.05, 0.05)))
    light = s.add_material(Material(DIFFUSE_LIGHT, (15.0, 15.0, 15.0)))
    white = s.add_material(Material(LAMBERTIAN, (0.73, 0.73, 0.73)))
    light_quad = ((343.0, 554.0, 332.0), (-130.0, 0.0, 0.0), (0.0, 0.0, -105.0))
    walls = [
        (((555.0, 0.0, 0.0), (0.0, 555.0, 0.0), (0.0, 0.0, 555.0)), green),
        (((0.0, 0.0, 0.0), (0.0, 555.0, 0.0), (0.0, 0.0, 555.0)), red),
        (light_quad, light),
        (((0.0, 0.0, 0.0), (555.0, 0.0, 0.0), (0.0, 0.0, 555.0)), white),
        (((555.0, 555.0, 555.0), (-555.0, 0.0, 0.0), (0.0, 0.0, -555.0)), white),
        (((0.0, 0.0, 555.0), (555.0, 0.0, 0.0), (0.0, 555.0, 0.0)), white),
    ]
    for (c, u, v), m in walls:
        s.add(QUAD, [*c, *u, *v], m)
    for size, angle, offset in (((165.0, 330.0, 165.0), 15.0, (265.0, 0.0, 295.0)),
                                ((165.0, 165.0, 165.0), -18.0, (130.0, 0.0, 65.0))):
        rad = angle * math.pi / 180.0
        sn, cs = math.sin(rad), math.cos(rad)
        quads = []
        for c, u, v in _box_quads((0.0, 0.0, 0.0), size):
            rc = _rotate_y(c, sn, cs)
            quads.append([rc[0] + offset[0], rc[1] + offset[1], rc[2] + offset[2],
                          *_rotate_y(u, sn, cs), *_rotate_y(v, sn, cs)])
        s.add(QUAD, quads, white)
    lights = [(QUAD, np.array([*light_quad[0], *light_quad[1], *light_quad[2]]), 1.0)]
    cam = dict(width=256, aspect=1.0, grid=8, depth=8, vfov=40.0, lookfrom=(278.0, 278.0, -800.0),
               lookat=(278.0, 278.0, 0.0), background=(0.0, 0.0, 0.0), seed=1, nee=True)
    return s, lights, cam


def height_map_triangles(heights: np.ndarray, x_size: float, z_size: float, height_scale: float,
                         y_offset: float) -> np.ndarray:
    """PolygonMatrix::from_height_map (polygon_matrix.rs:510-541): two
    triangles (p00, p01, p11) and (p00, p11, p10) per grid cell."""
    rows, cols = heights.shape
    xs = (np.arange(cols) / (cols - 1) - 0.5) * x_size
    zs = (np.arange(rows) / (rows - 1) - 0.5) * z_size
    P = np.zeros((rows, cols, 3))
    P[..., 0] = xs[None, :]
    P[..., 1] = heights * height_scale + y_offset
    P[..., 2] = zs[:, None]
    p00, p10 = P[:-1, :-1], P[:-1, 1:]
    p01, p11 = P[1:, :-1], P[1:, 1:]
    t1 = np.concatenate([p00, p01, p11], axis=-1)
    t2 = np.concatenate([p00, p11, p10], axis=-1)
    return np.stack([t1, t2], axis=2).reshape(-1, 9)


def terrain(grid: int = 512, n_spheres: int = 2000, seed: int = 7):
    """Perlin terrain (grid x grid cells, 2 triangles each) with Lambertian
    albedo from height (8 bands) and n_spheres spheres: 50% Lambertian, 25%
    Metal (fuzz U(0, 0.3)), 25% Dielectric (1.5). Constant sky 1.0."""
    perlin = Perlin(seed)
    size, hscale = 100.0, 6.0
    n = grid + 1
    gx, gz = np.meshgrid(np.linspace(0.0, 4.0, n), np.linspace(0.0, 4.0, n))
    pts = np.stack([gx, np.full_like(gx, 0.5), gz], axis=-1)
    h = perlin.turbulence(pts, 7)
    tris = height_map_triangles(h, size, size, hscale, 0.0)
    s = SceneData()
    bands = 8
    ramp = [(0.20 + 0.08 * k, 0.35 + 0.05 * k, 0.15 + 0.07 * k) for k in range(bands)]
    mats = [s.add_material(Material(LAMBERTIAN, c)) for c in ramp]
    cy = tris[:, [1, 4, 7]].mean(axis=1)
    band = np.clip(((cy - cy.min()) / max(cy.max() - cy.min(), 1e-9) * bands).astype(np.int64), 0, bands - 1)
    s.add(TRIANGLE, tris, np.array(mats, np.uint32)[band])
    rng = SampleRng(seed * 7919 + 11)
    spheres, sm = [], []
    for _ in range(n_spheres):
        u, v = rng.random_double(), rng.random_double()
        r = rng.random_range(0.2, 1.0)
        fx, fz = u * grid, v * grid
        ix, iz = min(int(fx), grid - 1), min(int(fz), grid - 1)
        y = h[iz, ix] * hscale
        x, z = (u - 0.5) * size, (v - 0.5) * size
        pick = rng.random_double()
        if pick < 0.5:
            m = Material(LAMBERTIAN, rng.random_vector_range(0.1, 0.9))
        elif pick < 0.75:
            m = Material(METAL, rng.random_vector_range(0.5, 1.0), rng.random_range(0.0, 0.3))
        else:
            m = Material(DIELECTRIC, (1.0, 1.0, 1.0), 1.5)
        spheres.append([x, y + r, z, r])
        sm.append(s.add_material(m))
    s.add(SPHERE, spheres, np.array(sm, np.uint32))
    cam = dict(width=1280, aspect=16.0 / 9.0, spp=32, depth=6, vfov=45.0, lookfrom=(0.0, 28.0, 62.0),
               lookat=(0.0, 2.0, 0.0), background=(1.0, 1.0, 1.0), seed=1)
    return s, cam


def sky_environment(width: int = 2048, height: int = 1024, sun_radiance: float = 5e4,
                    sun_diameter_deg: float = 0.5, sun_elevation_deg: float = 35.0,
                    sun_azimuth_deg: float = 60.0) -> tuple:
    """Procedural FP32 lat-long sky: a vertical gradient plus one sun disc.
    Rows run from +y (theta = 0) down to -y; column x covers
    phi = 2 pi (x + 0.5) / width measured from +x towards +z, the lookup of
    EnvironmentLight (environment.rs:193-198). Returns (rgb, sun direction)."""
    y = (np.arange(height) + 0.5) / height
    theta = np.pi * y
    dy = np.cos(theta)
    zen = np.array([0.18, 0.32, 0.75])
    hor = np.array([0.85, 0.88, 0.95])
    gnd = np.array([0.25, 0.22, 0.20])
    t_up = np.clip(dy, 0.0, 1.0)[:, None]
    sky = hor[None, :] * (1.0 - t_up) + zen[None, :] * t_up
    col = np.where((dy >= 0.0)[:, None], sky, gnd[None, :] * (0.5 + 0.5 * np.clip(dy + 1.0, 0, 1))[:, None])
    rgb = np.repeat(col[:, None, :], width, axis=1).astype(np.float64)
    el, az = math.radians(sun_elevation_deg), math.radians(sun_azimuth_deg)
    sun = np.array([math.cos(el) * math.cos(az), math.sin(el), math.cos(el) * math.sin(az)])
    phi = 2.0 * np.pi * (np.arange(width) + 0.5) / width
    st = np.sin(theta)[:, None]
    dirs = np.stack([st * np.cos(phi)[None, :], np.repeat(np.cos(theta)[:, None], width, 1),
                     st * np.sin(phi)[None, :]], axis=-1)
    cosang = dirs @ sun
    disc = cosang >= math.cos(math.radians(sun_diameter_deg * 0.5))
    rgb[disc] = sun_radiance
    return rgb.astype(np.float32), sun


def ggx_environment(n_random: int = 1000, seed: int = 3, env_size=(2048, 1024), sun_elevation_deg: float = 35.0,
                    sun_azimuth_deg: float = 60.0):
    """10x10 GGX sphere grid (roughness 0.02..1.0 across, metallic 0/1 by
    row) + n_random Dielectric/Metal/Lambertian spheres on a ground sphere.
    Metallic 1 = GgxMicrofacet(base colour, r); metallic 0 =
    GgxMicrofacet(grey F0 0.04, r). Lit only by the HDR environment (NEE on)."""
    s = SceneData()
    ground = s.add_material(Material(LAMBERTIAN, (0.45, 0.45, 0.45)))
    spheres, mids = [[0.0, -1000.0, 0.0, 1000.0]], [ground]
    rough = np.linspace(0.02, 1.0, 10)
    bases = [(0.95, 0.64, 0.54), (0.91, 0.92, 0.92), (0.95, 0.93, 0.88), (0.56, 0.57, 0.58), (1.00, 0.71, 0.29)]
    for row in range(10):
        metallic = row % 2 == 1
        for col in range(10):
            color = bases[row // 2] if metallic else (0.04, 0.04, 0.04)
            mids.append(s.add_material(Material(GGX, color, float(rough[col]))))
            spheres.append([(col - 4.5) * 1.0, 0.4, (row - 4.5) * 1.0, 0.4])
    rng = SampleRng(seed)
    placed = 0
    while placed < n_random:
        x, z = rng.random_range(-16.0, 16.0), rng.random_range(-16.0, 16.0)
        r = rng.random_range(0.1, 0.25)
        if abs(x) < 5.3 and abs(z) < 5.3:
            continue
        pick = rng.random_double()
        if pick < 1.0 / 3.0:
            m = Material(DIELECTRIC, (1.0, 1.0, 1.0), 1.5)
        elif pick < 2.0 / 3.0:
            m = Material(METAL, rng.random_vector_range(0.5, 1.0), rng.random_range(0.0, 0.4))
        else:
            m = Material(LAMBERTIAN, rng.random_vector_range(0.05, 0.9))
        mids.append(s.add_material(m))
        spheres.append([x, r, z, r])
        placed += 1
    s.add(SPHERE, spheres, np.array(mids, np.uint32))
    env, _ = sky_environment(*env_size, sun_elevation_deg=sun_elevation_deg, sun_azimuth_deg=sun_azimuth_deg)
    cam = dict(width=1920, aspect=16.0 / 9.0, spp=16, depth=10, vfov=32.0, lookfrom=(0.0, 5.5, 15.0),
               lookat=(0.0, 0.3, 0.0), seed=1, nee=True)
    return s, env, cam


def c3_style_scene(seed: int, width: int = 256, height: int = 256):
    """A training scene in the style of the GGX/environment workload, from
    SampleRng(seed): its own random sphere field (150-600 spheres), sun
    direction (elevation 12-70 degrees, any azimuth) and camera (orbit,
    height, distance, field of view, look-at jitter), rendered at
    width x height. Returns (SceneData, env rgb, camera spec)."""
    rng = SampleRng(seed)
    n_random = 150 + int(rng.random_double() * 450)
    elev = rng.random_range(12.0, 70.0)
    azim = rng.random_range(0.0, 360.0)
    data, env, cam = ggx_environment(n_random=n_random, seed=seed * 7 + 1, sun_elevation_deg=elev,
                                     sun_azimuth_deg=azim)
    ang = rng.random_range(0.0, 2.0 * math.pi)
    dist = rng.random_range(6.0, 18.0)
    eye = (dist * math.sin(ang), rng.random_range(1.0, 7.0), dist * math.cos(ang))
    at = (rng.random_range(-3.0, 3.0), rng.random_range(0.0, 0.8), rng.random_range(-3.0, 3.0))
    cam.update(width=width, aspect=width / height, vfov=rng.random_range(20.0, 50.0), lookfrom=eye, lookat=at,
               seed=1 + seed)
    return data, env, cam


def bench_scene():
    """benches/render_performance.rs:10-40 and its camera (:32-40)."""
    s = SceneData()
    ground = s.add_material(Material(LAMBERTIAN, (0.45, 0.47, 0.43)))
    warm = s.add_material(Material(LAMBERTIAN, (0.75, 0.45, 0.32)))
    cool = s.add_material(Material(LAMBERTIAN, (0.25, 0.45, 0.78)))
    s.add(SPHERE, [0.0, -1000.65, -4.0, 1000.0], ground)
    for row in range(8):
        for column in range(16):
            x = (column - 7.5) * 0.45
            z = -1.4 - row * 0.55
            y = -0.35 + 0.04 * ((row + column) % 3)
            s.add(SPHERE, [x, y, z, 0.16], warm if (row + column) % 2 == 0 else cool)
    cam = dict(width=96, aspect=16.0 / 9.0, spp=2, depth=5, vfov=35.0, lookfrom=(0.0, 1.1, 2.6),
               lookat=(0.0, -0.25, -3.4), seed=0x5EED2026)
    return s, cam


def randomized_ray_scene(seed: int) -> SceneData:
    """The reference test scene of raytracing.rs:242-289: 8 spheres, 4 quads
    and 4 triangles from SampleRng(seed), one Lambertian material."""
    rng = SampleRng(seed)
    s = SceneData()
    m = s.add_material(Material(LAMBERTIAN, (0.4, 0.5, 0.6)))
    for _ in range(8):
        c = (rng.random_range(-3.0, 3.0), rng.random_range(-2.0, 2.0), rng.random_range(-8.0, -1.0))
        r = rng.random_range(0.1, 0.45)
        s.add(SPHERE, [*c, r], m)
    for _ in range(4):
        c = (rng.random_range(-3.0, 3.0), rng.random_range(-2.0, 2.0), rng.random_range(-8.0, -1.0))
        half = rng.random_range(0.15, 0.5)
        s.add(QUAD, [c[0] - half, c[1] - half, c[2], 2.0 * half, 0.0, 0.0, 0.0, 2.0 * half, 0.0], m)
    for _ in range(4):
        c = (rng.random_range(-3.0, 3.0), rng.random_range(-2.0, 2.0), rng.random_range(-8.0, -1.0))
        sc = rng.random_range(0.2, 0.7)
        dz = rng.random_range(-0.2, 0.2)
        s.add(TRIANGLE, [c[0] - sc, c[1] - sc, c[2], c[0] + sc, c[1] - sc, c[2], c[0], c[1] + sc, c[2] + dz], m)
    return s


def final_scene(seed: int = 61) -> SceneData:
    """final_scene_ray_scene (scenes.rs:563-624)."""
    rng = SampleRng(seed)
    s = SceneData()

    def add(center, radius, m):
        s.add(SPHERE, [*center, radius], s.add_material(m))

    add((0.0, -1000.0, 0.0), 1000.0, Material(LAMBERTIAN, (0.5, 0.5, 0.5)))
    for a in range(-11, 11):
        for b in range(-11, 11):
            choose = rng.random_double()
            c = (a + 0.9 * rng.random_double(), 0.2, b + 0.9 * rng.random_double())
            if math.dist(c, (4.0, 0.2, 0.0)) <= 0.9:
                continue
            if choose < 0.8:
                v1, v2 = rng.random_vector(), rng.random_vector()
                add(c, 0.2, Material(LAMBERTIAN, tuple(x * y for x, y in zip(v1, v2))))
            elif choose < 0.95:
                alb = rng.random_vector_range(0.5, 1.0)
                add(c, 0.2, Material(METAL, alb, rng.random_range(0.0, 0.5)))
            else:
                add(c, 0.2, Material(DIELECTRIC, (1.0, 1.0, 1.0), 1.5))
    add((0.0, 1.0, 0.0), 1.0, Material(DIELECTRIC, (1.0, 1.0, 1.0), 1.5))
    add((-4.0, 1.0, 0.0), 1.0, Material(LAMBERTIAN, (0.4, 0.2, 0.1)))
    add((4.0, 1.0, 0.0), 1.0, Material(METAL, (0.7, 0.6, 0.5), 0.0))
    return s


# ------------------------------------------------------- batched scenes ---

def torus_triangles(center, r_small: float, r_big: float, steps: int) -> np.ndarray:
    """PolygonMatrix::add_torus (polygon_matrix.rs:795-864): steps x steps
    quads of (phi, theta) rings, two triangles (p0, p3, p2), (p0, p2, p1)
    each, 2 steps^2 triangles."""
    t = np.arange(steps) / steps * 2.0 * np.pi
    st, ct = np.sin(t), np.cos(t)
    sp, cp = np.sin(t), np.cos(t)
    rad = r_small * ct[None, :] + r_big                     # [1, theta]
    P = np.stack([cp[:, None] * rad + center[0],           # [phi, theta, xyz]
                  np.broadcast_to(r_small * st[None, :] + center[1], (steps, steps)),
                  -sp[:, None] * rad + center[2]], axis=-1)
    nxt = np.roll(np.arange(steps), -1)
    p0, p1 = P, P[:, nxt]
    p3, p2 = P[nxt], P[nxt][:, nxt]
    t1 = np.concatenate([p0, p3, p2], axis=-1)
    t2 = np.concatenate([p0, p2, p1], axis=-1)
    return np.stack([t1, t2], axis=2).reshape(-1, 9)


def crystal_triangles(center, sides: int, radius: float, height: float, yaw: float) -> np.ndarray:
    """PolygonMatrix::add_crystal_with_yaw (polygon_matrix.rs:680-704): a
    double pyramid, 2 sides triangles."""
    cx, cy, cz = center
    a = yaw + np.arange(sides) / sides * 2.0 * np.pi
    ring = np.stack([cx + np.cos(a) * radius, np.full(sides, cy), cz + np.sin(a) * radius], axis=-1)
    nxt = np.roll(ring, -1, axis=0)
    top = np.broadcast_to([cx, cy + height * 0.5, cz], ring.shape)
    bot = np.broadcast_to([cx, cy - height * 0.5, cz], ring.shape)
    up = np.concatenate([top, ring, nxt], axis=-1)
    down = np.concatenate([bot, nxt, ring], axis=-1)
    return np.stack([up, down], axis=1).reshape(-1, 9)


def crystal_field_triangles(centers: np.ndarray, sides: int, radius: np.ndarray, height: np.ndarray,
                            yaw: np.ndarray) -> np.ndarray:
    """crystal_triangles for many crystals at once, crystal after crystal."""
    c = np.asarray(centers, np.float64)
    a = yaw[:, None] + np.arange(sides)[None, :] / sides * 2.0 * np.pi       # [crystal, side]
    ring = np.stack([c[:, 0:1] + np.cos(a) * radius[:, None], np.broadcast_to(c[:, 1:2], a.shape),
                     c[:, 2:3] + np.sin(a) * radius[:, None]], axis=-1)
    nxt = np.roll(ring, -1, axis=1)
    top = np.broadcast_to((c + np.stack([0 * height, height * 0.5, 0 * height], -1))[:, None, :], ring.shape)
    bot = np.broadcast_to((c - np.stack([0 * height, height * 0.5, 0 * height], -1))[:, None, :], ring.shape)
    up = np.concatenate([top, ring, nxt], axis=-1)
    down = np.concatenate([bot, nxt, ring], axis=-1)
    return np.stack([up, down], axis=2).reshape(-1, 9)


BATCH_RESOLUTIONS = ((256, 256), (512, 512), (1280, 720), (1920, 1080))


def batch_scene(seed: int):
    """One scene of the 64-scene batched workload, from SampleRng(seed).

    Triangle count N = exp(U(ln 1e3, ln 2e6)) over one of three families: a
    torus (add_torus), a field of crystals (add_crystal_with_yaw), or a Perlin
    height map (from_height_map). Materials are mixed Lambertian / GGX /
    Dielectric, with 1-8 DiffuseLight quads overhead (also the NEE light
    list). The camera takes a resolution from {256^2, 512^2, 720p, 1080p}, at
    8 spp, depth 8, NEE. Returns (SceneData, lights, camera spec)."""
    rng = SampleRng(seed)
    n_target = math.exp(rng.random_range(math.log(1e3), math.log(2e6)))
    family = int(rng.random_double() * 3)
    s = SceneData()
    mats = []
    for k in range(6):
        pick = rng.random_double()
        if pick < 0.45:
            mats.append(s.add_material(Material(LAMBERTIAN, rng.random_vector_range(0.15, 0.85))))
        elif pick < 0.8:
            m = Material(GGX, rng.random_vector_range(0.3, 0.95), rng.random_range(0.05, 0.8))
            mats.append(s.add_material(m))
        else:
            mats.append(s.add_material(Material(DIELECTRIC, (1.0, 1.0, 1.0), 1.5)))
    floor = s.add_material(Material(LAMBERTIAN, (0.6, 0.6, 0.6)))
    extent = 10.0
    if family == 0:
        steps = max(int(math.sqrt(n_target / 2.0)), 8)
        tris = torus_triangles((0.0, 3.0, 0.0), 1.6, 4.5, steps)
        tri_mats = np.array(mats, np.uint32)[(np.arange(len(tris)) // max(len(tris) // 6, 1)) % 6]
    elif family == 1:
        sides = 6 + int(rng.random_double() * 10)
        count = max(int(n_target / (2 * sides)), 1)
        side = int(math.ceil(math.sqrt(count)))
        params = np.empty((count, 4))
        pick = np.empty(count, np.int64)
        for i in range(count):  # r, h, yaw, material in draw order
            r = extent / side * rng.random_range(0.25, 0.45)
            params[i] = (r, r * rng.random_range(1.5, 4.0), rng.random_range(0.0, math.pi), 0.0)
            pick[i] = int(rng.random_double() * 6) % 6
        i = np.arange(count)
        x = ((i % side) + 0.5) / side * 2.0 * extent - extent
        z = ((i // side) + 0.5) / side * 2.0 * extent - extent
        r, h, yaw = params[:, 0], params[:, 1], params[:, 2]
        tris = crystal_field_triangles(np.stack([x, h * 0.5, z], -1), sides, r, h, yaw)
        tri_mats = np.repeat(np.array(mats, np.uint32)[pick], 2 * sides)
    else:
        grid = max(int(math.sqrt(n_target / 2.0)), 2)
        perlin = Perlin(seed & 0xFFFF)
        n = grid + 1
        gx, gz = np.meshgrid(np.linspace(0.0, 3.0, n), np.linspace(0.0, 3.0, n))
        h = perlin.turbulence(np.stack([gx, np.full_like(gx, 0.25), gz], axis=-1), 5)
        tris = height_map_triangles(h, 2.0 * extent, 2.0 * extent, 4.0, 0.01)
        cy = tris[:, [1, 4, 7]].mean(axis=1)
        band = np.clip(((cy - cy.min()) / max(np.ptp(cy), 1e-9) * 6).astype(np.int64), 0, 5)
        tri_mats = np.array(mats, np.uint32)[band]
    s.add(QUAD, [[-3 * extent, 0.0, -3 * extent, 6 * extent, 0.0, 0.0, 0.0, 0.0, 6 * extent]], floor)
    s.add(TRIANGLE, tris, tri_mats)
    lights = []
    n_lights = 1 + int(rng.random_double() * 8)
    for _ in range(n_lights):
        size = rng.random_range(1.5, 4.0)
        x, z = rng.random_range(-extent, extent), rng.random_range(-extent, extent)
        y = rng.random_range(14.0, 20.0)
        power = rng.random_range(4.0, 12.0)
        lm = s.add_material(Material(DIFFUSE_LIGHT, (power, power * 0.95, power * 0.85)))
        q = [x - size / 2, y, z + size / 2, size, 0.0, 0.0, 0.0, 0.0, -size]  # faces -y
        s.add(QUAD, [q], lm)
        lights.append((QUAD, np.array(q, np.float64), 1.0))
    w, hh = BATCH_RESOLUTIONS[int(rng.random_double() * len(BATCH_RESOLUTIONS)) % len(BATCH_RESOLUTIONS)]
    ang = rng.random_range(0.0, 2.0 * math.pi)
    dist = rng.random_range(20.0, 30.0)
    cam = dict(width=w, aspect=w / hh, spp=8, depth=8, vfov=45.0,
               lookfrom=(dist * math.cos(ang), rng.random_range(8.0, 16.0), dist * math.sin(ang)),
               lookat=(0.0, 2.0, 0.0), background=(0.02, 0.02, 0.03), seed=1 + seed, nee=True)
    return s, lights, cam


def batch_scenes(n: int = 64, seed: int = 1000):
    """The batched workload: n scenes from consecutive seeds."""
    return [batch_scene(seed + i) for i in range(n)]


def smooth_sphere_mesh(center, radius: float, n_theta: int = 16, n_phi: int = 32):
    """A latitude-longitude tessellation of a sphere as mesh triangles with
    per-vertex normals (the sphere's radial directions), the kind of
    TriangleMesh with vertex normals mesh.rs:24-151 shades smoothly. Returns
    positions (n, 9) and normals (n, 9), outward-wound."""
    c = np.asarray(center, np.float64)
    th = np.linspace(0.0, math.pi, n_theta + 1)
    ph = np.linspace(0.0, 2.0 * math.pi, n_phi + 1)

    def vert(i, j):
        n = np.array([math.sin(th[i]) * math.cos(ph[j]), math.cos(th[i]), math.sin(th[i]) * math.sin(ph[j])])
        return c + radius * n, n

    pos, nrm = [], []
    for i in range(n_theta):
        for j in range(n_phi):
            (a, na), (b, nb) = vert(i, j), vert(i + 1, j)
            (cc, nc), (d, nd) = vert(i + 1, j + 1), vert(i, j + 1)
            for tri in ((a, na, d, nd, b, nb), (b, nb, d, nd, cc, nc)):
                p0, n0, p1, n1, p2, n2 = tri
                if np.linalg.norm(np.cross(p1 - p0, p2 - p0)) < 1e-12:  # the pole's degenerate triangles
                    continue
                pos.append(np.concatenate([p0, p1, p2]))
                nrm.append(np.concatenate([n0, n1, n2]))
    return np.array(pos), np.array(nrm)
