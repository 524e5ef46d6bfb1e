"""Synthetic workloads: the trefoil closest-hit benchmark, its ray sets, and the
path-tracing scenes (sphere room, full-feature scene, batched and served jobs).

Every generator is seeded and deterministic. Random draws come from PCG32
(O'Neill 2014), the generator the reference renders with
(sampling/pcg32.hpp:30-57); it is vectorised here with jump-ahead.
"""
import math

import numpy as np

from . import _abi
from .scene import (CameraSettings, CheckerTexture, ConductorMaterial, ConstantTexture, DielectricMaterial,
                    DiffuseMaterial, MeshGeometry, MeshInstance, MixMaterial, PerlinPattern, PointLight, Scene,
                    SkyLight, Sphere, Transform, TriangleMesh, FLAT, SMOOTH, fitted, grounded)

MASK64 = (1 << 64) - 1
PCG_MULT = 6364136223846793005


# ----------------------------------------------------------------------- PCG32

def _pcg_step(state, inc):
    return (state * PCG_MULT + inc) & MASK64


def _pcg_jump(delta, inc):
    """(mult, plus) such that state_{k+delta} = mult * state_k + plus."""
    acc_m, acc_p, cur_m, cur_p = 1, 0, PCG_MULT, inc
    while delta > 0:
        if delta & 1:
            acc_m = (acc_m * cur_m) & MASK64
            acc_p = (acc_p * cur_m + cur_p) & MASK64
        cur_p = ((cur_m + 1) * cur_p) & MASK64
        cur_m = (cur_m * cur_m) & MASK64
        delta >>= 1
    return acc_m, acc_p


def _pcg_output(old):
    """PCG-XSH-RR output of uint64 states (numpy array)."""
    xs = (((old >> np.uint64(18)) ^ old) >> np.uint64(27)).astype(np.uint32)
    rot = (old >> np.uint64(59)).astype(np.uint32)
    return (xs >> rot) | (xs << ((np.uint32(32) - rot) & np.uint32(31)))


def pcg32_stream(seed, stream, n):
    """The first n outputs of PCG32 seeded (seed, stream), as uint32."""
    inc = ((stream << 1) | 1) & MASK64
    state = 0
    state = _pcg_step(state, inc)
    state = (state + seed) & MASK64
    state = _pcg_step(state, inc)
    block = 4096
    base = np.empty(block, dtype=np.uint64)
    s = state
    for j in range(block):
        base[j] = s
        s = _pcg_step(s, inc)
    nblocks = (n + block - 1) // block
    mults = np.empty(nblocks, dtype=np.uint64)
    plus = np.empty(nblocks, dtype=np.uint64)
    m1, p1 = _pcg_jump(block, inc)
    m, p = 1, 0
    for i in range(nblocks):
        mults[i], plus[i] = m, p
        m, p = (m1 * m) & MASK64, (m1 * p + p1) & MASK64
    with np.errstate(over="ignore"):
        states = (mults[:, None] * base[None, :] + plus[:, None]).reshape(-1)[:n]
    return _pcg_output(states)


def pcg32_reals(seed, stream, n):
    """Reference nextReal(): u32 * 2^-32 in FP64."""
    return pcg32_stream(seed, stream, n).astype(np.float64) * (1.0 / 4294967296.0)


class Pcg32:
    """Scalar PCG32 for scene generation (reference semantics)."""

    def __init__(self, seed, stream):
        self.inc = ((stream << 1) | 1) & MASK64
        self.state = 0
        self.next()
        self.state = (self.state + seed) & MASK64
        self.next()

    def next(self):
        old = self.state
        self.state = (old * PCG_MULT + self.inc) & MASK64
        xs = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((xs >> rot) | (xs << ((32 - rot) & 31))) & 0xFFFFFFFF

    def real(self):
        return self.next() * (1.0 / 4294967296.0)

    def uniform(self, lo, hi):
        return lo + (hi - lo) * self.real()


# ------------------------------------------------------------------- trefoil

def trefoil_mesh(segments=512, sides=32, radius=0.35):
    """Tube of radius `radius` around x = sin t + 2 sin 2t, y = cos t - 2 cos 2t,
    z = -sin 3t, in a Frenet frame: segments x sides quads, two triangles each,
    wound so normals point out of the tube. FP64, rounded once to FP32."""
    t = np.arange(segments, dtype=np.float64) * (2 * math.pi / segments)
    c = np.stack([np.sin(t) + 2 * np.sin(2 * t), np.cos(t) - 2 * np.cos(2 * t), -np.sin(3 * t)], 1)
    d1 = np.stack([np.cos(t) + 4 * np.cos(2 * t), -np.sin(t) + 4 * np.sin(2 * t), -3 * np.cos(3 * t)], 1)
    d2 = np.stack([-np.sin(t) - 8 * np.sin(2 * t), -np.cos(t) + 8 * np.cos(2 * t), 9 * np.sin(3 * t)], 1)
    T = d1 / np.linalg.norm(d1, axis=1, keepdims=True)
    B = np.cross(d1, d2)
    B /= np.linalg.norm(B, axis=1, keepdims=True)
    N = np.cross(B, T)
    phi = np.arange(sides, dtype=np.float64) * (2 * math.pi / sides)
    ring = np.cos(phi)[None, :, None] * N[:, None, :] + np.sin(phi)[None, :, None] * B[:, None, :]
    pos = (c[:, None, :] + radius * ring).reshape(-1, 3).astype(np.float32)
    i = np.arange(segments)[:, None]
    j = np.arange(sides)[None, :]
    a = i * sides + j
    b = ((i + 1) % segments) * sides + j
    cc = ((i + 1) % segments) * sides + (j + 1) % sides
    d = i * sides + (j + 1) % sides
    f1 = np.stack([a, d, cc], -1).reshape(-1, 3)
    f2 = np.stack([a, cc, b], -1).reshape(-1, 3)
    faces = np.stack([f1, f2], 1).reshape(-1, 3).astype(np.uint32)
    return TriangleMesh(pos, faces)


def trefoil_scene(shading=FLAT):
    """The closest-hit benchmark scene: one trefoil instance, identity placement."""
    mesh = trefoil_mesh()
    geo = MeshGeometry(mesh, shading)
    s = Scene()
    s.add(MeshInstance(geo, Transform()), DiffuseMaterial(ConstantTexture((0.73, 0.73, 0.73))))
    return s, mesh


def _rays(org, d, tmin=0.0, tmax=np.inf):
    n = len(d)
    r = np.zeros((n, 8), dtype=np.float32)
    r[:, 0:3] = org
    r[:, 3] = tmin
    r[:, 4:7] = d
    r[:, 7] = tmax
    return r


def coherent_rays(mesh, n_side=4096, distance=8.0, vfov=50.0):
    """n_side^2 rays through pixel centres of a pinhole camera `distance` from the
    mesh-box centre, looking at it along -z. Returns float32 [n, 8] (ab_ray)."""
    lo, hi = mesh.bounds()
    centre = lo / 2 + hi / 2
    eye = centre + np.array([0.0, 0.0, distance])
    h = math.tan(math.radians(vfov) / 2)
    px = (np.arange(n_side, dtype=np.float64) + 0.5) / n_side
    fx, fy = np.meshgrid(2 * px - 1, 1 - 2 * px)
    d = np.stack([fx.reshape(-1) * h, fy.reshape(-1) * h, -np.ones(n_side * n_side)], 1)
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return _rays(eye.astype(np.float32), d.astype(np.float32))


def incoherent_rays(mesh, n, radius=6.0, seed=0x5EED):
    """n rays with origins uniform on a sphere of `radius` around the mesh-box
    centre (uniform-sphere warp of PCG32 stream 1, reference warps.cpp:37-45) aimed
    at points uniform in the mesh box (PCG32 stream 2)."""
    lo, hi = mesh.bounds()
    centre = lo / 2 + hi / 2
    u = pcg32_reals(seed, 1, 2 * n).reshape(n, 2)
    z = 1 - 2 * u[:, 0]
    r = np.sqrt(np.maximum(0.0, 1 - z * z))
    phi = 2 * math.pi * u[:, 1]
    org = centre + radius * np.stack([r * np.cos(phi), r * np.sin(phi), z], 1)
    v = pcg32_reals(seed, 2, 3 * n).reshape(n, 3)
    tgt = lo + v * (hi - lo)
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    return _rays(org.astype(np.float32), d.astype(np.float32))


# ---------------------------------------------------------------- room scene

GOLD = dict(eta=(0.156, 0.424, 1.383), k=(3.602, 2.472, 1.915))
COPPER = dict(eta=(0.200, 0.924, 1.102), k=(3.912, 2.452, 2.142))
SILVER = dict(eta=(0.155, 0.117, 0.138), k=(4.83, 3.12, 2.15))


def room_materials():
    return {
        "white": DiffuseMaterial(ConstantTexture((0.73, 0.73, 0.73))),
        "red": DiffuseMaterial(ConstantTexture((0.65, 0.05, 0.05))),
        "green": DiffuseMaterial(ConstantTexture((0.12, 0.45, 0.15))),
        "glass": DielectricMaterial(1.5, 0.0),
        "frosted": DielectricMaterial(1.5, 0.2),
        "gold": ConductorMaterial(GOLD["eta"], GOLD["k"], 0.25),
        "copper": ConductorMaterial(COPPER["eta"], COPPER["k"], 0.3),
        "silver": ConductorMaterial(SILVER["eta"], SILVER["k"], 0.0),
    }


FIVE_KINDS = ("glass", "frosted", "gold", "copper", "silver")
ROOM_CAMERA = CameraSettings((0.0, 2.4, 7.8), (0.0, 1.9, -0.5), (0.0, 1.0, 0.0), 38.0, 16.0 / 9.0)
WALL_R = 1e5


def _floor_positions(rng, count, radius, min_gap, x_range=(-2.1, 2.1), z_range=(-4.6, 1.8), tries=10000):
    pts = []
    for _ in range(tries):
        if len(pts) == count:
            break
        x, z = rng.uniform(*x_range), rng.uniform(*z_range)
        if all((x - px) ** 2 + (z - pz) ** 2 >= (2 * radius + min_gap) ** 2 for px, pz in pts):
            pts.append((x, z))
    return pts


def room_scene(seed=7, floor_sphere=True):
    """Sphere room: five radius-1e5 walls bounding x in [-3, 3], y in [0, 6] and
    the back wall z = -6 (front open); five radius-0.8 spheres at seeded floor
    positions (smooth glass, frosted glass, gold, copper, silver); a spherical area
    light, a point light and a sky. Light order: area, point, sky."""
    m = room_materials()
    s = Scene()
    if floor_sphere:
        s.add(Sphere((0.0, -WALL_R, 0.0), WALL_R), m["white"])
    s.add(Sphere((0.0, 6.0 + WALL_R, 0.0), WALL_R), m["white"])
    s.add(Sphere((0.0, 0.0, -6.0 - WALL_R), WALL_R), m["white"])
    s.add(Sphere((-3.0 - WALL_R, 0.0, 0.0), WALL_R), m["red"])
    s.add(Sphere((3.0 + WALL_R, 0.0, 0.0), WALL_R), m["green"])
    rng = Pcg32(seed, 11)
    for (x, z), kind in zip(_floor_positions(rng, 5, 0.8, 0.1), FIVE_KINDS):
        s.add(Sphere((x, 0.8, z), 0.8), m[kind])
    s.add_area_light(Sphere((0.0, 5.6, -1.2), 1.1), (9.0, 7.6, 5.8))
    s.add_light(PointLight((0.8, 4.2, 3.2), (1.2, 2.0, 3.6)))
    s.set_sky(SkyLight.uniform((0.05, 0.06, 0.08)))
    return s


def checker_floor(x0=-3.5, x1=3.5, z0=-6.5, z1=12.0):
    """Two triangles at y = 0 facing up, texture coordinates = (x, z)."""
    pos = np.array([[x0, 0, z1], [x1, 0, z1], [x1, 0, z0], [x0, 0, z0]], dtype=np.float32)
    faces = np.array([[0, 1, 2], [0, 2, 3]], dtype=np.uint32)
    uv = np.array([[x0, z1, x1, z1, x1, z0], [x0, z1, x1, z0, x0, z0]], dtype=np.float32)
    return TriangleMesh(pos, faces, uv)


def full_scene(seed=7, knot=None):
    """The room with a two-triangle checker floor (even 0.73, odd 0.3, frequency
    1.25), three smooth trefoil instances sharing one BLAS (gold, glass, copper;
    each fit to 2.6, turned about y by seeded angles, grounded), and a Perlin
    marble sphere mixing silver with porcelain veins."""
    m = room_materials()
    s = Scene()
    floor = MeshGeometry(checker_floor(), FLAT)
    s.add(MeshInstance(floor, Transform()),
          DiffuseMaterial(CheckerTexture((0.73, 0.73, 0.73), (0.3, 0.3, 0.3), 1.25)))
    s.add(Sphere((0.0, 6.0 + WALL_R, 0.0), WALL_R), m["white"])
    s.add(Sphere((0.0, 0.0, -6.0 - WALL_R), WALL_R), m["white"])
    s.add(Sphere((-3.0 - WALL_R, 0.0, 0.0), WALL_R), m["red"])
    s.add(Sphere((3.0 + WALL_R, 0.0, 0.0), WALL_R), m["green"])
    geo = MeshGeometry(knot if knot is not None else trefoil_mesh(), SMOOTH)
    rng = Pcg32(seed, 12)
    spots = [(-1.55, -3.3), (1.45, -2.2), (-0.2, -0.4)]
    for (x, z), kind in zip(spots, ("gold", "glass", "copper")):
        angle = rng.uniform(0.0, 360.0)
        place = fitted(Transform(), geo, 2.6)
        place = Transform.rotation((0.0, 1.0, 0.0), angle) * place
        place = grounded(place, geo)
        place = Transform.translation((x, 0.0, z)) * place
        s.add(MeshInstance(geo, place), m[kind])
    porcelain = DiffuseMaterial(ConstantTexture((0.9, 0.88, 0.85)))
    marble = MixMaterial(m["silver"], porcelain, PerlinPattern(3.0 + rng.uniform(0.0, 1.0), 6.0), 0.5)
    s.add(Sphere((2.0, 0.7, 1.4), 0.7), marble)
    s.add_area_light(Sphere((0.0, 5.6, -1.2), 1.1), (9.0, 7.6, 5.8))
    s.add_light(PointLight((0.8, 4.2, 3.2), (1.2, 2.0, 3.6)))
    s.set_sky(SkyLight.uniform((0.05, 0.06, 0.08)))
    return s


def sphere_variant_scene(seed, n_spheres, base="room", knot=None):
    """A seeded variant of the sphere room (base="room") or the full-feature scene
    (base="full") holding `n_spheres` extra random spheres, each made of one of the
    five room materials (smooth glass, frosted glass, gold, copper, silver) drawn
    uniformly."""
    s = room_scene(seed) if base == "room" else full_scene(seed, knot)
    m = room_materials()
    rng = Pcg32(seed, 21)
    for _ in range(n_spheres):
        r = rng.uniform(0.08, 0.35)
        x = rng.uniform(-2.8 + r, 2.8 - r)
        z = rng.uniform(-5.7 + r, 3.0)
        y = r if rng.real() < 0.6 else rng.uniform(r, 4.5)
        kind = FIVE_KINDS[min(int(rng.real() * 5), 4)]
        s.add(Sphere((x, y, z), r), m[kind])
    return s


RESOLUTIONS = ((256, 144), (640, 360), (960, 540), (1920, 1080))
SPAS = (2, 4, 8, 16)  # samples per axis: 4, 16, 64, 256 spp


def job_mix(n_jobs, seed=3, resolutions=RESOLUTIONS, spas=SPAS, max_spheres=200, knot=None):
    """The batched-rendering job list: resolutions and spp drawn uniformly from the
    given sets, scenes seeded variants of the room and the full scene holding
    5..max_spheres random spheres. Returns [(scene, (w, h), spa, max_depth, seed)]."""
    rng = Pcg32(seed, 31)
    jobs = []
    for i in range(n_jobs):
        w, h = resolutions[min(int(rng.real() * len(resolutions)), len(resolutions) - 1)]
        spa = spas[min(int(rng.real() * len(spas)), len(spas) - 1)]
        base = "room" if rng.real() < 0.5 else "full"
        n_sph = 5 + int(rng.real() * (max_spheres - 4))
        scene = sphere_variant_scene(1000 + i, n_sph, base, knot)
        jobs.append((scene, (w, h), spa, 8 if base == "room" else 16, i))
    return jobs


def render_params(width, height, samples_per_axis, max_depth, roulette_start=3, seed=0,
                  estimator=_abi.ESTIMATOR_PATH):
    p = _abi.RenderParams()
    p.width, p.height = width, height
    p.samples_per_axis = samples_per_axis
    p.max_depth = max_depth
    p.roulette_start = roulette_start
    p.estimator = estimator
    p.seed = seed
    return p
