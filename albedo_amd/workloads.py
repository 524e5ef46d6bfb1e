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


# ------------------------------------------------------- displaced icosphere

# Ken Perlin's published permutation ("Improving Noise", SIGGRAPH 2002), the table
# the reference's noise uses (texture/perlin_noise.cpp:18-32).
PERLIN_PERMUTATION = np.array([
    151, 160, 137, 91, 90, 15, 131, 13, 201, 95, 96, 53, 194, 233, 7, 225, 140, 36, 103, 30, 69,
    142, 8, 99, 37, 240, 21, 10, 23, 190, 6, 148, 247, 120, 234, 75, 0, 26, 197, 62, 94, 252,
    219, 203, 117, 35, 11, 32, 57, 177, 33, 88, 237, 149, 56, 87, 174, 20, 125, 136, 171, 168,
    68, 175, 74, 165, 71, 134, 139, 48, 27, 166, 77, 146, 158, 231, 83, 111, 229, 122, 60, 211,
    133, 230, 220, 105, 92, 41, 55, 46, 245, 40, 244, 102, 143, 54, 65, 25, 63, 161, 1, 216, 80,
    73, 209, 76, 132, 187, 208, 89, 18, 169, 200, 196, 135, 130, 116, 188, 159, 86, 164, 100,
    109, 198, 173, 186, 3, 64, 52, 217, 226, 250, 124, 123, 5, 202, 38, 147, 118, 126, 255, 82,
    85, 212, 207, 206, 59, 227, 47, 16, 58, 17, 182, 189, 28, 42, 223, 183, 170, 213, 119, 248,
    152, 2, 44, 154, 163, 70, 221, 153, 101, 155, 167, 43, 172, 9, 129, 22, 39, 253, 19, 98,
    108, 110, 79, 113, 224, 232, 178, 185, 112, 104, 218, 246, 97, 228, 251, 34, 242, 193, 238,
    210, 144, 12, 191, 179, 162, 241, 81, 51, 145, 235, 249, 14, 239, 107, 49, 192, 214, 31,
    181, 199, 106, 157, 184, 84, 204, 176, 115, 121, 50, 45, 127, 4, 150, 254, 138, 236, 205,
    93, 222, 114, 67, 29, 24, 72, 243, 141, 128, 195, 78, 66, 215, 61, 156, 180], dtype=np.int64)


def perlin_noise(p):
    """Improved Perlin noise of points p [n, 3] (FP64), the reference's
    PerlinNoise::noise (texture/perlin_noise.cpp:70-145) vectorised: lattice cell
    wrapped to 0..255, quintic fade, the 12-gradient table by hash & 15."""
    p = np.asarray(p, dtype=np.float64).reshape(-1, 3)
    out = np.zeros(len(p))
    ok = np.all(np.isfinite(p), axis=1)
    q = p[ok]
    f = np.floor(q)
    idx = np.mod(f, 256.0).astype(np.int64)
    x, y, z = (q - f).T
    xi, yi, zi = idx.T
    perm = PERLIN_PERMUTATION

    def h(i):
        return perm[i % 256]

    def fade(t):
        return t * t * t * (t * (t * 6 - 15) + 10)

    def grad(hash_, gx, gy, gz):
        hh = hash_ & 15
        u = np.where(hh < 8, gx, gy)
        v = np.where(hh < 4, gy, np.where((hh == 12) | (hh == 14), gx, gz))
        return np.where(hh & 1, -u, u) + np.where(hh & 2, -v, v)

    def lerp(t, a, b):
        return a + t * (b - a)

    u, v, w = fade(x), fade(y), fade(z)
    a = h(xi) + yi
    aa, ab = h(a) + zi, h(a + 1) + zi
    b = h(xi + 1) + yi
    ba, bb = h(b) + zi, h(b + 1) + zi
    out[ok] = lerp(w,
                   lerp(v, lerp(u, grad(h(aa), x, y, z), grad(h(ba), x - 1, y, z)),
                        lerp(u, grad(h(ab), x, y - 1, z), grad(h(bb), x - 1, y - 1, z))),
                   lerp(v, lerp(u, grad(h(aa + 1), x, y, z - 1), grad(h(ba + 1), x - 1, y, z - 1)),
                        lerp(u, grad(h(ab + 1), x, y - 1, z - 1), grad(h(bb + 1), x - 1, y - 1, z - 1))))
    return out


def fbm(p, octaves=6, lacunarity=2.0, gain=0.5):
    """Fractional Brownian motion: sum over octaves o of gain^o * noise(lacunarity^o * p)."""
    p = np.asarray(p, dtype=np.float64).reshape(-1, 3)
    total = np.zeros(len(p))
    weight, scale = 1.0, 1.0
    for _ in range(octaves):
        total += weight * perlin_noise(p * scale)
        weight *= gain
        scale *= lacunarity
    return total


def icosphere(level):
    """Unit icosphere: the icosahedron with each face split into four `level` times,
    new vertices pushed onto the sphere. 20 * 4^level faces, 10 * 4^level + 2
    vertices, faces wound counter-clockwise seen from outside. FP64 positions."""
    t = (1.0 + math.sqrt(5.0)) / 2.0
    v = np.array([[-1, t, 0], [1, t, 0], [-1, -t, 0], [1, -t, 0], [0, -1, t], [0, 1, t],
                  [0, -1, -t], [0, 1, -t], [t, 0, -1], [t, 0, 1], [-t, 0, -1], [-t, 0, 1]], dtype=np.float64)
    v /= np.linalg.norm(v, axis=1, keepdims=True)
    f = np.array([[0, 11, 5], [0, 5, 1], [0, 1, 7], [0, 7, 10], [0, 10, 11], [1, 5, 9], [5, 11, 4],
                  [11, 10, 2], [10, 7, 6], [7, 1, 8], [3, 9, 4], [3, 4, 2], [3, 2, 6], [3, 6, 8],
                  [3, 8, 9], [4, 9, 5], [2, 4, 11], [6, 2, 10], [8, 6, 7], [9, 8, 1]], dtype=np.int64)
    for _ in range(level):
        nv = len(v)
        e = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
        key = np.minimum(e[:, 0], e[:, 1]) * nv + np.maximum(e[:, 0], e[:, 1])
        uniq, inv = np.unique(key, return_inverse=True)
        lo, hi = uniq // nv, uniq % nv
        mid = v[lo] + v[hi]
        mid /= np.linalg.norm(mid, axis=1, keepdims=True)
        v = np.concatenate([v, mid])
        m = (inv.reshape(3, -1) + nv).T  # midpoints of edges (0,1), (1,2), (2,0)
        a, b, c = f[:, 0], f[:, 1], f[:, 2]
        m01, m12, m20 = m[:, 0], m[:, 1], m[:, 2]
        f = np.concatenate([np.stack([a, m01, m20], 1), np.stack([b, m12, m01], 1),
                            np.stack([c, m20, m12], 1), np.stack([m01, m12, m20], 1)])
    return v, f


def displaced_icosphere(level=8, amplitude=0.15, frequency=2.0, seed=17):
    """The serving scene's mesh: an icosphere of `level` (level 8: 1,310,720
    triangles) whose vertices move along the radius to 1 + amplitude * fbm, with
    6-octave fBm of the reference's Perlin noise evaluated at frequency * p + offset.
    The reference's permutation is fixed, so the seed picks the offset (PCG32 stream 41)."""
    v, f = icosphere(level)
    rng = Pcg32(seed, 41)
    offset = np.array([rng.uniform(0.0, 256.0) for _ in range(3)])
    r = 1.0 + amplitude * fbm(v * frequency + offset)
    return TriangleMesh((v * r[:, None]).astype(np.float32), f.astype(np.uint32))


ICOSPHERE_CAMERA = CameraSettings((0.0, 6.5, 15.0), (0.0, 0.4, 0.0), (0.0, 1.0, 0.0), 40.0, 16.0 / 9.0)


def icosphere_scene(seed=17, level=8, mesh=None):
    """64 instances of the displaced icosphere (one shared BLAS) on an 8 x 8 grid over
    a checker floor (even 0.73, odd 0.3, frequency 1.25): each fit to a seeded size
    in [0.9, 1.3], turned about a seeded axis, grounded on y = 0 and jittered inside
    its 1.5-unit cell, with a material drawn from the five room kinds plus white and
    red diffuse. A spherical area light overhead, a point light and a blue-grey sky."""
    mesh = mesh if mesh is not None else displaced_icosphere(level, seed=seed)
    geo = MeshGeometry(mesh, SMOOTH)
    m = room_materials()
    kinds = FIVE_KINDS + ("white", "red")
    s = Scene()
    floor = MeshGeometry(checker_floor(-7.0, 7.0, -7.0, 7.0), FLAT)
    s.add(MeshInstance(floor, Transform()),
          DiffuseMaterial(CheckerTexture((0.73, 0.73, 0.73), (0.3, 0.3, 0.3), 1.25)))
    rng = Pcg32(seed, 43)
    for i in range(64):
        gx, gz = i % 8, i // 8
        size = rng.uniform(0.9, 1.3)
        axis = np.array([rng.uniform(-1.0, 1.0), 1.0, rng.uniform(-1.0, 1.0)])
        angle = rng.uniform(0.0, 360.0)
        kind = kinds[min(int(rng.real() * len(kinds)), len(kinds) - 1)]
        jx, jz = rng.uniform(-0.15, 0.15), rng.uniform(-0.15, 0.15)
        place = fitted(Transform(), geo, size)
        place = Transform.rotation(axis / np.linalg.norm(axis), angle) * place
        place = grounded(place, geo)
        place = Transform.translation(((gx - 3.5) * 1.5 + jx, 0.0, (gz - 3.5) * 1.5 + jz)) * place
        s.add(MeshInstance(geo, place), m[kind])
    s.add_area_light(Sphere((0.0, 9.0, 2.0), 1.5), (12.0, 11.0, 9.5))
    s.add_light(PointLight((-5.0, 6.0, 6.0), (6.0, 6.0, 7.0)))
    s.set_sky(SkyLight.uniform((0.08, 0.09, 0.12)))
    return s


def serving_library(seed=17, level=8, knot=None):
    """The 8 scenes the serving workload keeps resident: the displaced-icosphere
    scene, three sphere-room variants and four full-scene variants holding 5..200
    random spheres. Returns [(name, scene, camera settings, max_depth)]."""
    knot = knot if knot is not None else trefoil_mesh()
    lib = [("icosphere", icosphere_scene(seed, level), ICOSPHERE_CAMERA, 8)]
    rng = Pcg32(seed, 51)
    for i in range(7):
        base = "room" if i < 3 else "full"
        n_sph = 5 + int(rng.real() * 196)
        lib.append(("%s+%d" % (base, n_sph), sphere_variant_scene(2000 + i, n_sph, base, knot), ROOM_CAMERA,
                    8 if base == "room" else 16))
    return lib


def serving_jobs(n_jobs, n_scenes=8, seed=23, resolutions=RESOLUTIONS, spas=SPAS):
    """Jobs for the serving workload: scene index uniform over the library,
    resolution uniform over `resolutions` (256x144..1920x1080), samples per axis
    uniform over `spas` (4..256 spp, perfect squares as stratification needs).
    Returns [(scene index, width, height, samples per axis, seed)]."""
    rng = Pcg32(seed, 53)
    jobs = []
    for i in range(n_jobs):
        k = min(int(rng.real() * n_scenes), n_scenes - 1)
        w, h = resolutions[min(int(rng.real() * len(resolutions)), len(resolutions) - 1)]
        spa = spas[min(int(rng.real() * len(spas)), len(spas) - 1)]
        jobs.append((k, w, h, spa, 100 + i))
    return jobs


def poisson_arrivals(n, rate_hz, seed=29):
    """n arrival times (seconds from 0) of a Poisson process of rate `rate_hz`:
    exponential gaps by inversion of PCG32 stream 55."""
    u = pcg32_reals(seed, 55, n)
    return np.cumsum(-np.log1p(-u) / rate_hz)


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
