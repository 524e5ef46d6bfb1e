"""Host-side scene API mirroring the reference's C++ types.

Reference interfaces mirrored here (names, argument meaning, validation):
  Sphere                 geometry/sphere.hpp, sphere.cpp:21-41
  TriangleMesh           geometry/triangle_mesh.hpp, triangle_mesh.cpp
  MeshGeometry           geometry/mesh_geometry.hpp (Flat / Smooth shading)
  MeshInstance           geometry/mesh_instance.hpp
  Transform              core/transform.hpp:38-93, transform.cpp:93-225
  Constant/Checker/PerlinTexture, Diffuse/Conductor/Dielectric/MixMaterial
                         texture/*.hpp, material/*.hpp
  PointLight, SkyLight   light/point_light.hpp, light/sky_light.hpp
  Scene                  scene/scene.hpp:31-98 (add, addAreaLight, addLight, setSky)
  CameraSettings         camera/perspective_camera.hpp:16-29

A Scene flattens into an ab_scene_desc (include/albedo_gpu.h) for the library.
Geometry the device sees is FP32; everything here is kept in FP64 as the
reference keeps it, and rounded once when the description is made.
"""
import ctypes
import math
from dataclasses import dataclass, field
from typing import List, Optional, Sequence, Tuple, Union

import numpy as np

from . import _abi

Vec3 = Tuple[float, float, float]


# ---------------------------------------------------------------- transforms

class Transform:
    """Affine map with its exact inverse kept alongside (core/transform.hpp:38-93)."""

    def __init__(self, linear=None, offset=None, inv_linear=None, inv_offset=None):
        self.linear = np.eye(3) if linear is None else np.asarray(linear, dtype=np.float64)
        self.offset = np.zeros(3) if offset is None else np.asarray(offset, dtype=np.float64)
        self.inv_linear = np.eye(3) if inv_linear is None else np.asarray(inv_linear, dtype=np.float64)
        self.inv_offset = np.zeros(3) if inv_offset is None else np.asarray(inv_offset, dtype=np.float64)

    @staticmethod
    def translation(v: Sequence[float]) -> "Transform":
        v = np.asarray(v, dtype=np.float64)
        if not np.all(np.isfinite(v)):
            raise ValueError("transform translation component is not finite")
        return Transform(offset=v, inv_offset=-v)

    @staticmethod
    def scaling(f: Sequence[float]) -> "Transform":
        f = np.asarray(f, dtype=np.float64)
        with np.errstate(divide="ignore"):
            inv = 1.0 / f
        if not (np.all(np.isfinite(f)) and np.all(f != 0) and np.all(np.isfinite(inv))):
            raise ValueError("transform scale factor must be finite and non-zero with a finite reciprocal")
        return Transform(linear=np.diag(f), inv_linear=np.diag(inv))

    @staticmethod
    def rotation(axis: Sequence[float], degrees: float) -> "Transform":
        a = np.asarray(axis, dtype=np.float64)
        l2 = float(a @ a)
        if not (math.isfinite(degrees) and l2 > 0 and math.isfinite(l2) and l2 >= 2.2250738585072014e-308):
            raise ValueError("transform rotation axis has no usable direction")
        x, y, z = a / math.sqrt(l2)
        r = degrees * (math.pi / 180)
        c, s = math.cos(r), math.sin(r)
        t = 1 - c
        m = np.array([[t * x * x + c, t * x * y - s * z, t * x * z + s * y],
                      [t * x * y + s * z, t * y * y + c, t * y * z - s * x],
                      [t * x * z - s * y, t * y * z + s * x, t * z * z + c]])
        return Transform(linear=m, inv_linear=m.T.copy())

    def __mul__(self, first: "Transform") -> "Transform":
        """self applied after `first` (reference Transform::operator*)."""
        lin = self.linear @ first.linear
        off = self.linear @ first.offset + self.offset
        ilin = first.inv_linear @ self.inv_linear
        ioff = first.inv_linear @ self.inv_offset + first.inv_offset
        out = Transform(lin, off, ilin, ioff)
        if not all(np.all(np.isfinite(x)) for x in (lin, off, ilin, ioff)):
            raise ValueError("composed transform has a non-finite entry")
        return out

    def inverse(self) -> "Transform":
        return Transform(self.inv_linear, self.inv_offset, self.linear, self.offset)

    def apply_points(self, p: np.ndarray) -> np.ndarray:
        return p @ self.linear.T + self.offset

    def apply_box(self, lo, hi):
        """Arvo's transformed box (core/transform.cpp:118-132)."""
        lo = np.asarray(lo, np.float64)
        hi = np.asarray(hi, np.float64)
        a = self.linear * lo[None, :]
        b = self.linear * hi[None, :]
        return np.minimum(a, b).sum(1) + self.offset, np.maximum(a, b).sum(1) + self.offset

    def rows(self, inverse=False):
        lin, off = (self.inv_linear, self.inv_offset) if inverse else (self.linear, self.offset)
        return [float(v) for r in range(3) for v in (lin[r, 0], lin[r, 1], lin[r, 2], off[r])]


# ------------------------------------------------------------------ textures

@dataclass
class ConstantTexture:
    color: Vec3 = (0.0, 0.0, 0.0)


@dataclass
class CheckerTexture:
    even: Vec3 = (0.0, 0.0, 0.0)
    odd: Vec3 = (0.0, 0.0, 0.0)
    frequency: float = 1.0


@dataclass
class PerlinTexture:
    low: Vec3 = (0.0, 0.0, 0.0)
    high: Vec3 = (1.0, 1.0, 1.0)
    scale: float = 1.0
    veins: float = 0.0


Texture = Union[ConstantTexture, CheckerTexture, PerlinTexture]


# ----------------------------------------------------------------- materials

@dataclass
class DiffuseMaterial:
    albedo: Texture = field(default_factory=ConstantTexture)


@dataclass
class ConductorMaterial:
    eta: Vec3 = (1.0, 1.0, 1.0)
    k: Vec3 = (0.0, 0.0, 0.0)
    roughness: float = 0.0


@dataclass
class DielectricMaterial:
    ior: float = 1.5
    roughness: float = 0.0


@dataclass
class PerlinPattern:
    scale: float = 1.0
    veins: float = 0.0


@dataclass
class MixMaterial:
    first: Union[DiffuseMaterial, ConductorMaterial, DielectricMaterial]
    second: Union[DiffuseMaterial, ConductorMaterial, DielectricMaterial]
    pattern: PerlinPattern = field(default_factory=PerlinPattern)
    threshold: float = 0.5


Material = Union[DiffuseMaterial, ConductorMaterial, DielectricMaterial, MixMaterial]


# ------------------------------------------------------------------ geometry

@dataclass
class Sphere:
    center: Vec3
    radius: float

    def __post_init__(self):
        r = float(self.radius)
        if not (r > 0 and math.isfinite(r) and r * r >= 2.2250738585072014e-308 and math.isfinite(r * r)):
            raise ValueError("sphere radius must be positive and its square a normal number, got %r" % r)
        if not all(math.isfinite(float(c)) for c in self.center):
            raise ValueError("sphere center must be finite")


DEFAULT_TRIANGLE_UV = ((0.0, 0.0), (1.0, 0.0), (1.0, 1.0))


def forms_triangle(p0, p1, p2) -> bool:
    """True when Triangle(p0, p1, p2) would not raise: finite coordinates and an
    edge cross product whose squared length is a normal number (reference
    geometry/triangle.cpp:198-205)."""
    p = np.array([p0, p1, p2], dtype=np.float64)
    if not np.all(np.isfinite(p)):
        return False
    with np.errstate(over="ignore", invalid="ignore"):
        n = np.cross(p[1] - p[0], p[2] - p[0])
        l2 = float((n * n).sum())
    return math.isfinite(l2) and l2 >= 2.2250738585072014e-308


@dataclass
class Triangle:
    """A world-space triangle with per-corner texture coordinates (reference
    geometry/triangle.hpp:22-78). Traced as a one-face flat mesh under the
    identity map, which intersects it exactly as the same face of any mesh."""
    p0: Vec3
    p1: Vec3
    p2: Vec3
    uv: Tuple[Tuple[float, float], ...] = DEFAULT_TRIANGLE_UV

    def __post_init__(self):
        if not np.all(np.isfinite(np.array([self.p0, self.p1, self.p2], dtype=np.float64))):
            raise ValueError("triangle vertices must have finite coordinates")
        if not forms_triangle(self.p0, self.p1, self.p2):
            raise ValueError("triangle has no computable area")
        if len(self.uv) != 3 or not np.all(np.isfinite(np.array(self.uv, dtype=np.float64))):
            raise ValueError("triangle texture coordinates must be three finite pairs")

    def normal(self):
        p = np.array([self.p0, self.p1, self.p2], dtype=np.float64)
        n = np.cross(p[1] - p[0], p[2] - p[0])
        return tuple(n / np.sqrt((n * n).sum()))


class TriangleMesh:
    """Positions + triangle faces (+ optional per-face corner uv). Immutable once
    built (its arrays are read-only), so derived sets are computed once."""

    def __init__(self, positions, faces, uv=None):
        self.positions = np.array(positions, dtype=np.float32).reshape(-1, 3)
        self.faces = np.array(faces, dtype=np.uint32).reshape(-1, 3)
        self.uv = None if uv is None else np.array(uv, dtype=np.float32).reshape(-1, 6)
        if not np.all(np.isfinite(self.positions)):
            raise ValueError("mesh position has a non-finite coordinate")
        if self.faces.size and int(self.faces.max()) >= len(self.positions):
            raise IndexError("mesh face uses a position the mesh does not have")
        for a in (self.positions, self.faces, self.uv):
            if a is not None:
                a.flags.writeable = False
        self._kept = None
        self._used = None

    def kept_faces(self):
        """Faces whose triangle has a computable area (triangle_mesh.cpp:76-90)."""
        if self._kept is None:
            p = self.positions.astype(np.float64)
            a, b, c = p[self.faces[:, 0]], p[self.faces[:, 1]], p[self.faces[:, 2]]
            n = np.cross(b - a, c - a)
            l2 = (n * n).sum(1)
            self._kept = np.nonzero(np.isfinite(l2) & (l2 >= 2.2250738585072014e-308))[0]
            self._kept.flags.writeable = False
        return self._kept

    def used_positions(self):
        """Indices of the positions some kept face uses, ascending."""
        if self._used is None:
            self._used = np.unique(self.faces[self.kept_faces()].reshape(-1))
            self._used.flags.writeable = False
        return self._used

    def bounds(self):
        pts = self.positions[self.used_positions()].astype(np.float64)
        return pts.min(0), pts.max(0)


FLAT, SMOOTH = 0, 1


class MeshGeometry:
    """A mesh shared by instances, shaded flat or smooth."""

    def __init__(self, mesh: TriangleMesh, shading: int = FLAT):
        if len(mesh.kept_faces()) == 0:
            raise ValueError("mesh geometry needs at least one triangle with area")
        self.mesh = mesh
        self.shading = shading


@dataclass
class MeshInstance:
    geometry: MeshGeometry
    object_to_world: Transform


@dataclass
class PointLight:
    position: Vec3
    intensity: Vec3


@dataclass
class SkyLight:
    nadir: Vec3 = (0.0, 0.0, 0.0)
    zenith: Vec3 = (0.0, 0.0, 0.0)

    @staticmethod
    def uniform(c: Vec3) -> "SkyLight":
        return SkyLight(tuple(c), tuple(c))


@dataclass
class CameraSettings:
    look_from: Vec3 = (0.0, 0.0, 0.0)
    look_at: Vec3 = (0.0, 0.0, -1.0)
    up: Vec3 = (0.0, 1.0, 0.0)
    vertical_fov_degrees: float = 60.0
    aspect_ratio: float = 16.0 / 9.0

    def to_abi(self) -> _abi.Camera:
        c = _abi.Camera()
        c.look_from[:] = [float(v) for v in self.look_from]
        c.look_at[:] = [float(v) for v in self.look_at]
        c.up[:] = [float(v) for v in self.up]
        c.vfov_degrees = float(self.vertical_fov_degrees)
        c.aspect = float(self.aspect_ratio)
        return c


# --------------------------------------------------------------------- scene

class Scene:
    """Objects and lights; mirrors the reference Scene's builder calls."""

    def __init__(self):
        self.objects: List[tuple] = []  # (shape, material, light index or -1)
        self.lights: List[tuple] = []   # ("point", PointLight) | ("area", object) | ("sky", SkyLight)
        self.sky_index: Optional[int] = None
        self._mesh_counts = {}  # object index -> triangles of an added TriangleMesh

    def add(self, shape, material: Material):
        """Adds a Sphere, Triangle or MeshInstance as one object, or every
        triangle of a TriangleMesh with a computable area, in face order
        (reference scene/scene.cpp:25-60). The mesh's triangles become one flat
        mesh under the identity map: the same faces hit exactly as separate
        triangles would, and the (t, object, face) tie order is the same because
        the faces keep their order and stay together in the object order. A mesh
        with no such triangle adds nothing."""
        if isinstance(shape, (Sphere, Triangle, MeshInstance)):
            self.objects.append((shape, material, -1))
        elif isinstance(shape, TriangleMesh):
            if len(shape.kept_faces()):
                self.objects.append((MeshInstance(MeshGeometry(shape, FLAT), Transform()), material, -1))
                self._mesh_counts[len(self.objects) - 1] = len(shape.kept_faces())
        else:
            raise TypeError("Scene.add takes a Sphere, Triangle, MeshInstance or TriangleMesh")

    def add_area_light(self, sphere: Sphere, radiance: Vec3, material: Material = None):
        if material is None:
            material = DiffuseMaterial(ConstantTexture((0.0, 0.0, 0.0)))
        light = len(self.lights)
        self.lights.append(("area", len(self.objects), tuple(radiance)))
        self.objects.append((sphere, material, light))
        return light

    def add_light(self, light: PointLight) -> int:
        self.lights.append(("point", light))
        return len(self.lights) - 1

    def set_sky(self, sky: SkyLight):
        if self.sky_index is None:
            self.sky_index = len(self.lights)
            self.lights.append(("sky", sky))
        else:
            self.lights[self.sky_index] = ("sky", sky)

    def shape_count(self) -> int:
        """Objects as the reference counts them: one per shape, one per kept
        triangle of an added TriangleMesh."""
        return len(self.objects) - len(self._mesh_counts) + sum(self._mesh_counts.values())

    # ------------------------------------------------------------ flattening

    def to_desc(self) -> "SceneDescription":
        return SceneDescription(self)


def _tex(t: Texture) -> _abi.TextureDesc:
    d = _abi.TextureDesc()
    if isinstance(t, ConstantTexture):
        d.kind = _abi.TEX_CONSTANT
        d.a[:] = list(t.color)
    elif isinstance(t, CheckerTexture):
        d.kind = _abi.TEX_CHECKER
        d.a[:] = list(t.even)
        d.b[:] = list(t.odd)
        d.frequency = t.frequency
    elif isinstance(t, PerlinTexture):
        d.kind = _abi.TEX_PERLIN
        d.a[:] = list(t.low)
        d.b[:] = list(t.high)
        d.frequency = t.scale
        d.veins = t.veins
    else:
        raise TypeError("unknown texture %r" % (t,))
    return d


class SceneDescription:
    """ab_scene_desc for a Scene, holding every buffer it points into alive."""

    def __init__(self, scene: Scene):
        self._keep = []
        mats: List[_abi.MaterialDesc] = []
        mat_index = {}

        def basic(m) -> _abi.MaterialDesc:
            d = _abi.MaterialDesc()
            if isinstance(m, DiffuseMaterial):
                d.kind = _abi.MAT_DIFFUSE
                d.albedo = _tex(m.albedo)
            elif isinstance(m, ConductorMaterial):
                d.kind = _abi.MAT_CONDUCTOR
                d.eta[:] = list(m.eta)
                d.k[:] = list(m.k)
                d.roughness = m.roughness
            elif isinstance(m, DielectricMaterial):
                d.kind = _abi.MAT_DIELECTRIC
                d.ior = m.ior
                d.roughness = m.roughness
            else:
                raise TypeError("unknown material %r" % (m,))
            return d

        def material(m) -> int:
            key = id(m)
            if key in mat_index:
                return mat_index[key]
            if isinstance(m, MixMaterial):
                i1 = material(m.first)
                i2 = material(m.second)
                d = _abi.MaterialDesc()
                d.kind = _abi.MAT_MIX
                d.first, d.second = i1, i2
                d.scale = m.pattern.scale
                d.veins = m.pattern.veins
                d.threshold = m.threshold
            else:
                d = basic(m)
            mats.append(d)
            mat_index[key] = len(mats) - 1
            return mat_index[key]

        meshes: List[_abi.MeshDesc] = []
        mesh_index = {}

        def mesh(g: MeshGeometry) -> int:
            if id(g) in mesh_index:
                return mesh_index[id(g)]
            md = _abi.MeshDesc()
            pos = g.mesh.positions
            fac = g.mesh.faces
            self._keep += [pos, fac]
            md.positions = pos.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            md.num_vertices = len(pos)
            md.indices = fac.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
            md.num_faces = len(fac)
            if g.mesh.uv is not None:
                self._keep.append(g.mesh.uv)
                md.uv = g.mesh.uv.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            md.smooth = 1 if g.shading == SMOOTH else 0
            meshes.append(md)
            mesh_index[id(g)] = len(meshes) - 1
            return mesh_index[id(g)]

        def one_face(t: Triangle) -> int:
            pos = np.array([t.p0, t.p1, t.p2], dtype=np.float32)
            fac = np.array([0, 1, 2], dtype=np.uint32)
            uv = np.array(t.uv, dtype=np.float32).reshape(6)
            self._keep += [pos, fac, uv]
            md = _abi.MeshDesc()
            md.positions = pos.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            md.num_vertices = 3
            md.indices = fac.ctypes.data_as(ctypes.POINTER(ctypes.c_uint32))
            md.num_faces = 1
            md.uv = uv.ctypes.data_as(ctypes.POINTER(ctypes.c_float))
            md.smooth = 0
            meshes.append(md)
            return len(meshes) - 1

        objs: List[_abi.ObjectDesc] = []
        for shape, mat, light in scene.objects:
            o = _abi.ObjectDesc()
            o.material = material(mat)
            o.light = light
            if isinstance(shape, Sphere):
                o.kind = _abi.OBJ_SPHERE
                o.center[:] = [float(c) for c in shape.center]
                o.radius = float(shape.radius)
            elif isinstance(shape, Triangle):
                o.kind = _abi.OBJ_INSTANCE
                o.mesh = one_face(shape)
                o.to_world[:] = Transform().rows(False)
                o.to_object[:] = Transform().rows(True)
            else:
                o.kind = _abi.OBJ_INSTANCE
                o.mesh = mesh(shape.geometry)
                o.to_world[:] = shape.object_to_world.rows(False)
                o.to_object[:] = shape.object_to_world.rows(True)
            objs.append(o)

        lights: List[_abi.LightDesc] = []
        for entry in scene.lights:
            l = _abi.LightDesc()
            if entry[0] == "point":
                l.kind = _abi.LIGHT_POINT
                l.position[:] = [float(v) for v in entry[1].position]
                l.intensity[:] = list(entry[1].intensity)
            elif entry[0] == "area":
                l.kind = _abi.LIGHT_AREA
                l.object = entry[1]
                l.radiance[:] = list(entry[2])
            else:
                l.kind = _abi.LIGHT_SKY
                l.nadir[:] = list(entry[1].nadir)
                l.zenith[:] = list(entry[1].zenith)
            lights.append(l)

        def arr(T, items):
            a = (T * max(1, len(items)))(*items)
            self._keep.append(a)
            return a

        self.desc = _abi.SceneDesc()
        self.desc.meshes = ctypes.cast(arr(_abi.MeshDesc, meshes), ctypes.POINTER(_abi.MeshDesc))
        self.desc.num_meshes = len(meshes)
        self.desc.objects = ctypes.cast(arr(_abi.ObjectDesc, objs), ctypes.POINTER(_abi.ObjectDesc))
        self.desc.num_objects = len(objs)
        self.desc.materials = ctypes.cast(arr(_abi.MaterialDesc, mats), ctypes.POINTER(_abi.MaterialDesc))
        self.desc.num_materials = len(mats)
        self.desc.lights = ctypes.cast(arr(_abi.LightDesc, lights), ctypes.POINTER(_abi.LightDesc))
        self.desc.num_lights = len(lights)

    def ptr(self):
        return ctypes.byref(self.desc)


# ------------------------------------------------------- instance placement

def fitted(placement: Transform, geometry: MeshGeometry, size: float) -> Transform:
    """`placement`, then centred and scaled so the box's longest side is `size`
    (reference scene_file.cpp fitted())."""
    lo, hi = geometry.mesh.bounds()
    blo, bhi = placement.apply_box(lo, hi)
    ext = bhi - blo
    scale = size / float(ext.max())
    if not (math.isfinite(scale) and scale > 0):
        raise ValueError("the instance is too large or too small to fit")
    center = blo / 2 + bhi / 2
    return Transform.scaling((scale, scale, scale)) * Transform.translation(-center) * placement


def grounded(placement: Transform, geometry: MeshGeometry) -> Transform:
    """`placement`, then lifted so its lowest point is at y = 0 (scene_file.cpp grounded())."""
    m = geometry.mesh
    ys = placement.apply_points(m.positions[m.used_positions()].astype(np.float64))[:, 1]
    return Transform.translation((0.0, -float(ys.min()), 0.0)) * placement
