"""A scene read from a file feeds the GPU renderer (reference readScene ->
renderImage, apps/albedo/main.cpp:85-106): a sphere room in the shape of the
sphere-room benchmark scene plus a smooth OBJ mesh instance placed with
`scale`, `rotate`, `translate` and `ground`, written as a .scene file beside its
.obj, read by the drop-in readScene, rendered by renderImage, and compared with
the same scene built through the Python mirror -- rendered on the GPU and by the
oracle with the same random numbers. The CLI renders the same file once."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from albedo_amd import workloads as W
from albedo_amd.scene import (CameraSettings, CheckerTexture, ConductorMaterial, ConstantTexture, DielectricMaterial,
                              DiffuseMaterial, MeshGeometry, MeshInstance, PointLight, Scene, SkyLight, Sphere,
                              Transform, TriangleMesh, SMOOTH, grounded)

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "scene_check")
CLI = os.path.join(ROOT, "albedo_amd", "bin", "albedo")

# A tent (coordinates exact in FP32, so the FP64 reader and the FP32 mirror hold
# the same mesh) with texture coordinates for the checker.
TENT_OBJ = ("v 0 0 -1\nv 0 0 1\nv -1 -1 -1\nv -1 -1 1\nv 1 -1 -1\nv 1 -1 1\n"
            "vt 0 0\nvt 1 0\nvt 0 1\nvt 1 1\n"
            "f 3/1 4/2 2/4\nf 3/1 2/4 1/3\nf 1/3 2/4 5/1\nf 2/4 6/2 5/1\n")

SCENE = """# the sphere room with a tent
camera from 0 2.4 7.8 at 0 1.9 -0.5 fov 38
sky 0.05 0.06 0.08
material white diffuse color 0.73 0.73 0.73
material red diffuse color 0.65 0.05 0.05
material green diffuse color 0.12 0.45 0.15
material glass dielectric ior 1.5
material frosted dielectric ior 1.5 roughness 0.2
material gold conductor eta 0.156 0.424 1.383 k 3.602 2.472 1.915 roughness 0.25
material silver conductor eta 0.155 0.117 0.138 k 4.83 3.12 2.15
material floor checker even 0.73 0.73 0.73 odd 0.3 0.3 0.3 frequency 4
sphere 0 -100000 0 radius 100000 material white
sphere 0 100006 0 radius 100000 material white
sphere 0 0 -100006 radius 100000 material white
sphere -100003 0 0 radius 100000 material red
sphere 100003 0 0 radius 100000 material green
sphere -1.8 0.8 -2 radius 0.8 material glass
sphere 0 0.8 -3 radius 0.8 material frosted
sphere 1.8 0.8 -2 radius 0.8 material gold
sphere 1.2 0.8 0.2 radius 0.8 material silver
mesh tent tent.obj smooth
instance tent material floor scale 0.75 rotate 0 1 0 30 translate -0.5 0 1 ground
area_light 0 5.6 -1.2 radius 1.1 radiance 9 7.6 5.8
point_light 0.8 4.2 3.2 intensity 1.2 2 3.6
"""


def mirror():
    """The same scene through the Python mirror, statement by statement."""
    s = Scene()
    s.set_sky(SkyLight.uniform((0.05, 0.06, 0.08)))

    def diffuse(r, g, b):
        return DiffuseMaterial(ConstantTexture((r, g, b)))

    white = diffuse(0.73, 0.73, 0.73)
    for c, m in [((0, -1e5, 0), white), ((0, 6 + 1e5, 0), white), ((0, 0, -6 - 1e5), white),
                 ((-3 - 1e5, 0, 0), diffuse(0.65, 0.05, 0.05)), ((3 + 1e5, 0, 0), diffuse(0.12, 0.45, 0.15))]:
        s.add(Sphere(c, 1e5), m)
    s.add(Sphere((-1.8, 0.8, -2), 0.8), DielectricMaterial(1.5, 0.0))
    s.add(Sphere((0, 0.8, -3), 0.8), DielectricMaterial(1.5, 0.2))
    s.add(Sphere((1.8, 0.8, -2), 0.8), ConductorMaterial((0.156, 0.424, 1.383), (3.602, 2.472, 1.915), 0.25))
    s.add(Sphere((1.2, 0.8, 0.2), 0.8), ConductorMaterial((0.155, 0.117, 0.138), (4.83, 3.12, 2.15), 0.0))
    pos = [(0, 0, -1), (0, 0, 1), (-1, -1, -1), (-1, -1, 1), (1, -1, -1), (1, -1, 1)]
    uvs = {1: (0, 0), 2: (1, 0), 3: (0, 1), 4: (1, 1)}
    faces = [((3, 1), (4, 2), (2, 4)), ((3, 1), (2, 4), (1, 3)), ((1, 3), (2, 4), (5, 1)), ((2, 4), (6, 2), (5, 1))]
    tent = TriangleMesh(pos, [[p - 1 for p, _ in f] for f in faces],
                        uv=[[c for _, t in f for c in uvs[t]] for f in faces])
    geo = MeshGeometry(tent, SMOOTH)
    place = Transform.translation((-0.5, 0, 1)) * Transform.rotation((0, 1, 0), 30) * \
        Transform.scaling((0.75, 0.75, 0.75))
    s.add(MeshInstance(geo, grounded(place, geo)),
          DiffuseMaterial(CheckerTexture((0.73, 0.73, 0.73), (0.3, 0.3, 0.3), 4.0)))
    s.add_area_light(Sphere((0, 5.6, -1.2), 1.1), (9.0, 7.6, 5.8))
    s.add_light(PointLight((0.8, 4.2, 3.2), (1.2, 2.0, 3.6)))
    return s


def rel(a, b):
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


@pytest.fixture()
def scene_file(tmp_path):
    (tmp_path / "tent.obj").write_text(TENT_OBJ)
    path = tmp_path / "room.scene"
    path.write_text(SCENE)
    return path


def test_scene_read_from_a_file_renders_like_its_description(ctx, scene_file, tmp_path):
    w, h, spa, depth, seed = 64, 36, 3, 8, 5
    out = tmp_path / "img.f32"
    r = subprocess.run([CHECK, str(scene_file), "--render", str(w), str(h), str(spa), str(depth), str(seed), str(out)],
                       capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and r.stdout.startswith("rendered objects=11")
    img = np.fromfile(out, np.float32).reshape(h, w, 3)
    s = mirror()
    cam = CameraSettings((0, 2.4, 7.8), (0, 1.9, -0.5), (0, 1, 0), 38.0, w / h)
    py = ctx.scene(s).render(cam, w, h, spa, max_depth=depth, roulette_start=3, seed=seed)
    ref, _ = oracle.OracleScene(s.to_desc()).render(cam.to_abi(), W.render_params(w, h, spa, depth, 3, seed),
                                                    sample_mode=1)
    print("file scene vs python mirror %.2e, vs oracle %.4f" % (rel(img, py), rel(img, ref)))
    # The reader and the mirror build the instance's transform in different
    # operation orders (last-bit differences in FP64, then rounded to FP32).
    assert rel(img, py) < 1e-3
    assert rel(img, ref) < 0.03


def test_cli_renders_the_file(scene_file, tmp_path):
    ppm = tmp_path / "out.ppm"
    r = subprocess.run([CLI, str(scene_file), str(ppm), "32"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0 and "wrote" in r.stdout
    data = ppm.read_bytes()
    head = b"P6\n32 18\n255\n"
    assert data.startswith(head) and len(data) == len(head) + 32 * 18 * 3
    pixels = np.frombuffer(data[len(head):], np.uint8)
    assert 10 < pixels.mean() < 245
    bad = subprocess.run([CLI, str(scene_file), str(ppm), "8"], capture_output=True, text=True, timeout=60)
    assert bad.returncode != 0 and "invalid width" in bad.stderr
