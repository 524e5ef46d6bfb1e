"""The C++ drop-in API (albedo_amd/host/albedo_dropin.hpp) builds against the
C ABI and, on a GPU, renders through it with the reference's error behaviour,
renderImage's accumulating semantics and RenderJob's default estimator; its
image matches the oracle on the same scene and random numbers."""
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_render")


def test_dropin_header_compiles_and_links():
    if not os.path.exists(BIN):
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(BIN)


def dropin_scene():
    """The scene tests/cpp/dropin_render.cpp builds, through the Python mirror."""
    from albedo_amd.scene import (CheckerTexture, ConductorMaterial, ConstantTexture, DielectricMaterial,
                                  DiffuseMaterial, MeshGeometry, MeshInstance, MixMaterial, PerlinPattern, PointLight,
                                  Scene, SkyLight, Sphere, Transform, TriangleMesh, FLAT)
    s = Scene()
    white = DiffuseMaterial(ConstantTexture((0.73, 0.73, 0.73)))
    s.add(Sphere((0.0, -1e5, 0.0), 1e5), white)
    s.add(Sphere((0.0, 6.0 + 1e5, 0.0), 1e5), white)
    s.add(Sphere((0.0, 0.0, -6.0 - 1e5), 1e5), white)
    s.add(Sphere((-3.0 - 1e5, 0.0, 0.0), 1e5), DiffuseMaterial(ConstantTexture((0.65, 0.05, 0.05))))
    s.add(Sphere((3.0 + 1e5, 0.0, 0.0), 1e5), DiffuseMaterial(ConstantTexture((0.12, 0.45, 0.15))))
    s.add(Sphere((-0.9, 1.0, -1.0), 1.0), DielectricMaterial(1.5, 0.0))
    s.add(Sphere((1.6, 0.8, -2.2), 0.8), ConductorMaterial((0.156, 0.424, 1.383), (3.602, 2.472, 1.915), 0.25))
    s.add(Sphere((-2.2, 0.5, -2.8), 0.5),
          MixMaterial(ConductorMaterial((0.155, 0.117, 0.138), (4.83, 3.12, 2.15), 0.0),
                      DiffuseMaterial(ConstantTexture((0.9, 0.9, 0.85))), PerlinPattern(3.0, 6.0), 0.5))
    quad = TriangleMesh([(-1, 0.01, 1), (1, 0.01, 1), (1, 0.01, -1), (-1, 0.01, -1)], [(0, 1, 2), (0, 2, 3)],
                        uv=[(-1, 1, 1, 1, 1, -1), (-1, 1, 1, -1, -1, -1)])
    s.add(MeshInstance(MeshGeometry(quad, FLAT), Transform.translation((0.5, 0.0, 1.0)) *
                       Transform.rotation((0, 1, 0), 30)),
          DiffuseMaterial(CheckerTexture((0.73, 0.73, 0.73), (0.3, 0.3, 0.3), 1.25)))
    s.add_area_light(Sphere((0.0, 5.6, -1.2), 1.1), (9.0, 7.6, 5.8))
    s.add_light(PointLight((0.8, 4.2, 3.2), (1.2, 2.0, 3.6)))
    s.set_sky(SkyLight.uniform((0.05, 0.06, 0.08)))
    return s


@pytest.mark.gpu
def test_dropin_render_runs_on_the_gpu(ctx, tmp_path):
    import oracle
    from albedo_amd import workloads as W
    from albedo_amd.scene import CameraSettings
    out = tmp_path / "dropin.f32"
    r = subprocess.run([BIN, "64", str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lines = dict(l.split(" ", 1) for l in r.stdout.strip().splitlines() if " " in l)
    assert lines["errors"] == "3" and lines["size_errors"] == "3"
    assert lines["accumulate_mismatches"] == "0" and lines["add_sample_ok"] == "1"
    assert lines["default_is_scatter_only"] == "1"
    # Triangle shapes, an added TriangleMesh and the identity instance: one image.
    assert lines["triangle_forms_equal"] == "1" and lines["shape_counts_ok"] == "1"
    assert lines["triangle_errors"] == "4" and lines["triangle_ok"] == "1"
    w, h = 64, 36
    img = np.fromfile(out, np.float32).reshape(h, w, 3)
    s = dropin_scene()
    cam = CameraSettings((0.0, 2.4, 7.8), (0.0, 1.9, -0.5), (0, 1, 0), 38.0, w / h)
    # The same scene through the Python mirror and ab_render: the descriptions
    # differ at most in the last bits of the instance's transform.
    py = ctx.scene(s).render(cam, w, h, 4, max_depth=8, roulette_start=3, seed=7)
    # The oracle's FP64 path estimator fed the same per-sample random numbers.
    ref, _ = oracle.OracleScene(s.to_desc()).render(cam.to_abi(), W.render_params(w, h, 4, 8, 3, 7), sample_mode=1)

    def rel(a, b):
        return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))

    print("drop-in vs python path %.2e, vs oracle %.4f" % (rel(img, py), rel(img, ref)))
    assert rel(img, py) < 1e-3
    assert rel(img, ref) < 0.03
