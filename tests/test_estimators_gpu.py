"""GPU parity of the two other estimators (reference estimator/radiance_estimator.hpp:36-38):
NormalsEstimator (normals_estimator.cpp:15-40) and ScatterOnlyEstimator
(scatter_only_estimator.cpp:10-48), as modes of the wavefront path tracer.

Bar (written in each test):
  * normals view against the oracle fed the same jitter: within 1e-4 per channel
    on 99.5 % of pixels (the rest are silhouette pixels whose FP32 and FP64 hits
    differ);
  * scatter-only against the oracle's FP64 estimator on the same random numbers:
    relative RMSE below 2 %, as for the path estimator;
  * scatter-only and the path estimator are unbiased estimates of the same image
    in a scene without point lights: their 1024-spp images agree in mean within 2 %.
"""
import numpy as np
import pytest

import oracle
from albedo_amd import _abi, AlbedoError
from albedo_amd import workloads as W
from albedo_amd.scene import CameraSettings, ConstantTexture, DiffuseMaterial, Scene, SkyLight, Sphere

pytestmark = pytest.mark.gpu


def rel_rmse(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.sqrt(np.mean((a - b) ** 2)) / np.sqrt(np.mean(b ** 2)))


@pytest.fixture(scope="module")
def room(ctx):
    s = W.room_scene()
    return s, ctx.scene(s), oracle.OracleScene(s.to_desc())


def cam_for(w, h):
    c = W.ROOM_CAMERA
    return CameraSettings(c.look_from, c.look_at, (0, 1, 0), 38.0, w / h)


def test_normals_view_matches_oracle(room):
    s, g, o = room
    w, h, n = 64, 36, 2
    cam = cam_for(w, h)
    img, st = g.render(cam, w, h, n, max_depth=8, seed=3, stats=True, estimator=_abi.ESTIMATOR_NORMALS)
    ref, counts = o.render(cam.to_abi(), W.render_params(w, h, n, 8, 3, 3, _abi.ESTIMATOR_NORMALS), sample_mode=1)
    assert st["camera_rays"] == w * h * n * n == counts[0]
    assert st["extension_rays"] == 0 and st["shadow_rays"] == 0
    assert img.min() >= 0.0 and img.max() <= 1.0
    diff = np.abs(img.astype(np.float64) - ref).max(axis=2).ravel()
    print("normals view: max diff %.3g, 99.5%% quantile %.3g" % (diff.max(), np.quantile(diff, 0.995)))
    assert np.quantile(diff, 0.995) <= 1e-4
    assert diff.mean() <= 1e-3


def test_scatter_only_same_streams_match_fp64_estimator(room):
    s, g, o = room
    w, h, n, limit = 64, 36, 4, 8
    cam = cam_for(w, h)
    img, st = g.render(cam, w, h, n, max_depth=limit, seed=9, stats=True, estimator=_abi.ESTIMATOR_SCATTER_ONLY)
    ref, counts = o.render(cam.to_abi(), W.render_params(w, h, n, limit, 3, 9, _abi.ESTIMATOR_SCATTER_ONLY),
                           sample_mode=1)
    assert st["shadow_rays"] == 0 and counts[2] == 0
    err = rel_rmse(img, ref)
    print("scatter-only same-stream relative RMSE %.5f" % err)
    assert err < 0.02
    assert abs(int(st["extension_rays"]) - int(counts[1])) <= 0.01 * int(counts[1])


def test_scatter_only_and_path_estimator_agree_without_point_lights(ctx):
    # A diffuse box under a large area light and a sky: every light can be hit,
    # so both estimators converge to the same image.
    white = DiffuseMaterial(ConstantTexture((0.7, 0.7, 0.7)))
    red = DiffuseMaterial(ConstantTexture((0.6, 0.1, 0.1)))
    s = Scene()
    r = 1e5
    s.add(Sphere((0.0, -r, 0.0), r), white)
    s.add(Sphere((0.0, 0.0, -6.0 - r), r), white)
    s.add(Sphere((-3.0 - r, 0.0, 0.0), r), red)
    s.add(Sphere((0.0, 1.0, -2.0), 1.0), white)
    s.add_area_light(Sphere((0.0, 5.0, -1.0), 1.8), (3.0, 3.0, 3.0))
    s.set_sky(SkyLight.uniform((0.2, 0.25, 0.3)))
    g = ctx.scene(s)
    w, h, depth = 32, 18, 6
    cam = cam_for(w, h)
    scatter = g.render(cam, w, h, 32, max_depth=depth, seed=1, estimator=_abi.ESTIMATOR_SCATTER_ONLY)
    path = g.render(cam, w, h, 32, max_depth=depth, roulette_start=3, seed=2)
    ms, mp = float(scatter.mean()), float(path.mean())
    err = rel_rmse(scatter, path)
    print("scatter-only mean %.5f, path mean %.5f, relative RMSE %.4f" % (ms, mp, err))
    assert abs(ms - mp) <= 0.02 * mp
    assert err < 0.15


def test_unknown_estimator_is_rejected(room):
    s, g, o = room
    with pytest.raises(AlbedoError, match="AB_INVALID_ARGUMENT"):
        g.render(cam_for(8, 8), 8, 8, 1, estimator=7)
