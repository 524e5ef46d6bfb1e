"""The reference's path-estimator closed forms on the HIP path tracer.

Each case of tests/estimator/path_estimator_test.cpp that fixes a radiance along
one ray is rendered by ab_render through a 0.01-degree camera aimed along that
ray (a pixel spans about 1e-6 radians), so every sample follows the reference
test's ray up to that spread; the image mean is compared with the closed form
or with the other estimator, at the reference's tolerances. The same cases pin
the oracle (tests/test_oracle_pins.py).
"""
import math

import numpy as np
import pytest

import oracle
from albedo_amd import _abi
from albedo_amd.scene import (CameraSettings, MeshGeometry, MeshInstance, PointLight, Scene,
                              SkyLight, Sphere, Transform, SMOOTH)
from test_oracle_pins import (AT_GROUND, GLASS_RAY, GLOSSY_RAY, add_ground, f32, glossy_metal_scene, grey,
                              luminance, rough_glass_scene, tent_mesh)

pytestmark = pytest.mark.gpu


def ray_camera(origin, direction):
    d = np.asarray(direction, np.float64)
    up = (0.0, 0.0, -1.0) if abs(d[1]) > 0.99 else (0.0, 1.0, 0.0)
    return CameraSettings(tuple(origin), tuple(np.asarray(origin) + d), up, 0.01, 1.0)


def mean_along(ctx, scene, ray, w=8, spa=32, estimator=_abi.ESTIMATOR_PATH, max_depth=64, rr=3, seed=0):
    g = ctx.scene(scene)
    img = g.render(ray_camera(*ray), w, w, spa, max_depth=max_depth, roulette_start=rr, seed=seed,
                   estimator=estimator)
    return img.reshape(-1, 3).astype(np.float64).mean(0), img


def test_diffuse_ground_under_an_area_light_is_albedo_l_sin2(ctx):
    # path_estimator_test.cpp:148-166: rho L sin^2(theta_max) = 0.4 within 0.004
    # (65 536 samples here against the reference's 20 000).
    s = Scene()
    add_ground(s, 0.5)
    s.add_area_light(Sphere((0.0, 5.0, 0.0), 2.0), (5.0, 5.0, 5.0), grey(0.0))
    got, _ = mean_along(ctx, s, AT_GROUND)
    assert np.all(np.abs(got - 0.4) < 0.004), got


def test_point_light_on_the_smooth_ridge_is_exact(ctx):
    # path_estimator_test.cpp:94-111: every sample rho/pi I / 4 (FP32: rtol 1e-4).
    s = Scene()
    s.add(MeshInstance(MeshGeometry(tent_mesh(), SMOOTH), Transform()), grey(0.5))
    s.add_light(PointLight((0.0, 2.0, 0.0), (8.0, 6.0, 4.0)))
    _, img = mean_along(ctx, s, AT_GROUND, w=4, spa=4)
    want = np.array([8.0, 6.0, 4.0]) * (f32(0.5) / math.pi / 4)
    np.testing.assert_allclose(img.reshape(-1, 3), np.tile(want, (16, 1)), rtol=1e-4)


def test_point_light_behind_an_occluder_is_dark(ctx):
    # path_estimator_test.cpp:113-126: exactly black.
    s = Scene()
    add_ground(s, 0.5)
    s.add(Sphere((0.0, 1.5, 0.0), 0.2), grey(0.0))
    s.add_light(PointLight((0.0, 2.0, 0.0), (8.0, 8.0, 8.0)))
    d = np.array([-0.5, -1.0, 0.0]) / math.sqrt(1.25)
    _, img = mean_along(ctx, s, ((0.5, 1.0, 0.0), d), w=4, spa=8)
    assert np.all(img == 0.0)


def test_point_light_and_sky_add_up(ctx):
    # path_estimator_test.cpp:304-324: rho S + rho/pi I / d^2 within 0.01.
    s = Scene()
    add_ground(s, 0.5)
    s.add_light(PointLight((0.0, 2.0, 0.0), (8.0, 8.0, 8.0)))
    s.set_sky(SkyLight.uniform((0.4, 0.4, 0.4)))
    got, _ = mean_along(ctx, s, AT_GROUND)
    want = f32(0.4) * f32(0.5) + 8.0 * (f32(0.5) / math.pi / 4)
    assert np.all(np.abs(got - want) < 0.01), (got, want)


@pytest.mark.parametrize("make,ray", [(glossy_metal_scene, GLOSSY_RAY), (rough_glass_scene, GLASS_RAY)],
                         ids=["glossy_metal", "rough_glass"])
def test_path_estimator_agrees_with_scatter_only_and_the_oracle(ctx, make, ray):
    # path_estimator_test.cpp:208-233, 277-302: the HIP path estimator (light
    # sampling with power-heuristic MIS on a glossy lobe) against the HIP
    # scatter-only estimator and the oracle's path estimator, 102 400 samples
    # each, within 0.03 lum + 0.01.
    s = make()
    path, _ = mean_along(ctx, s, ray, spa=40)
    scatter, _ = mean_along(ctx, s, ray, spa=40, estimator=_abi.ESTIMATOR_SCATTER_ONLY, max_depth=32, seed=1)
    ref = oracle.OracleScene(s.to_desc()).radiance_est(*ray, 0, seed=17, count=100_000)
    tol = 0.03 * luminance(ref) + 0.01
    print("path %s scatter-only %s oracle %s tol %.4f" % (path, scatter, ref, tol))
    assert np.all(np.abs(path - scatter) < tol)
    assert np.all(np.abs(path - ref) < tol)
