"""GPU tests of the device BLAS build (AB_BUILD_DEVICE_BLAS): a scene built with
it answers every query exactly as the host-built scene does (closest hits are
the minimum of (t, object, prim), which no tree changes), so hits and rendered
images are bit-identical; only the tree statistics differ."""
import ctypes

import numpy as np
import pytest

from albedo_amd import AlbedoError
from albedo_amd import workloads as W
from albedo_amd.scene import CameraSettings

pytestmark = pytest.mark.gpu


def rays_around(lo, hi, n, seed):
    rng = np.random.default_rng(seed)
    c, e = (lo + hi) / 2, (hi - lo)
    org = c + 1.5 * e * rng.uniform(-1, 1, (n, 3))
    tgt = c + 0.5 * e * rng.uniform(-1, 1, (n, 3))
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = org
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


def test_trefoil_hits_identical(ctx):
    scene, mesh = W.trefoil_scene()
    host, dev = ctx.scene(scene), ctx.scene(scene, device_blas=True)
    assert host.hash != dev.hash
    hi, di = host.info(), dev.info()
    assert hi["triangles"] == di["triangles"] == 32768 and di["blas_nodes"] > 0
    for rays in (W.coherent_rays(mesh, 512), W.incoherent_rays(mesh, 1 << 18), rays_around(*mesh.bounds(), 4096, 3)):
        assert np.array_equal(dev.trace_closest(rays), host.trace_closest(rays))
        assert np.array_equal(dev.trace_any(rays), host.trace_any(rays))


def test_instanced_million_triangle_scene_identical(ctx, icosphere):
    s, host, _ = icosphere
    dev = ctx.scene(s, device_blas=True)
    rays = rays_around(np.array([-7.0, 0.0, -7.0]), np.array([7.0, 2.0, 7.0]), 1 << 17, 5)
    hits = dev.trace_closest(rays)
    assert np.array_equal(hits, host.trace_closest(rays))
    assert len(np.unique(hits[:, 1])) > 60
    assert np.array_equal(dev.trace_any(rays), host.trace_any(rays))
    c = W.ICOSPHERE_CAMERA
    cam = CameraSettings(c.look_from, c.look_at, (0, 1, 0), c.vertical_fov_degrees, 96 / 54)
    a = dev.render(cam, 96, 54, 2, 8, 3, 6)
    b = host.render(cam, 96, 54, 2, 8, 3, 6)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_full_scene_render_identical(ctx):
    s = W.full_scene()
    host, dev = ctx.scene(s), ctx.scene(s, device_blas=True)
    c = W.ROOM_CAMERA
    cam = CameraSettings(c.look_from, c.look_at, (0, 1, 0), 38.0, 64 / 36)
    a = dev.render(cam, 64, 36, 3, 16, 3, 2)
    b = host.render(cam, 64, 36, 3, 16, 3, 2)
    assert np.array_equal(a.view(np.uint32), b.view(np.uint32))


def test_unknown_build_flag_rejected(ctx):
    s = W.room_scene()
    d = s.to_desc()
    d.desc.build_flags = 8  # the first bit no AB_BUILD_* flag uses
    h = ctypes.c_void_p()
    with pytest.raises(AlbedoError, match="build flags"):
        ctx.check(ctx.lib.ab_scene_create(ctx.handle, d.ptr(), ctypes.byref(h)))
