"""Ray queries over the four-wide BVH (AB_BUILD_WIDE_BVH) against the oracle.

The wide form collapses the host's binary SAH trees (TLAS and every BLAS) into
nodes of four children and traverses them with a stack of up to three entries a
visit. The answer of a closest-hit query is the lexicographic minimum of
(t, object, prim), which no tree shape changes, so the bar is the binary tree's:
object, prim and t bit-exact against the oracle's FP32 mirror, any-hit equal.
"""
import numpy as np
import pytest

import albedo_amd
import oracle
from albedo_amd import workloads as W

from test_trace_gpu import MISS, assert_bit_exact, room_rays

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def wide_trefoil(ctx):
    scene, mesh = W.trefoil_scene()
    g = ctx.scene(scene, wide_bvh=True)
    yield g, oracle.OracleScene(scene.to_desc()), mesh
    g.release()


def test_trefoil_sets_bit_exact(wide_trefoil):
    g, o, mesh = wide_trefoil
    for label, rays in [("coherent", W.coherent_rays(mesh, 1024)), ("incoherent", W.incoherent_rays(mesh, 1 << 20))]:
        hits = g.trace_closest(rays)
        ref = o.trace_closest(rays, fp32=True)
        assert (ref[1] != MISS).mean() > 0.25
        assert_bit_exact(hits, ref, "wide " + label)
    small = W.incoherent_rays(mesh, 4096)
    assert_bit_exact(g.trace_closest(small), o.trace_closest(small, fp32=True, brute=True), "wide brute")


def test_trefoil_any_hit_and_ranges(wide_trefoil):
    g, o, mesh = wide_trefoil
    rays = W.incoherent_rays(mesh, 1 << 18)
    rng = np.random.default_rng(15)
    rays[:, 7] = rng.uniform(0.5, 12.0, len(rays)).astype(np.float32)
    rays[: 1 << 14, 3] = rng.uniform(0.0, 4.0, 1 << 14).astype(np.float32)
    ref = o.trace_any(rays, fp32=True)
    assert 0.1 < ref.mean() < 0.9
    assert np.array_equal(g.trace_any(rays), ref)
    assert_bit_exact(g.trace_closest(rays), o.trace_closest(rays, fp32=True), "wide ranges")


@pytest.mark.parametrize("which", ["room", "full"])
def test_sphere_and_instance_scenes_bit_exact(ctx, which):
    scene = W.room_scene() if which == "room" else W.full_scene()
    g = ctx.scene(scene, wide_bvh=True)
    o = oracle.OracleScene(scene.to_desc())
    rays = room_rays(1 << 16, seed=11)
    assert_bit_exact(g.trace_closest(rays), o.trace_closest(rays, fp32=True), "wide " + which)
    assert np.array_equal(g.trace_any(rays), o.trace_any(rays, fp32=True))
    g.release()


def test_million_triangle_instanced_scene_equals_binary_tree(ctx, icosphere):
    # The deepest trees the library traces: 64 instances of a 1,310,720-triangle
    # BLAS. The binary-tree scene is checked against the oracle in test_trace_gpu.
    s, g2, _ = icosphere
    g = ctx.scene(s, wide_bvh=True)
    rng = np.random.default_rng(13)
    n = 1 << 16
    org = np.tile(np.array([0.0, 6.5, 15.0]), (n, 1))
    tgt = np.stack([rng.uniform(-6, 6, n), rng.uniform(0.0, 1.5, n), rng.uniform(-6, 6, n)], 1)
    d = tgt - org
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    hits = g.trace_closest(rays)
    assert np.array_equal(hits, g2.trace_closest(rays))
    assert len(np.unique(hits[:, 1])) > 30
    assert np.array_equal(g.trace_any(rays), g2.trace_any(rays))
    g.release()


def test_wide_and_device_blas_rejected_together(ctx):
    scene, _ = W.trefoil_scene()
    with pytest.raises(albedo_amd.AlbedoError, match="AB_BUILD_WIDE_BVH"):
        ctx.scene(scene, device_blas=True, wide_bvh=True)
