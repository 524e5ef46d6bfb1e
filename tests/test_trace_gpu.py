"""GPU parity of the ray queries (ab_trace_closest / ab_trace_any) against the
oracle, on the trefoil closest-hit benchmark and on sphere / instance scenes.

Bar:
  * against the oracle's FP32 mirror: hit object, prim and t bit-exact, for every
    ray, at the benchmark's full size (2^24 coherent + 2^24 incoherent rays);
  * against the FP64 restatement: t within 128 ulp of the FP32 t (99.99% within
    16 ulp) wherever the ids match, and every id mismatch is a tie across an edge
    shared by the two triangles, with FP64 t's within 128 ulp;
  * any-hit equals the mirror's existence test.
"""
import math

import numpy as np
import pytest

import oracle
from albedo_amd import workloads as W
from albedo_amd import hits_t

pytestmark = pytest.mark.gpu

MISS = 0xFFFFFFFF


@pytest.fixture(scope="module")
def trefoil(ctx):
    scene, mesh = W.trefoil_scene()
    g = ctx.scene(scene)
    o = oracle.OracleScene(scene.to_desc())
    return g, o, mesh


def assert_bit_exact(hits, ref, label):
    t_ref, obj, prim, b1, b2 = ref
    gt = hits_t(hits)
    gobj, gprim = hits[:, 1], hits[:, 2]
    bad_obj = np.nonzero(gobj != obj)[0]
    bad_prim = np.nonzero(gprim != prim)[0]
    t32 = t_ref.astype(np.float32)
    hit = obj != MISS
    bad_t = np.nonzero(hit & (gt.view(np.uint32) != t32.view(np.uint32)))[0]
    assert len(bad_obj) == 0, "%s: %d object mismatches, first %s" % (label, len(bad_obj), bad_obj[:5])
    assert len(bad_prim) == 0, "%s: %d prim mismatches, first %s" % (label, len(bad_prim), bad_prim[:5])
    assert len(bad_t) == 0, "%s: %d t mismatches, first %s" % (label, len(bad_t), bad_t[:5])
    assert np.all(np.isinf(gt[~hit]))
    # barycentrics: unorm16 of the same FP32 values (one unit for rounding of the packing)
    gb1 = (hits[:, 3] & 0xFFFF).astype(np.int64)
    gb2 = (hits[:, 3] >> 16).astype(np.int64)
    tri = hit & (prim != MISS)
    eb1 = np.round(np.clip(b1, 0, 1) * 65535).astype(np.int64)
    eb2 = np.round(np.clip(b2, 0, 1) * 65535).astype(np.int64)
    assert np.all(np.abs(gb1[tri] - eb1[tri]) <= 1) and np.all(np.abs(gb2[tri] - eb2[tri]) <= 1)


def test_small_sets_match_brute_force_mirror(trefoil):
    g, o, mesh = trefoil
    for label, rays in [("coherent", W.coherent_rays(mesh, 64)), ("incoherent", W.incoherent_rays(mesh, 4096))]:
        hits = g.trace_closest(rays)
        ref = o.trace_closest(rays, fp32=True, brute=True)
        assert (ref[1] != MISS).mean() > 0.25
        assert_bit_exact(hits, ref, label)


@pytest.mark.parametrize("which", ["coherent", "incoherent"])
def test_full_benchmark_sets_bit_exact(trefoil, which):
    g, o, mesh = trefoil
    rays = W.coherent_rays(mesh, 4096) if which == "coherent" else W.incoherent_rays(mesh, 1 << 24)
    assert len(rays) == 1 << 24
    hits = g.trace_closest(rays)
    ref = o.trace_closest(rays, fp32=True)
    assert_bit_exact(hits, ref, which)


def adjacent(faces, kept, p, q):
    a = set(faces[kept[p]].tolist())
    b = set(faces[kept[q]].tolist())
    return len(a & b) >= 1


def test_against_fp64_within_stated_bound(trefoil):
    g, o, mesh = trefoil
    rays = np.concatenate([W.coherent_rays(mesh, 512), W.incoherent_rays(mesh, 1 << 18)])
    hits = g.trace_closest(rays)
    t64, obj64, prim64, _, _ = o.trace_closest(rays, fp32=False)
    gt = hits_t(hits)
    same = (hits[:, 1] == obj64) & (hits[:, 2] == prim64)
    both = same & (obj64 != MISS)
    ulp = np.abs(gt[both].astype(np.float64) - t64[both]) / np.spacing(gt[both])
    assert ulp.max() <= 128, ulp.max()
    assert np.quantile(ulp, 0.9999) <= 16
    kept = mesh.kept_faces()
    mism = np.nonzero(~same)[0]
    assert len(mism) <= 1e-4 * len(rays)
    for i in mism:
        # Both answers hit, on triangles sharing an edge or vertex, at t's within the bound.
        assert hits[i, 1] != MISS and obj64[i] != MISS, i
        assert adjacent(mesh.faces, kept, hits[i, 2], prim64[i]), i
        assert abs(float(gt[i]) - t64[i]) <= 128 * np.spacing(gt[i]), i


def test_any_hit_matches_mirror(trefoil):
    g, o, mesh = trefoil
    rays = W.incoherent_rays(mesh, 1 << 20)
    rng = np.random.default_rng(5)
    rays[:, 7] = rng.uniform(0.5, 12.0, len(rays)).astype(np.float32)  # shadow-ray style ranges
    rays[: 1 << 16, 3] = rng.uniform(0.0, 4.0, 1 << 16).astype(np.float32)
    occ = g.trace_any(rays)
    ref = o.trace_any(rays, fp32=True)
    assert 0.1 < ref.mean() < 0.9
    assert np.array_equal(occ, ref)
    small = rays[:4096]
    assert np.array_equal(g.trace_any(small), o.trace_any(small, fp32=True, brute=True))


def test_ranges_and_degenerate_rays(trefoil):
    g, o, mesh = trefoil
    rays = W.incoherent_rays(mesh, 8192)
    rng = np.random.default_rng(9)
    rays[:2048, 3] = rng.uniform(0, 8, 2048).astype(np.float32)          # tmin > 0
    rays[2048:4096, 7] = rng.uniform(0, 8, 2048).astype(np.float32)      # finite tmax
    rays[4096:4200, 4:7] = 0.0                                           # zero direction: never hits
    rays[4200:4400, 4] = 0.0                                             # axis-parallel components
    rays[4400:4600, 5:7] = 0.0
    rays[4400:4600, 4] = np.where(rng.uniform(size=200) < 0.5, -1.0, 1.0)
    rays[4600:4700, 0:3] = (mesh.positions[mesh.faces[:100, 0]] * 0.5 +
                            mesh.positions[mesh.faces[:100, 1]] * 0.5)  # origins on shared edges
    hits = g.trace_closest(rays)
    ref = o.trace_closest(rays, fp32=True, brute=True)
    assert_bit_exact(hits, ref, "edge cases")
    assert np.all(hits[4096:4200, 1] == MISS)
    assert g.trace_closest(rays[:0]).shape == (0, 4)


def test_rays_through_shared_edges_never_slip_through(trefoil):
    # Watertightness at benchmark scale (reference triangle_test.cpp:291-322):
    # rays that cross the surface at a point of a shared edge, coming in along
    # the face normal from outside, must hit something. (Rays that only graze
    # a silhouette edge may miss in any precision, so they are not used.)
    g, o, mesh = trefoil
    f = mesh.faces[mesh.kept_faces()]
    p = mesh.positions.astype(np.float64)
    rng = np.random.default_rng(2)
    sel = rng.integers(0, len(f), 20000)
    s = rng.uniform(0.01, 0.99, 20000)[:, None]
    tgt = p[f[sel, 0]] * (1 - s) + p[f[sel, 1]] * s
    n = np.cross(p[f[sel, 1]] - p[f[sel, 0]], p[f[sel, 2]] - p[f[sel, 0]])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    org = tgt + 0.2 * n
    d = tgt - org
    rays = np.zeros((20000, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    hits = g.trace_closest(rays)
    assert np.all(hits[:, 1] != MISS)
    assert_bit_exact(hits, o.trace_closest(rays, fp32=True, brute=True), "edges")


def room_rays(n, seed=4):
    rng = np.random.default_rng(seed)
    org = np.stack([rng.uniform(-2.9, 2.9, n), rng.uniform(0.05, 5.9, n), rng.uniform(-5.9, 7.0, n)], 1)
    d = rng.normal(size=(n, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    r = np.zeros((n, 8), np.float32)
    r[:, 0:3] = org
    r[:, 4:7] = d
    r[:, 7] = np.inf
    return r


def test_sphere_room_bit_exact_and_precise(ctx):
    scene = W.room_scene()
    g = ctx.scene(scene)
    o = oracle.OracleScene(scene.to_desc())
    rays = room_rays(1 << 16)
    hits = g.trace_closest(rays)
    assert_bit_exact(hits, o.trace_closest(rays, fp32=True, brute=True), "room")
    t64, obj64, _, _, _ = o.trace_closest(rays, fp32=False)
    assert np.array_equal(hits[:, 1], obj64)
    gt = hits_t(hits)
    hit = obj64 != MISS
    # FP64 sphere arithmetic: t is the FP64 root rounded once (radius-1e5 walls included).
    assert np.all(np.abs(gt[hit] - t64[hit]) <= 0.5 * np.spacing(gt[hit]))
    occ = g.trace_any(rays)
    assert np.array_equal(occ, o.trace_any(rays, fp32=True, brute=True))


def test_two_level_instances_bit_exact(ctx):
    scene = W.full_scene()
    g = ctx.scene(scene)
    o = oracle.OracleScene(scene.to_desc())
    rays = room_rays(1 << 17, seed=8)
    hits = g.trace_closest(rays)
    assert_bit_exact(hits, o.trace_closest(rays, fp32=True), "instances")
    small = rays[:2048]
    assert_bit_exact(g.trace_closest(small), o.trace_closest(small, fp32=True, brute=True), "instances brute")
    assert len(np.unique(hits[:, 1])) > 8  # knots, floor, walls, spheres all appear
    occ = g.trace_any(rays)
    assert np.array_equal(occ, o.trace_any(rays, fp32=True))


def test_device_pointer_entry_point_and_streams(ctx, trefoil):
    torch = pytest.importorskip("torch")
    g, o, mesh = trefoil
    rays = W.incoherent_rays(mesh, 1 << 16)
    d_rays = torch.from_numpy(rays).cuda()
    d_hits = torch.zeros((len(rays), 4), dtype=torch.int32, device="cuda")
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        g.trace_closest_device(d_rays.data_ptr(), d_hits.data_ptr(), len(rays), s.cuda_stream)
    s.synchronize()
    host = d_hits.cpu().numpy().view(np.uint32)
    assert np.array_equal(host, g.trace_closest(rays))


def test_million_triangle_instanced_scene_bit_exact(icosphere):
    # 64 instances of a 1,310,720-triangle BLAS: the deepest stacks the library
    # traces (TLAS depth 9 + BLAS depth 25, the 64-entry stack class).
    s, g, o = icosphere
    rng = np.random.default_rng(12)
    n = 1 << 17
    org = np.stack([rng.uniform(-7, 7, n), rng.uniform(0.05, 4.0, n), rng.uniform(-7, 7, n)], 1)
    org[: n // 2] = (0.0, 6.5, 15.0)  # half from the serving camera, aimed at the grid
    tgt = np.stack([rng.uniform(-6, 6, n), rng.uniform(0.0, 1.5, n), rng.uniform(-6, 6, n)], 1)
    d = np.where(np.arange(n)[:, None] < n // 2, tgt - org, rng.normal(size=(n, 3)))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    rays = np.zeros((n, 8), np.float32)
    rays[:, 0:3] = org
    rays[:, 4:7] = d
    rays[:, 7] = np.inf
    hits = g.trace_closest(rays)
    assert_bit_exact(hits, o.trace_closest(rays, fp32=True), "icosphere instances")
    assert len(np.unique(hits[:, 1])) > 60  # the floor and nearly every instance
    occ = g.trace_any(rays)
    assert np.array_equal(occ, o.trace_any(rays, fp32=True))
