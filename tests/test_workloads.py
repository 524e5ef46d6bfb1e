"""The serving workload's generators: the displaced icosphere's noise against the
reference's Perlin golden values and the oracle, and the icosphere's topology
against the closed forms of a subdivided icosahedron."""
import numpy as np
import pytest

import oracle
from albedo_amd import workloads as W

L = oracle.lib()


@pytest.mark.parametrize("p,expected", [((3.14, 42, 7), 0.13691995878400012),
                                        ((-0.3, -5.7, 2.25), 0.40750939192218755),
                                        ((-300.6, 17.2, -0.45), -0.11245507194489657)])
def test_vectorised_perlin_matches_reference_golden_values(p, expected):
    # tests/texture/perlin_noise_test.cpp:79-87
    assert abs(W.perlin_noise(np.array([p]))[0] - expected) < 1e-12


def test_vectorised_perlin_matches_oracle_on_random_points():
    pts = np.random.default_rng(5).uniform(-600.0, 600.0, (3000, 3))
    pts[:10] = np.round(pts[:10])  # lattice points, where noise is zero
    got = W.perlin_noise(pts)
    ref = np.array([L.or_perlin(*q) for q in pts])
    assert np.max(np.abs(got - ref)) < 1e-12
    assert np.all(got[:10] == 0.0)


def test_non_finite_points_give_zero_noise():
    got = W.perlin_noise(np.array([[np.nan, 0, 0], [np.inf, 1, 2], [0.5, 0.25, 0.125]]))
    assert got[0] == 0.0 and got[1] == 0.0
    assert got[2] == L.or_perlin(0.5, 0.25, 0.125)


def test_fbm_is_octave_sum_of_oracle_noise():
    p = np.array([[0.3, 1.7, -2.2], [11.5, -3.25, 7.0]])
    got = W.fbm(p)
    for row, g in zip(p, got):
        s, w, q = 0.0, 1.0, row.copy()
        for _ in range(6):
            s += w * L.or_perlin(*q)
            w *= 0.5
            q = 2 * q
        assert abs(g - s) < 1e-12


@pytest.mark.parametrize("level", [0, 1, 2, 4])
def test_icosphere_is_a_closed_outward_manifold(level):
    v, f = W.icosphere(level)
    assert len(f) == 20 * 4 ** level
    assert len(v) == 10 * 4 ** level + 2
    assert np.allclose(np.linalg.norm(v, axis=1), 1.0, atol=1e-15)
    e = np.sort(np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]]), 1)
    _, counts = np.unique(e, axis=0, return_counts=True)
    assert np.all(counts == 2) and len(counts) == 30 * 4 ** level  # Euler: V - E + F = 2
    p = v[f]
    n = np.cross(p[:, 1] - p[:, 0], p[:, 2] - p[:, 0])
    assert np.all(np.einsum("ij,ij->i", n, p.mean(1)) > 0)
    # Directed edges appear once each: consistent winding.
    d = np.concatenate([f[:, [0, 1]], f[:, [1, 2]], f[:, [2, 0]]])
    assert len(np.unique(d, axis=0)) == len(d)


def test_displaced_icosphere_radius_follows_fbm():
    m = W.displaced_icosphere(level=3, amplitude=0.15, seed=17)
    v, _ = W.icosphere(3)
    r = np.linalg.norm(m.positions.astype(np.float64), axis=1)
    rng = W.Pcg32(17, 41)
    offset = np.array([rng.uniform(0.0, 256.0) for _ in range(3)])
    expect = 1.0 + 0.15 * W.fbm(v * 2.0 + offset)
    assert np.max(np.abs(r - expect)) < 1e-6
    assert len(m.faces) == 1280


def test_icosphere_scene_has_64_instances_on_the_floor():
    s = W.icosphere_scene(level=2)
    d = s.to_desc()
    assert d.desc.num_objects == 66  # floor, 64 instances, the area light's sphere
    assert d.desc.num_meshes == 2    # the floor and one BLAS shared by every instance
    for shape, _, _ in s.objects[1:65]:
        p = shape.object_to_world.apply_points(shape.geometry.mesh.positions.astype(np.float64))
        assert abs(p[:, 1].min()) < 1e-9
        assert np.all(np.abs(p[:, [0, 2]]) <= 7.0)
