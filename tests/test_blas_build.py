"""CPU check of the device BLAS build: its per-element steps (blas_core.hpp)
run on the CPU in kernel order, and the tree they produce is checked by
tests/cpp/blas_check.cpp (structure, boxes, levels, closest hits = brute force)."""
import os
import subprocess

import numpy as np
import pytest

from albedo_amd import workloads as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "blas_check")


@pytest.fixture(scope="module")
def checker():
    r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp"), CHECK], capture_output=True, text=True)
    assert r.returncode == 0, r.stdout + r.stderr
    return CHECK


def run(checker, tmp_path, positions, faces, rays, top=None):
    path = tmp_path / "mesh.bin"
    with open(path, "wb") as f:
        np.array([len(positions), len(faces)], np.uint32).tofile(f)
        np.asarray(positions, np.float32).tofile(f)
        np.asarray(faces, np.uint32).tofile(f)
    args = [checker, str(path), str(rays)] + ([] if top is None else [str(top)])
    r = subprocess.run(args, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    words = r.stdout.split()
    return {words[i]: int(words[i + 1]) for i in range(1, len(words) - 1, 2)}


def test_trefoil(checker, tmp_path):
    m = W.trefoil_mesh()
    out = run(checker, tmp_path, m.positions, m.faces, 400)
    assert out["n"] == 32768 and out["hits"] > 100


@pytest.mark.parametrize("top", [None, -1, 0, 6])
def test_displaced_icosphere_level_5(checker, tmp_path, top):
    # The device's split keys (None), plain Morton keys (-1), positions only (0).
    m = W.displaced_icosphere(level=5)
    out = run(checker, tmp_path, m.positions, m.faces, 400, top)
    assert out["n"] == 20480 and out["levels"] <= 32


def test_serving_mesh_full_size(checker, tmp_path):
    # The 1,310,720-triangle serving mesh: structure in full, 30 rays by brute force.
    m = W.displaced_icosphere(level=8)
    out = run(checker, tmp_path, m.positions, m.faces, 30)
    # 25 levels keep the serving scene (TLAS 9 levels) within the 32-entry stack.
    assert out["n"] == 1310720 and out["levels"] <= 25


def test_many_triangles_with_one_centroid(checker, tmp_path):
    tri = np.array([[0, 0, 0], [1, 0, 0], [0, 1, 0]], np.float32)
    pos = np.concatenate([tri] * 200 + [tri + np.float32(i) for i in range(1, 9)])
    out = run(checker, tmp_path, pos, np.arange(len(pos), dtype=np.uint32).reshape(-1, 3), 300)
    assert out["n"] == 208


@pytest.mark.parametrize("n", [5, 6, 9, 17])
def test_small_meshes(checker, tmp_path, n):
    pos = np.random.default_rng(n).normal(size=(3 * n, 3))
    out = run(checker, tmp_path, pos, np.arange(3 * n, dtype=np.uint32).reshape(-1, 3), 200)
    assert out["n"] == n and out["nodes"] >= 1
