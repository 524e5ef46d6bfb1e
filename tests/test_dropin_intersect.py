"""Scene::intersect / Scene::occluded of the C++ drop-in (reference
scene/scene.cpp:107-165), run by tests/cpp/dropin_intersect.cpp on the GPU and
checked against the oracle's FP64 traversal of the same scene built through the
Python mirror: the same objects hit (numbered as the reference numbers them, each
face of an added TriangleMesh an object), t within FP32-ray rounding, points on
the ray, unit normals facing against it, the sphere normal and texture
coordinates of the closed forms, triangle texture coordinates from the oracle's
barycentrics, and occlusion exactly where a closest hit exists."""
import math
import os
import subprocess

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_intersect")
MISS = 0xFFFFFFFF


def ball():
    p = [(1.0, 0.0, 0.0), (-1.0, 0.0, 0.0), (0.0, 1.0, 0.0), (0.0, -1.0, 0.0), (0.0, 0.0, 1.0), (0.0, 0.0, -1.0)]
    f = [(0, 2, 4), (2, 1, 4), (1, 3, 4), (3, 0, 4), (2, 0, 5), (1, 2, 5), (3, 1, 5), (0, 3, 5)]
    out = []

    def mid(a, b):
        m = [(p[a][k] + p[b][k]) / 2 for k in range(3)]
        l = math.sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2])
        p.append(tuple(c / l for c in m))
        return len(p) - 1

    for t in f:
        a, b, c = mid(t[0], t[1]), mid(t[1], t[2]), mid(t[2], t[0])
        out += [(t[0], a, c), (a, t[1], b), (c, b, t[2]), (a, b, c)]
    return p, out


def mirror_scene():
    from albedo_amd.scene import (ConstantTexture, DiffuseMaterial, MeshGeometry, MeshInstance, SMOOTH, Scene,
                                  Sphere, Transform, Triangle, TriangleMesh)
    grey = DiffuseMaterial(ConstantTexture((0.6, 0.6, 0.6)))
    s = Scene()
    s.add(Sphere((0.0, -1e5, 0.0), 1e5), grey)
    s.add(Sphere((-1.5, 0.6, 0.0), 0.6), grey)
    s.add(Triangle((0.5, 0.0, -1.0), (1.5, 0.0, -1.0), (1.0, 1.2, -1.2)), grey)
    s.add(TriangleMesh([(-0.5, 0.0, 1.0), (0.5, 0.0, 1.0), (0.0, 1.0, 1.1), (1.0, 1.0, 1.1), (2.0, 2.0, 1.2)],
                       [(0, 1, 2), (1, 3, 2), (0, 0, 4)]), grey)
    p, f = ball()
    place = (Transform.translation((0.4, 0.9, 0.2)) * Transform.rotation((1, 1, 0), 35) *
             Transform.scaling((0.5, 0.7, 0.6)))
    s.add(MeshInstance(MeshGeometry(TriangleMesh(p, f), SMOOTH), place), grey)
    s.add(Triangle((-2.5, 0.0, 2.0), (-1.5, 0.0, 2.0), (-2.0, 1.0, 2.0), ((0.1, 0.2), (0.9, 0.1), (0.5, 0.8))),
          grey)
    return s


def read_records(path):
    rays, recs = [], []
    with open(path) as fh:
        for line in fh:
            w = line.split()
            if w[0] == "ray":
                rays.append([float(x) for x in w[1:]])
            elif w[0] == "hit":
                recs.append({"t": float(w[1]), "object": int(w[2]), "p": np.array([float(x) for x in w[3:6]]),
                             "n": np.array([float(x) for x in w[6:9]]), "s": np.array([float(x) for x in w[9:12]]),
                             "front": w[12] == "1", "uv": (float(w[13]), float(w[14])), "occ": w[15] == "1"})
            else:
                recs.append({"t": None, "occ": w[1] == "1"})
    return np.array(rays), recs


@pytest.mark.gpu
def test_intersect_and_occluded_match_the_oracle(tmp_path):
    import oracle
    out = tmp_path / "records.txt"
    r = subprocess.run([BIN, str(out)], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    rays, recs = read_records(out)
    assert len(rays) == 4000 and len(recs) == 4000
    s = mirror_scene()
    assert s.shape_count() == 7  # the added mesh counts its 2 faces, as the reference does
    o = oracle.OracleScene(s.to_desc())
    r32 = np.zeros((len(rays), 8), np.float32)
    r32[:, 0:3] = rays[:, 0:3]
    r32[:, 4:7] = rays[:, 3:6]
    r32[:, 7] = np.inf
    t64, obj, prim, b1, b2 = o.trace_closest(r32, fp32=False)
    base = [0, 1, 2, 3, 5, 6]  # reference object index of each description object

    def ref_index(k, pr):
        return base[k] + (pr if k == 3 else 0)

    hit = obj != MISS
    mine = np.array([rec["t"] is not None for rec in recs])
    assert hit.mean() > 0.5
    assert (mine != hit).sum() <= 0.005 * len(rays)
    same = 0
    kinds = {"sphere": 0, "floor": 0, "triangle": 0, "mesh": 0, "instance": 0, "uv_triangle": 0}
    for i, rec in enumerate(recs):
        # Occlusion over (0, inf) exactly where the drop-in found a closest hit.
        assert rec["occ"] == (rec["t"] is not None), i
        if rec["t"] is None or not hit[i]:
            continue
        if rec["object"] != ref_index(int(obj[i]), int(prim[i])):
            continue
        same += 1
        o_, d = rays[i, 0:3], rays[i, 3:6]
        assert abs(rec["t"] - t64[i]) <= 1e-5 * max(1.0, t64[i]), (i, rec["t"], t64[i])
        assert np.linalg.norm(rec["p"] - (o_ + rec["t"] * d)) <= 1e-6 * max(1.0, np.linalg.norm(rec["p"])), i
        assert abs(np.linalg.norm(rec["n"]) - 1) < 1e-12 and abs(np.linalg.norm(rec["s"]) - 1) < 1e-12, i
        assert np.dot(d, rec["n"]) <= 0, i  # faces against the ray
        k = int(obj[i])
        if k in (0, 1):
            c, rad = ((0.0, -1e5, 0.0), 1e5) if k == 0 else ((-1.5, 0.6, 0.0), 0.6)
            out_n = (rec["p"] - np.array(c)) / np.linalg.norm(rec["p"] - np.array(c))
            sign = 1 if rec["front"] else -1
            assert np.allclose(rec["n"], sign * out_n, atol=1e-9), i
            u = (math.atan2(-out_n[2], out_n[0]) + math.pi) / (2 * math.pi)
            v = math.acos(max(-1.0, min(1.0, -out_n[1]))) / math.pi
            assert abs(rec["uv"][0] - u) < 1e-9 and abs(rec["uv"][1] - v) < 1e-9, i
            kinds["floor" if k == 0 else "sphere"] += 1
        elif k in (2, 3):
            # Default corner uv (0,0), (1,0), (1,1): uv = (b1 + b2, b2).
            assert abs(rec["uv"][0] - (b1[i] + b2[i])) < 1e-5 and abs(rec["uv"][1] - b2[i]) < 1e-5, i
            assert np.allclose(rec["s"], rec["n"]), i  # flat
            kinds["triangle" if k == 2 else "mesh"] += 1
        elif k == 4:
            assert np.dot(rec["s"], rec["n"]) > 0.5, i  # the blended normal leans, on the same side
            kinds["instance"] += 1
        else:
            b0 = 1 - b1[i] - b2[i]
            uv = (b0 * 0.1 + b1[i] * 0.9 + b2[i] * 0.5, b0 * 0.2 + b1[i] * 0.1 + b2[i] * 0.8)
            assert abs(rec["uv"][0] - uv[0]) < 1e-5 and abs(rec["uv"][1] - uv[1]) < 1e-5, i
            kinds["uv_triangle"] += 1
    print("agreeing hits", same, kinds)
    assert same >= 0.995 * int((hit & mine).sum())
    assert all(v > 0 for v in kinds.values()), kinds
