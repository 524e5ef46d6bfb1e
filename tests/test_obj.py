"""The drop-in OBJ reader (albedo_amd/host/scene_file.hpp readObj) against the
reference's own OBJ tests (tests/geometry/obj_test.cpp), case by case: what it
reads, the line each error names, its wording, the line-length rules, limits,
and stream failures. CPU only."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "obj_check")
TRI = "v 0 0 0\nv 1 0 0\nv 0 1 0\nf 1 2 3\n"
SQUARE = "v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\n"
THREE = "v 0 0 0\nv 1 0 0\nv 0 1 0\n"


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CHECK):
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def obj(text, *args):
    r = subprocess.run([CHECK] + [str(a) for a in args], input=text.encode("latin-1"), capture_output=True,
                       timeout=60)
    out = r.stdout.decode("latin-1").strip()
    kind, _, rest = out.partition(" ")
    if kind == "ok":
        f = dict(kv.split("=", 1) for kv in rest.split())
        d = {k: int(f[k]) for k in ("positions", "faces", "uv", "uvfaces", "triangles")}
        d["f"] = [tuple(int(x) for x in t.split(",")) for t in f["f"].split(";") if t]
        d["tf"] = [tuple(int(x) for x in t.split(",")) for t in f["tf"].split(";") if t]
        d["p"] = [tuple(float(x) for x in t.split(",")) for t in f["p"].split(";") if t]
        d["t"] = [tuple(float(x) for x in t.split(",")) for t in f["t"].split(";") if t]
        return {"ok": True, **d}
    if kind == "error":
        line, _, msg = rest.partition(" ")
        return {"ok": False, "kind": "parse", "line": int(line.split("=")[1]), "message": msg}
    return {"ok": False, "kind": kind, "line": None, "message": rest}


def line_of(text, *args):
    r = obj(text, *args)
    return 0 if r["ok"] else r["line"]


# obj_test.cpp:49-105
def test_reads_a_triangle_and_empty_input():
    m = obj(TRI)
    assert m["positions"] == 3 and m["p"][1] == (1.0, 0.0, 0.0) and m["f"] == [(0, 1, 2)]
    e = obj("")
    assert e["ok"] and e["positions"] == 0 and e["faces"] == 0


def test_ignores_comments_blank_lines_and_other_statements():
    m = obj("# exported by a tool\n\nmtllib scene.mtl\no Thing\ng group\ns 1\nusemtl red\n" + THREE +
            "vn 0 0 1\nvt 0.5 0.5\nl 1 2\nf 1 2 3\n")
    assert m["positions"] == 3 and m["faces"] == 1
    assert obj("v 0 0 0 # origin\nv 1 0 0\nv 0 1 0\nf 1 2 3 # the only face\n")["faces"] == 1


def test_every_face_vertex_form():
    m = obj(THREE + "vt 0 0\nvn 0 0 1\nf 1/1 2/1 3/1\nf 1//1 2//1 3//1\nf 1/1/1 2/1/1 3/1/1\n")
    assert m["f"] == [(0, 1, 2)] * 3


def test_negative_indices_and_fans():
    assert obj("v 9 9 9\n" + THREE + "f -3 -2 -1\n")["f"] == [(1, 2, 3)]
    assert obj(SQUARE + "v -1 1 0\nf 1 2 3 4 5\n")["f"] == [(0, 1, 2), (0, 2, 3), (0, 3, 4)]


# obj_test.cpp:134-163
def test_line_endings_weights_signs_and_zero_area():
    m = obj("v\t0 0 0\r\nv 1\t0 0\r\n  v 0 1 0\r\nf 1 2 3")
    assert m["positions"] == 3 and m["faces"] == 1
    m = obj("v 0 0 0 1\nv 1 0 0 0.5 0.5 0.5\nv 0 1 0\nf 1 2 3\n")
    assert m["positions"] == 3 and m["p"][1] == (1.0, 0.0, 0.0)
    assert obj("v +1.5e0 -2E-1 .25\n")["p"] == [(1.5, -0.2, 0.25)]
    m = obj("v 0 0 0\nv 1 0 0\nf 1 1 2\n")
    assert m["faces"] == 1 and m["triangles"] == 0  # kept for the mesh to skip


# obj_test.cpp:165-242
def test_errors_name_the_line():
    r = obj("v 0 0 0\nv 1 0\n")
    assert r["line"] == 2 and "line 2" in r["message"]


@pytest.mark.parametrize("text,line", [
    ("v 0 0 0\nv 1 0\n", 2), ("v\n", 1),
    ("v 1.0abc 0 0\n", 1), ("v 1..2 0 0\n", 1), ("v 0 0 0\nv - 0 0\n", 2), ("v ++1 0 0\n", 1), ("v 0 0 0 x\n", 1),
    ("v nan 0 0\n", 1), ("v 0 inf 0\n", 1), ("v 0 0 -infinity\n", 1), ("v 1e400 0 0\n", 1),
    ("v 0 0 0\nv 1 0 0\nf 1 2\n", 3), ("f\n", 1),
    (THREE + "f 0 1 2\n", 4), (THREE + "f 1 2 4\n", 4), ("v 0 0 0\nv 1 0 0\nf 1 2 3\nv 0 1 0\n", 3),
    (THREE + "f -1 -2 -4\n", 4),
    (THREE + "f 1 2 99999999999999999999999\n", 4), (THREE + "f 1 2 -99999999999999999999999\n", 4),
    (THREE + "f 1 2 -9223372036854775808\n", 4)])
def test_rejected_lines(text, line):
    assert line_of(text) == line


@pytest.mark.parametrize("corner", ["a", "1/2/3/4", "1/x", "/1", "1/1/y", "1.5", "+1", ""])
def test_malformed_face_vertices(corner):
    # obj_test.cpp:228-238
    assert line_of(THREE + "f 1 2 " + corner + ("/\n" if corner == "" else "\n")) == 4


# obj_test.cpp:240-300: lines of exactly 16 bytes fit a 16-byte limit with any ending.
@pytest.mark.parametrize("text", ["v 0.00000001 0 0\n", "v 0.00000001 0 0\r\n", "v 0.00000001 0 0",
                                  "v 0.00000001 0 0\r"])
def test_lines_at_the_length_limit(text):
    assert obj(text, "max_line_length=16")["ok"]


@pytest.mark.parametrize("text", ["v 0 0 0\nv 0.000000000001 0 0\n", "v 0 0 0\nv 0.000000001 0 0\n",
                                  "v 0 0 0\nv 0.000000001 0 0", "v 0 0 0\nv 0.000000001 0 0\r\n"])
def test_lines_over_the_length_limit(text):
    assert line_of(text, "max_line_length=16") == 2


def test_overlong_line_message():
    assert obj("v 0 0 0\nv 0.000000001 0 0\n", "max_line_length=16")["message"] == \
        "OBJ line 2: line longer than 16 bytes"


def test_stream_failures():
    r = obj(TRI, "--failed")
    assert r["kind"] == "failure" and r["message"] == "OBJ input stream is not readable"
    r = obj("v 0 0 0\n", "--read-error")
    assert r["kind"] == "failure" and r["message"] == "failed to read OBJ input"
    r = obj(TRI, "--huge-limit")
    assert r["kind"] == "invalid" and r["message"] == "OBJ line length limit is too large"


# obj_test.cpp:309-345: limits
def test_limits():
    assert obj(TRI, "max_vertices=3")["ok"]
    assert line_of(SQUARE, "max_vertices=3") == 4
    assert obj(SQUARE + "f 1 2 3\nf 1 2 3 4\n", "max_triangles=3")["faces"] == 3
    assert line_of(SQUARE + "f 1 2 3 4\nf 1 2 3 4\n", "max_triangles=3") == 6
    assert line_of(SQUARE + "f 1 2 3 4\n", "max_triangles=1") == 5
    assert obj("vt 0 0\nvt 1 0\n", "max_texture_coordinates=2")["ok"]
    assert line_of("vt 0 0\nvt 1 0\nvt 0 1\n", "max_texture_coordinates=2") == 3


# obj_test.cpp:347-460: texture coordinates
def test_texture_coordinates():
    assert obj(THREE + "f 1/ 2// 3//\n")["faces"] == 1
    m = obj("vt 0.25\nvt 0.5 0.75\nvt 0.125 1 0.3\n")
    assert m["t"] == [(0.25, 0.0), (0.5, 0.75), (0.125, 1.0)]
    assert obj("vt 0.5 0.75 0 7 8\n")["t"] == [(0.5, 0.75)]
    m = obj(THREE + "vt 0 0\nvt 1 0\nvt 0 1\nf 1/3 2/1 3/2\nf 1/1/1 2/2/1 3/3/1\n")
    assert m["tf"] == [(2, 0, 1), (0, 1, 2)]
    m = obj(SQUARE + "vt 0 0\nvt 1 0\nvt 1 1\nvt 0 1\nf 1/4 2/3 3/2 4/1\n")
    assert m["tf"] == [(3, 2, 1), (3, 1, 0)]
    m = obj(THREE + "vt 9 9\nvt 0 0\nvt 1 0\nvt 0 1\nf 1/-3 2/-2 3/-1\n")
    assert m["tf"] == [(1, 2, 3)]
    m = obj("vt 0 0\n" + TRI)
    assert m["faces"] == 1 and m["uvfaces"] == 0 and m["uv"] == 1


@pytest.mark.parametrize("faces", ["f 1/1 2/1 3/1\nf 1 2 3\n", "f 1/1 2/1 3/1\nf 1/1 2 3/1\n",
                                   "f 1/1 2/1 3/1\nf 1// 2//1 3//\n", "f 1 2 3\nf 1/1 2/1 3/1\n"])
def test_one_untextured_face_drops_texture_faces_for_the_mesh(faces):
    m = obj(THREE + "vt 0 0\n" + faces)
    assert m["faces"] == 2 and m["uvfaces"] == 0 and m["uv"] == 1


def test_polygon_with_an_untextured_middle_corner():
    m = obj(SQUARE + "vt 0 0\nf 1/1 2/1 3 4/1\n")
    assert m["faces"] == 2 and m["uvfaces"] == 0


def test_texture_errors():
    assert "texture coordinate" in obj(THREE + "vt 0 0\nf 1/1 2/1 3/2\n")["message"]
    assert line_of("vt 0 0\nvt\n") == 2
    for text in ["vt a\n", "vt 0 b\n", "vt 0 0 c\n", "vt nan\n", "vt 0 inf\n", "vt 1e999\n", "vt 0 0 0 x\n"]:
        assert line_of(text) == 1, text
    assert line_of(THREE + "vt 0 0\nf 1/0 2/1 3/1\n") == 5
    assert line_of(THREE + "vt 0 0\nf 1/1 2/1 3/2\n") == 5
    assert line_of(THREE + "f 1/1 2/1 3/1\nvt 0 0\n") == 4
    assert line_of(THREE + "vt 0 0\nf 1/-1 2/-1 3/-2\n") == 5
    for idx in ["99999999999999999999999", "-99999999999999999999999", "-9223372036854775808"]:
        assert line_of(THREE + "vt 0 0\nf 1/1 2/1 3/" + idx + "\n") == 5


# obj_test.cpp:517-556: files
def test_files(tmp_path):
    good = tmp_path / "t.obj"
    good.write_text(TRI)
    assert obj("", "--file", good)["faces"] == 1
    r = obj("", "--file", tmp_path / "missing_file.obj")
    assert r["kind"] == "failure" and "missing_file.obj" in r["message"]
    bad = tmp_path / "bad.obj"
    bad.write_text("v 0 0 0\nf 1 2 3\n")
    assert obj("", "--file", bad)["kind"] == "parse"
