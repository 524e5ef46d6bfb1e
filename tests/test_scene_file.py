"""The drop-in scene-file reader (albedo_amd/host/scene_file.hpp) against the
reference's own parser tests (tests/scene/scene_file_test.cpp): accepted
statements, the line each error names, its wording, limits, OBJ meshes and
mesh-path rules. CPU only: reading a scene makes no device call."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "scene_check")
CAMERA = "camera from 0 0 0 at 0 0 -1\n"


@pytest.fixture(scope="module", autouse=True)
def built():
    if not os.path.exists(CHECK):
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr


def read(text, *args):
    r = subprocess.run([CHECK, "-"] + [str(a) for a in args], input=text.encode("latin-1"), capture_output=True,
                       timeout=60)
    out = r.stdout.decode("latin-1").strip()
    if out.startswith("ok"):
        fields = dict(kv.split("=", 1) for kv in out.split()[1:])
        return {"ok": True, **fields}
    if out.startswith("error"):
        head, msg = out.split(" ", 2)[1], out.split(" ", 2)[2]
        return {"ok": False, "line": int(head.split("=")[1]), "message": msg}
    return {"ok": False, "line": None, "message": out}


def line_of(text, *args):
    r = read(text, *args)
    return None if r["ok"] else r["line"]


# camera (scene_file_test.cpp:86-140)
def test_a_scene_needs_only_a_camera():
    r = read(CAMERA)
    assert r["ok"] and r["objects"] == "0" and r["fov"] == "60"


def test_camera_fields_in_any_order_with_defaults():
    r = read("camera at 0 0 -1 fov 40 from 1 2 3 up 0 1 0\n")
    assert r["ok"] and r["from"] == "1,2,3" and r["fov"] == "40"


@pytest.mark.parametrize("text,line", [("sky 0.1 0.1 0.1\n", 0), ("", 0), (CAMERA + "camera from 1 0 0 at 0 0 0\n", 2),
                                       ("camera at 0 0 -1\n", 1), ("camera from 0 0 0\n", 1),
                                       ("camera from 1 1 1 at 1 1 1\n", 1),
                                       ("camera from 0 0 0 at 0 0 -1 fov 180\n", 1),
                                       ("camera from 0 0 0 at 0 0 -1 up 0 0 1\n", 1)])
def test_camera_errors(text, line):
    assert line_of(text) == line


def test_missing_camera_message():
    assert read("")["message"] == "scene: the scene has no camera"


# sky
@pytest.mark.parametrize("rest", ["sky 1 1 1\nsky 1 1 1\n", "sky 1 1 1 2\n", "sky nadir 1 1 1 colour 1 1 1\n",
                                  "sky nadir 1 1 1\n", "sky inf 1 1\n"])
def test_sky_errors(rest):
    assert line_of(CAMERA + rest) in (2, 3)


def test_sky_forms():
    assert read(CAMERA + "sky 0.1 0.2 0.3\n")["ok"]
    assert read(CAMERA + "sky zenith 0 0 1 nadir 1 1 1\n")["ok"]


# materials (scene_file_test.cpp:194-500)
def test_reads_every_material_and_sphere():
    text = CAMERA + ("material red diffuse color 1 0 0\n"
                     "material c checker even 1 1 1 odd 0 0 0 frequency 2\n"
                     "material p perlin low 0 0 0 high 1 1 1 scale 3 veins 4\n"
                     "material g conductor eta 0.2 0.4 1.3 k 3.6 2.4 1.9 roughness 0.2\n"
                     "material d dielectric ior 1.5 roughness 0.1\n"
                     "material mx mix first g second red scale 2 veins 5 threshold 0.3\n"
                     "sphere 0 0 -5 radius 1 material mx\n"
                     "sphere 0 0 -5 material red radius 2\n")
    assert read(text)["objects"] == "2"


@pytest.mark.parametrize("statement", [
    "material m diffuse color -0.1 0 0", "material m checker even 1 1 1 odd 0 0 0 frequency 0",
    "material m checker even 1 1 1 odd 0 0 0 frequency -2", "material m perlin low 0 0 0 high 1 1 1 scale 0",
    "material m perlin low 0 0 0 high 1 1 1 scale 1 veins -1", "material m perlin low 0 0 0 high 1 1 1 scale 1 veins 101",
    "material m perlin low -1 0 0 high 1 1 1 scale 1", "material m perlin low 0 0 0 high 1 1 1.5 scale 1",
    "material m conductor eta 0 1 1 k 1 1 1", "material m conductor eta 1 1 1 k -1 1 1",
    "material m conductor eta 1 1 1 k 1 1 1 roughness -0.5", "material m dielectric ior 0",
    "material m dielectric ior 1.5 roughness -1",
    # rejectsValuesThatWouldBreakRendering
    "material m conductor eta 1 1 1 k 1 1 1 roughness 1.5", "material m conductor eta 1 1 1 k 1 1 1 roughness 1e200",
    "material m dielectric ior 1.5 roughness 1e200", "material m dielectric ior 1e-310", "material m dielectric ior 0.001",
    "material m dielectric ior 1e300", "material m conductor eta 1e-310 1 1 k 1 1 1",
    "material m conductor eta 1e300 1 1 k 1 1 1", "material m conductor eta 1 1 1 k 1e300 1 1",
    "material m diffuse color 1.1 0.5 0.5", "material m checker even 1 1 1 odd 2 0 0 frequency 1",
    "material m conductor eta 0.5 1 0.5 k 1 0 1",
    # missing name/type/fields
    "material", "material m", "material m dielectric ior 1.5 ior 1.3", "material m diffuse",
    "material m checker even 1 1 1 odd 0 0 0", "material m conductor eta 1 1 1", "material m dielectric roughness 0",
    "material m perlin high 1 1 1 scale 1", "material m perlin low 0 0 0 scale 1", "material m perlin low 0 0 0 high 1 1 1",
    "material red plastic color 1 0 0", "material 9lives diffuse color 1 1 1"])
def test_material_errors_name_line_two(statement):
    assert line_of(CAMERA + statement + "\n") == 2


def test_values_at_the_edges_are_accepted():
    assert read(CAMERA + "material a conductor eta 0.001 1000 1 k 0 1000 1 roughness 1\n"
                         "material b dielectric ior 0.01 roughness 0\n"
                         "material c dielectric ior 100\n"
                         "material d diffuse color 0 1 1\n"
                         "material e checker even 1 1 1 odd 0 0 0 frequency 1\n"
                         "material f perlin low 0 0 0 high 1 1 1 scale 1 veins 0\n"
                         "material g perlin low 0 0 0 high 1 1 1 scale 1 veins 100\n"
                         "sky 1e6 1e6 1e6\n"
                         "point_light 0 0 0 intensity 1e6 0 1e6\n")["ok"]


def test_mix_rules():
    base = CAMERA + "material a diffuse color 1 1 1\nmaterial b dielectric ior 1.5\n"
    assert read(base + "material m mix first a second a scale 1\n")["ok"]
    assert read(base + "material m mix first a second b scale 1 threshold 0\n")["ok"]
    assert read(base + "material m mix first a second b scale 1 threshold 1\n")["ok"]
    r = read(base + "material m mix first a second b scale 1\nmaterial n mix first m second a scale 1\n")
    assert r["line"] == 5 and "'m' is a mix, and a mix cannot hold another mix" in r["message"]
    assert line_of(base + "material m mix first a second zz scale 1\n") == 4
    assert line_of(base + "material m mix first a second b\n") == 4
    assert line_of(base + "material m mix first a second b scale 1 threshold 2\n") == 4


def test_names_and_definition_order():
    assert line_of(CAMERA + "sphere 0 0 0 radius 1 material m\n") == 2
    r = read(CAMERA + "material m diffuse color 1 1 1\nmaterial m diffuse color 1 1 1\n")
    assert r["line"] == 3 and "material 'm' is already defined" in r["message"]
    assert read(CAMERA + "material _a-1 diffuse color 1 1 1\n")["ok"]


# objects and lights
@pytest.mark.parametrize("statement", ["sphere 0 0 0 radius 0 material w", "sphere 0 0 0 radius -1 material w",
                                       "sphere 0 0 0 radius 1e200 material w", "sphere 0 0 0 material w",
                                       "sphere 0 0 0 radius 1", "sphere 0 0 radius 1 material w",
                                       "area_light 0 0 0 radius 1", "area_light 0 0 0 radiance 1 1 1",
                                       "area_light 0 0 0 radius 1 radiance 1 1 1 material nope",
                                       "point_light 0 0 0 intensity -1 1 1", "area_light 0 0 0 radius 1 radiance 1 -1 1",
                                       "point_light 0 0 0 intensity 1e7 1 1", "sky 1 2e6 1",
                                       "cube 0 0 0", "point_light 0 0 0 intensity 1 1 1 colour 1 1 1",
                                       "point_light 0 0 0 intensity 1 1 1 intensity 2 2 2", "point_light 0 0",
                                       "point_light 0 0 0 intensity 1 1"])
def test_object_errors(statement):
    assert line_of(CAMERA + "material w diffuse color 1 1 1\n" + statement + "\n") == 3


def test_lights_and_objects():
    r = read(CAMERA + "material w diffuse color 1 1 1\npoint_light 0 1 0 intensity 1 2 3\n"
                      "area_light 0 5 0 radius 1 radiance 1 1 1\narea_light 0 5 0 radius 1 radiance 1 1 1 material w\n")
    assert r["ok"] and r["objects"] == "2"


@pytest.mark.parametrize("value", ["x", "1.5.2", "nan", "inf", "1e999", "0x10", "--1", "+-1"])
def test_malformed_numbers(value):
    assert line_of(CAMERA + "point_light 0 " + value + " 0 intensity 1 1 1\n") == 2


def test_comments_blank_lines_crlf_and_no_final_newline():
    assert read("# a comment\n\ncamera from 0 0 0 at 0 0 -1 # trailing\r\n   \nsky 1 1 1")["ok"]


def test_error_messages():
    m = read("camera from 0 x 0 at 0 0 -1\n")["message"]
    assert "the y coordinate of the camera position" in m
    m = read(CAMERA + "\x1b[2Jcube" + "x" * 200 + "\n")["message"]
    assert "\x1b" not in m and "\\x1B[2Jcube" in m and len(m) < 100 and "'..." in m
    m = read(CAMERA + "cube 0 0 0\n")["message"]
    assert m == "scene line 2: unknown statement 'cube'"
    m = read(CAMERA + "material m dielectric ior 200\n")["message"]
    assert m == "scene line 2: the index of refraction must be between 0.01 and 100, got 200"
    m = read(CAMERA + "material m checker even 1 1 1 odd 0 0 0 frequency -2\n")["message"]
    assert m == "scene line 2: the checker frequency must be positive, got -2"
    m = read(CAMERA + "material m diffuse\n")["message"]
    assert m == "scene line 2: missing 'color'"


# limits (scene_file_test.cpp:657-700)
def test_limits():
    assert line_of(CAMERA + "point_light 0 0 0 intensity 1 1 1\n", "max_line_length=30") == 2
    two = CAMERA + "point_light 0 0 0 intensity 1 1 1\n" * 2
    assert read(two, "max_objects=2")["ok"]
    assert line_of(CAMERA + "point_light 0 0 0 intensity 1 1 1\n" * 3, "max_objects=2") == 4
    mats = CAMERA + "".join("material m%d diffuse color 1 1 1\n" % i for i in range(3))
    assert read(mats, "max_materials=3")["ok"]
    assert line_of(mats + "material x diffuse color 1 1 1\n", "max_materials=3") == 5
    assert line_of(CAMERA + "material w diffuse color 1 1 1\narea_light 0 0 0 radius 1 radiance 1 1 1\n"
                   "area_light 0 0 0 radius 1 radiance 1 1 1\n", "max_objects=1") == 4


# meshes and instances (scene_file_test.cpp:803-1240)
TENT = ("v -1 0 0\nv 1 0 0\nv 0 1 0\nv 0 0 1\nvt 0 0\nvt 1 0\nvt 0 1\n"
        "f 1/1 2/2 3/3\nf 1/1 4/2 2/3\n")


@pytest.fixture()
def meshdir(tmp_path):
    (tmp_path / "tent.obj").write_text(TENT)
    (tmp_path / "quad.obj").write_text("v 0 0 0\nv 1 0 0\nv 1 1 0\nv 0 1 0\nf 1 2 3 4\n")
    (tmp_path / "flat.obj").write_text("v 0 0 0\nv 1 0 0\nv 2 0 0\nf 1 2 3\n")
    (tmp_path / "bad.obj").write_text("v 0 0 0\nv 1 0 0\nf 1 2 7\n")
    (tmp_path / "sub").mkdir()
    (tmp_path / "sub" / "t.obj").write_text(TENT)
    return tmp_path


def test_mesh_and_instances(meshdir):
    text = CAMERA + ("material w diffuse color 1 1 1\nmesh tent tent.obj smooth\nmesh q quad.obj\n"
                     "mesh s sub/t.obj\n"
                     "instance tent material w translate 0 0 -5 scale 2 rotate 0 1 0 30 fit 1.5 ground\n"
                     "instance q material w scale 1 2 3\ninstance tent scale 2 material w\ninstance s material w\n")
    r = read(text, meshdir)
    assert r["ok"] and r["objects"] == "4", r


@pytest.mark.parametrize("rest,line", [
    ("mesh t tent.obj\nmesh t tent.obj\n", 3), ("instance nope material w\n", 2),
    ("mesh t tent.obj\ninstance t\n", 3), ("mesh t tent.obj\ninstance t material w material w\n", 3),
    ("mesh t tent.obj\ninstance t material w spin 1\n", 3), ("mesh t tent.obj\ninstance t material w fit 0\n", 3),
    ("mesh t tent.obj\ninstance t material w scale 0\n", 3), ("mesh t tent.obj\ninstance t material w scale 1 x 1\n", 3),
    ("mesh t tent.obj\ninstance t material w rotate 0 0 0 10\n", 3), ("mesh t tent.obj extra\n", 2),
    ("mesh f flat.obj\n", 2), ("mesh b bad.obj\n", 2), ("mesh m missing.obj\n", 2), ("mesh m ../x.obj\n", 2),
    ("mesh m /etc/passwd\n", 2), ("mesh m nul.obj\n", 2), ("mesh m com1\n", 2), ("mesh m sub\n", 2)])
def test_mesh_errors(meshdir, rest, line):
    assert line_of(CAMERA + "material w diffuse color 1 1 1\n" + rest, meshdir) == line + 1


def test_mesh_messages(meshdir):
    base = CAMERA + "material w diffuse color 1 1 1\n"
    assert "mesh 'bad.obj': OBJ line 3: a face refers to a vertex not yet defined" in read(base + "mesh b bad.obj\n", meshdir)["message"]
    assert "mesh file 'missing.obj' does not exist" in read(base + "mesh m missing.obj\n", meshdir)["message"]
    assert "must be relative to the scene and not use '..'" in read(base + "mesh m ../x.obj\n", meshdir)["message"]
    assert "names a device" in read(base + "mesh m aux.txt\n", meshdir)["message"]
    assert "is not a regular file" in read(base + "mesh m sub\n", meshdir)["message"]
    assert "has no triangles" in read(base + "mesh f flat.obj\n", meshdir)["message"]
    assert read(base + "mesh m con1.obj\n", meshdir)["message"].endswith("does not exist")  # only resembles a device
    assert read(CAMERA + "mesh t tent.obj\n")["message"] == \
        "scene line 2: meshes can only be read from a scene file, which they sit beside"


def test_mesh_limits(meshdir):
    base = CAMERA + "material w diffuse color 1 1 1\n"
    assert line_of(base + "mesh a tent.obj\nmesh b tent.obj\n", meshdir, "max_meshes=1") == 4
    assert line_of(base + "mesh a tent.obj\nmesh b tent.obj\n", meshdir, "max_vertices=6") == 4
    assert line_of(base + "mesh a tent.obj\ninstance a material w " + "translate 0 0 0 " * 3 + "\n", meshdir,
                   "max_instance_operations=2") == 4
    assert line_of(base + "mesh a tent.obj\ninstance a material w ground\ninstance a material w ground\n", meshdir,
                   "max_ground_points=6") == 5


def test_links_leaving_the_directory_are_rejected(meshdir, tmp_path_factory):
    outside = tmp_path_factory.mktemp("outside")
    (outside / "x.obj").write_text(TENT)
    os.symlink(outside / "x.obj", meshdir / "link.obj")
    r = read(CAMERA + "material w diffuse color 1 1 1\nmesh l link.obj\n", meshdir)
    assert r["line"] == 3 and "leads outside the scene's directory" in r["message"]


def test_reads_a_scene_file_with_meshes_beside_it(meshdir):
    (meshdir / "s.scene").write_text(CAMERA + "material w diffuse color 1 1 1\nmesh t tent.obj\ninstance t material w\n")
    r = subprocess.run([CHECK, str(meshdir / "s.scene")], capture_output=True, text=True)
    assert r.stdout.startswith("ok objects=1")
    r = subprocess.run([CHECK, str(meshdir / "none.scene")], capture_output=True, text=True)
    assert "cannot open scene file" in r.stdout
    r = subprocess.run([CHECK, str(meshdir / "sub")], capture_output=True, text=True)
    assert "is not a regular file" in r.stdout
