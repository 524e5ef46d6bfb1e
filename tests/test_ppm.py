"""PPM / sRGB output of the drop-in (albedo_amd/host/scene_file.hpp) against the
reference's cases (tests/color/srgb_test.cpp, tests/image/ppm_test.cpp): the sRGB
curve and its breakpoint, byte rounding and clamping, P6 header and row order,
stream and file failures. CPU only."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CHECK = os.path.join(ROOT, "tests", "cpp", "ppm_check")


def test_srgb_and_ppm_match_the_reference_cases(tmp_path):
    if not os.path.exists(CHECK):
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True)
        assert r.returncode == 0, r.stderr
    r = subprocess.run([CHECK, str(tmp_path)], capture_output=True, text=True, timeout=60)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "failures 0" in r.stdout and "checks 17" in r.stdout
