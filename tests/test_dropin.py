"""The C++ drop-in API (albedo_amd/host/albedo_dropin.hpp) builds against the
C ABI and, on a GPU, renders through it with the reference's error behaviour."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "dropin_render")


def test_dropin_header_compiles_and_links():
    if not os.path.exists(BIN):
        r = subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "cpp")], capture_output=True, text=True)
        assert r.returncode == 0, r.stdout + r.stderr
    assert os.path.exists(BIN)


@pytest.mark.gpu
def test_dropin_render_runs_on_the_gpu():
    r = subprocess.run([BIN, "64"], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    lines = dict(l.split(" ", 1) for l in r.stdout.strip().splitlines())
    assert float(lines["mean"]) > 0.01
    assert lines["errors"] == "3"
