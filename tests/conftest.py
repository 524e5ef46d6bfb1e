import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def ctx():
    """One library context for all GPU tests."""
    import albedo_amd
    c = albedo_amd.Context(0)
    yield c
    c.close()


def pytest_runtest_setup(item):
    # Name the case before it runs, so a GPU failure that ends the process
    # still says which test was running.
    if "gpu" in item.keywords:
        print("\n[gpu test] %s" % item.nodeid, flush=True)


@pytest.fixture(scope="session")
def icosphere(ctx):
    """The serving workload's 1,310,720-triangle displaced icosphere x 64 instances:
    (scene, device scene, oracle scene), built once per session."""
    import oracle
    from albedo_amd import workloads as W
    s = W.icosphere_scene()
    return s, ctx.scene(s), oracle.OracleScene(s.to_desc())
