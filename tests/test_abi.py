"""CPU checks of the C ABI: the library loads, exports every entry point the
header declares, and the ctypes mirrors have the header's layout."""
import ctypes
import os
import re

import pytest

from albedo_amd import _abi

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    text = open(os.path.join(ROOT, "include", "albedo_gpu.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(ab_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    L = _abi.lib()
    declared = header_functions()
    assert len(declared) >= 20
    missing = [f for f in declared if not hasattr(L, f)]
    assert not missing, missing


def test_ctypes_prototypes_cover_the_header():
    assert sorted(_abi.PROTOTYPES) == header_functions()


@pytest.mark.parametrize("struct,size", [(_abi.Ray, 32), (_abi.Hit, 16), (_abi.Camera, 88),
                                         (_abi.RenderParams, 32), (_abi.ObjectDesc, 240)])
def test_struct_sizes(struct, size):
    assert ctypes.sizeof(struct) == size


def test_header_constants_match_the_mirror():
    text = open(os.path.join(ROOT, "include", "albedo_gpu.h")).read()
    defines = {k: int(v) for k, v in re.findall(r"#define (AB_[A-Z_0-9]+) (\d+)u", text)}
    mirror = {"AB_TUNE_PLAIN": _abi.AB_TUNE_PLAIN, "AB_TUNE_BATCH": _abi.AB_TUNE_BATCH,
              "AB_ESTIMATOR_PATH": _abi.ESTIMATOR_PATH, "AB_ESTIMATOR_SCATTER_ONLY": _abi.ESTIMATOR_SCATTER_ONLY,
              "AB_ESTIMATOR_NORMALS": _abi.ESTIMATOR_NORMALS, "AB_SCHED_FIFO": _abi.SCHED_FIFO,
              "AB_SCHED_SHARE": _abi.SCHED_SHARE, "AB_SCHED_SMALL_FIRST": _abi.SCHED_SMALL_FIRST}
    for name, value in mirror.items():
        assert defines[name] == value, name


def test_abi_version():
    assert _abi.lib().ab_abi_version() == 1


def test_errors_without_device_are_reported_not_raised():
    # On a machine without a GPU, creating a context fails with a status and a
    # message; on one with a GPU, a bad device index does.
    h = ctypes.c_void_p()
    st = _abi.lib().ab_context_create(1 << 20, ctypes.byref(h))
    assert st != _abi.AB_OK
    assert _abi.lib().ab_last_error(None)
