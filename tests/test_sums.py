"""The progressive wire format's host helpers (ab_sums_merge, ab_sums_resolve):
closed forms, exact merging in any order, argument errors. Host code only."""
import ctypes

import numpy as np
import pytest

import albedo_amd
from albedo_amd import AlbedoError, _abi

ONE = 1 << _abi.SUMS_FRACTION_BITS


def test_resolve_closed_forms():
    # A pixel whose n samples each carried radiance k has sums k * n * 2^28.
    s = np.zeros((2, 3, 4), np.uint64)
    s[0, 0] = (3 * 4 * ONE, 0, ONE // 2 * 4, 4)        # (3, 0, 0.5) over 4 samples
    s[0, 1] = (7 * ONE, 7 * ONE, 7 * ONE, 1)            # one sample of 7
    s[0, 2] = (ONE, 2 * ONE, 3 * ONE, 0)                # no samples yet: black
    s[1, 0] = (1, 0, 0, 1)                              # the smallest step, 2^-28
    s[1, 1] = (100000 * 65536 * ONE, 0, 0, 65536)       # the largest clamped sum
    img = albedo_amd.sums_resolve(s)
    assert img.shape == (2, 3, 3) and img.dtype == np.float32
    assert tuple(img[0, 0]) == (3.0, 0.0, 0.5)
    assert tuple(img[0, 1]) == (7.0, 7.0, 7.0)
    assert tuple(img[0, 2]) == (0.0, 0.0, 0.0)
    assert img[1, 0, 0] == np.float32(2.0 ** -28)
    assert img[1, 1, 0] == 100000.0
    assert np.all(img[1, 2] == 0)


def test_resolve_rounds_once_to_float32():
    # 1/3: the FP64 quotient rounded to FP32 once, not an FP32 division of rounded parts.
    s = np.array([[ONE, 2 * ONE, 10 * ONE, 3]], np.uint64)
    img = albedo_amd.sums_resolve(s)
    assert img[0, 0] == np.float32(1.0 / 3.0)
    assert img[0, 1] == np.float32(2.0 / 3.0)
    assert img[0, 2] == np.float32(10.0 / 3.0)


def test_merge_is_exact_and_order_free():
    rng = np.random.default_rng(3)
    parts = [rng.integers(0, 1 << 40, size=(16, 9, 4), dtype=np.uint64) for _ in range(5)]
    a = np.zeros_like(parts[0])
    for p in parts:
        albedo_amd.sums_merge(a, p)
    b = np.zeros_like(parts[0])
    for p in reversed(parts):
        albedo_amd.sums_merge(b, p)
    assert np.array_equal(a, b)
    assert np.array_equal(a, sum(p.astype(object) for p in parts).astype(np.uint64))


def test_merged_snapshot_resolves_to_weighted_mean():
    # Two sample sets of one pixel: 3 samples of radiance 2, then 1 sample of 6.
    first = np.array([[3 * 2 * ONE, 0, 0, 3]], np.uint64)
    second = np.array([[6 * ONE, 0, 0, 1]], np.uint64)
    merged = albedo_amd.sums_merge(first.copy(), second)
    assert albedo_amd.sums_resolve(merged)[0, 0] == 3.0  # (3*2 + 6) / 4


def test_argument_errors():
    L = _abi.lib()
    assert L.ab_sums_merge(None, None, 1) == _abi.AB_INVALID_ARGUMENT
    assert L.ab_sums_resolve(None, 1, None) == _abi.AB_INVALID_ARGUMENT
    assert L.ab_sums_merge(None, None, 0) == _abi.AB_OK
    with pytest.raises(ValueError):
        albedo_amd.sums_merge(np.zeros((2, 4), np.uint64), np.zeros((3, 4), np.uint64))
