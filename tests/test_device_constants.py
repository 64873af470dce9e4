"""Constants of the device triangle tests that stand for the reference's double arithmetic
(shapes/Triangle.cu:50-128): checked here against their double definitions, and the float
division against the double quotient on random and edge-case operands."""
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHADOW_EPS = np.float32(0.0001)  # core/ErrorHelper.cuh:40


def _constant(name: str) -> np.float32:
    src = open(os.path.join(ROOT, "lighttransportsimulation_amd", "csrc", "pt_device.h")).read()
    m = re.search(r"constexpr float %s = ([0-9a-fA-Fx.p+-]+)f;" % name, src)
    assert m, name
    return np.float32(float.fromhex(m.group(1)))


def _largest_float_at_most(c: float) -> np.float32:
    f = np.float32(c)
    if np.float64(f) > c:
        f = np.nextafter(f, np.float32(-np.inf))
    return f


def test_barycentric_thresholds_are_the_largest_floats_below_the_double_bounds():
    for name, c in (("kOneMinusEpsF", 1.0 - np.float64(SHADOW_EPS)), ("kOnePlusEpsF", 1.0 + np.float64(SHADOW_EPS))):
        f = _constant(name)
        assert f == _largest_float_at_most(c)
        assert np.float64(f) <= c < np.float64(np.nextafter(f, np.float32(np.inf)))
        # x > f  <=>  (double)x > c, for floats on both sides of the bound
        xs = np.array([np.nextafter(f, np.float32(-np.inf)), f, np.nextafter(f, np.float32(np.inf)),
                       np.float32(1.0), np.float32(0.5), np.float32(2.0)], dtype=np.float32)
        assert np.array_equal(xs > f, xs.astype(np.float64) > c)


def test_float_division_equals_the_rounded_double_quotient():
    rng = np.random.default_rng(7)
    bits = rng.integers(0, 2**32, size=200000, dtype=np.uint64).astype(np.uint32)
    det = bits.view(np.float32)
    det = det[np.isfinite(det) & (det != 0)]
    det = np.concatenate([det, np.float32([1.0, -1.0, 3.0, 1e-38, 1e-45, 3.4e38, 7.0, 0.1])])
    with np.errstate(over="ignore", divide="ignore"):
        via_double = (np.float64(1.0) / det.astype(np.float64)).astype(np.float32)
        in_float = np.float32(1.0) / det
    assert np.array_equal(via_double.view(np.uint32), in_float.view(np.uint32))
