"""Constants the HIP sources hard-code in place of a computation, checked
against the computation they replace (CPU only)."""

import math
import os
import re

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PHYSICS = os.path.join(ROOT, "lbm_jax_autodiff_amd", "csrc", "lbm_physics.hpp")


def _guard_usq_literals():
    text = open(PHYSICS).read()
    f32 = re.search(r"guard_usq<float>\(\)\s*\{\s*return\s+(0x[0-9a-fA-Fp.+-]+)f;", text).group(1)
    f64 = re.search(r"guard_usq<double>\(\)\s*\{\s*return\s+(0x[0-9a-fA-Fp.+-]+);", text).group(1)
    return float.fromhex(f32), float.fromhex(f64)


def test_guard_speed_threshold_float32():
    """|u| > 0.4 (state.py:158-160) is tested as u.u > t: t must be the
    largest float32 whose correctly rounded sqrt is <= float32(0.4)."""
    t32, _ = _guard_usq_literals()
    t = np.float32(t32)
    assert float(t) == t32  # exactly a float32
    c = np.float32(0.4)
    assert np.sqrt(t) <= c
    assert np.sqrt(np.nextafter(t, np.float32(np.inf))) > c
    # every float32 in a neighbourhood agrees with the sqrt form
    lo = t
    for _ in range(200):
        lo = np.nextafter(lo, np.float32(0))
    grid = [lo]
    for _ in range(400):
        grid.append(np.nextafter(grid[-1], np.float32(1)))
    grid = np.array(grid, dtype=np.float32)
    assert np.array_equal(np.sqrt(grid) > c, grid > t)


def test_guard_speed_threshold_float64():
    _, t = _guard_usq_literals()
    assert math.sqrt(t) <= 0.4
    assert math.sqrt(np.nextafter(t, np.inf)) > 0.4
    grid = [t]
    for _ in range(200):
        grid.insert(0, float(np.nextafter(grid[0], 0.0)))
        grid.append(float(np.nextafter(grid[-1], 1.0)))
    assert all((math.sqrt(x) > 0.4) == (x > t) for x in grid)
