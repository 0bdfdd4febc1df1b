"""Seeded synthetic scenes: particle counts in closed form, canonical FP32
positions inside their cells, and the type predicate of each workload."""
import math

import numpy as np
import pytest

import pyoracle as O


@pytest.mark.parametrize("kind,n,want", [
    (O.SCENE_RT, 64, 62 ** 3 * 8),                 # 1,906,624
    (O.SCENE_RT, 256, 254 ** 3 * 8),               # 131,096,512
    (O.SCENE_BUBBLE, 128, 126 * 76 * 126 * 8),     # 9,652,608: rows j = 1..76 (centre y < 0.6)
    (O.SCENE_DAMBREAK, 256, 101 * 153 * 254 * 8),  # i = 1..101 (x < 0.4), j = 1..153 (y < 0.6)
])
def test_counts(kind, n, want):
    assert O.lib().orc_scene_count(kind, n, n, n) == want


def test_positions_jittered_inside_cells_and_types():
    n = 16
    dx = 1.0 / n
    x, y, z, t = O.scene(O.SCENE_RT, n, n, n, dx, 3)
    assert x.dtype == np.float32 and len(x) == 14 ** 3 * 8
    i = np.floor(x.astype(np.float64) / dx).astype(int)
    assert i.min() >= 1 and i.max() <= n - 2
    # every cell holds 8 particles, in (k, j, i) order
    cells = (i + n * (np.floor(y / dx).astype(int) + n * np.floor(z / dx).astype(int)))
    assert np.all(np.bincount(cells)[cells] == 8)
    yi = 0.5 + 0.05 * np.cos(2 * math.pi * x.astype(np.float64)) * np.cos(2 * math.pi * z.astype(np.float64))
    assert np.array_equal(t == 0, y.astype(np.float64) > yi)
    # jitter: not the reference's fixed sub-lattice
    frac = x.astype(np.float64) / dx - i
    assert 0.2 < frac.std() < 0.4
    # the seed changes positions
    x2, *_ = O.scene(O.SCENE_RT, n, n, n, dx, 4)
    assert not np.array_equal(x, x2)


def test_bubble_types():
    n = 32
    dx = 1.0 / n
    x, y, z, t = O.scene(O.SCENE_BUBBLE, n, n, n, dx, 1)
    r2 = (x - 0.5) ** 2 + (y - 0.25) ** 2 + (z - 0.5) ** 2
    assert np.array_equal(t == 1, r2.astype(np.float64) < 0.12 ** 2 + 0 * r2)
    assert y.max() < 0.6 + dx
