"""Host-side pieces of the learned initialiser (no GPU)."""

import numpy as np

from py4dstem_amd.learned_init import ARCH, neighbour_table


def test_neighbour_table_is_clamped_3x3_raster():
    grid = (4, 6)
    t = neighbour_table(grid, np.array([[0, 0], [2, 3], [3, 5]]))
    assert t.shape == (3, 9) and t.dtype == np.int32
    # interior: raster indices of the 3x3 block around (2, 3), row-major over (di, dj)
    assert list(t[1]) == [(2 + di) * 6 + (3 + dj) for di in (-1, 0, 1) for dj in (-1, 0, 1)]
    # corner (0, 0): neighbours clamp onto the edge
    assert list(t[0]) == [0, 0, 1, 0, 0, 1, 6, 6, 7]
    assert t[2].max() == 23 and t[2][4] == 3 * 6 + 5
    # per-sample base offsets (several scans in one frame array)
    t2 = neighbour_table(grid, np.array([[1, 1], [1, 1]]), base=np.array([0, 24]))
    assert np.array_equal(t2[1], t2[0] + 24)


def test_architecture_receptive_field_covers_the_patch():
    # receptive field on the 32x32 output grid after the stride-2 first layer
    assert ARCH[0][2] == 2
    rf = 1
    for ci, co, stride, dil, _ in ARCH[1:]:
        rf += 2 * dil
    assert rf >= 32
    assert len(ARCH) == 6 and all(co in (16, 64) for _, co, _, _, _ in ARCH)
