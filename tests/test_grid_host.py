"""Host side of the lattice queries: the point order and rounding the device lattice follows."""
import numpy as np

from ebgeometry_amd import grid_points


def test_grid_points_layout_and_rounding():
    origin = np.array([-1.3, 0.25, 7.0])
    spacing = np.array([0.1, 1.0 / 3.0, 2.5e-3])
    dims = (5, 4, 3)
    p = grid_points(origin, spacing, dims)
    assert p.shape == (60, 3) and p.dtype == np.float32
    n = 0
    for k in range(3):
        for j in range(4):
            for i in range(5):
                # FP64 product and sum, each rounded, then one rounding to FP32
                want = [np.float32(origin[c] + idx * spacing[c]) for c, idx in enumerate((i, j, k))]
                assert list(p[n]) == want, (i, j, k)
                n += 1


def test_grid_points_empty():
    assert grid_points([0, 0, 0], [1, 1, 1], (0, 3, 3)).shape == (0, 3)
