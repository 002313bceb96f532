"""Point clouds for the nearest-neighbour tests and the golden-fixture generator: a seeded cloud of
uniform points, tight clusters and exact duplicates (ties at distance 0), and query points around
it."""
from __future__ import annotations

import numpy as np


def cloud(n=6000, seed=3):
    rng = np.random.default_rng(seed)
    a = rng.uniform(-1, 1, (n // 2, 3))
    centres = rng.uniform(-1, 1, (20, 3))
    b = centres[rng.integers(0, 20, n // 2 - 100)] + rng.normal(0, 0.02, (n // 2 - 100, 3))
    dup = a[rng.integers(0, a.shape[0], 100)]  # coincident with other cloud points
    return np.concatenate([a, b, dup]).astype(np.float32)


def queries(n=2500, seed=4):
    rng = np.random.default_rng(seed)
    return rng.uniform(-1.3, 1.3, (n, 3)).astype(np.float32)
