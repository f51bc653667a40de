"""Scene inputs shared by the GPU parity tests (seeded, deterministic)."""
import numpy as np

import gpr


def soup(n, seed=2, sigma=0.005):
    return gpr.gen_soup(seed, 0, n, sigma)


def duplicate_heavy(n, seed=9):
    """Many triangles share centroid cells: exercises the equal-key midpoint rule."""
    rng = np.random.default_rng(seed)
    base = rng.random((max(n // 8, 1), 3)).astype(np.float32)
    c = base[rng.integers(0, len(base), n)]
    jitter = rng.normal(0, 1e-4, (n, 3, 3)).astype(np.float32)
    return (c[:, None, :] + jitter).reshape(n, 9).astype(np.float32)


def identical(n):
    t = np.array([0.1, 0.2, 0.3, 0.5, 0.2, 0.3, 0.1, 0.6, 0.3], np.float32)
    return np.tile(t, (n, 1))


def scene_from_soup(pos):
    return gpr.Scene.soup(pos).commit()
