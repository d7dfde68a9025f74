"""Seeded synthetic point clouds for the benchmark configurations.

All generators use ``numpy.random.default_rng(seed)`` and return an (N, 3)
float64 array of collocation points.
"""
from __future__ import annotations

import numpy as np


def sphere(n: int, seed: int = 0) -> np.ndarray:
    """N points uniform on the unit sphere (normal vectors, normalised)."""
    rng = np.random.default_rng(seed)
    p = rng.standard_normal((n, 3))
    return p / np.linalg.norm(p, axis=1, keepdims=True)


def cube_surface(n: int, seed: int = 0) -> np.ndarray:
    """N points on the surface of the unit cube [0,1]^3: face uniform, then
    uniform on the face."""
    rng = np.random.default_rng(seed)
    face = rng.integers(0, 6, n)
    uv = rng.random((n, 2))
    p = np.empty((n, 3))
    axis = face // 2
    side = (face % 2).astype(np.float64)
    for a in range(3):
        m = axis == a
        others = [b for b in range(3) if b != a]
        p[m, a] = side[m]
        p[m, others[0]] = uv[m, 0]
        p[m, others[1]] = uv[m, 1]
    return p


def ellipsoid(n: int, axes, seed: int = 0) -> np.ndarray:
    """N points on the ellipsoid with semi-axes ``axes`` (sphere points scaled)."""
    return sphere(n, seed) * np.asarray(axes, np.float64)[None, :]


def ellipsoid_area(axes) -> float:
    """Surface area of an ellipsoid (Knud Thomsen's approximation, p = 1.6075)."""
    a, b, c = (float(v) for v in axes)
    p = 1.6075
    return 4 * np.pi * (((a * b) ** p + (a * c) ** p + (b * c) ** p) / 3) ** (1 / p)


def uneven_batch(n_mat: int = 64, seed: int = 2024, n_min: int = 4096, n_max: int = 65536):
    """The batch of independent problems: per system N log-uniform in
    [n_min, n_max], an ellipsoid with semi-axes (1, a, b), a, b ~ U(1, 4)
    (points seeded 1000 + s), and kappa ~ U(0, 5).  Returns a list of
    (points, kappa, quadrature weight = area / N)."""
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_mat):
        N = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
        a, b = rng.uniform(1, 4, 2)
        kappa = float(rng.uniform(0, 5))
        axes = (1.0, a, b)
        out.append((ellipsoid(N, axes, seed=1000 + s), kappa, ellipsoid_area(axes) / N))
    return out
