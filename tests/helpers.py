"""Shared helpers for the parity tests: H-matrices built by the oracle (the
checker) and handed to the product through the C ABI."""
import functools

import numpy as np

import oracle as O
from ahmed_amd import geometry
from conftest import golden_points, load_golden


@functools.lru_cache(maxsize=None)
def oracle_api():
    import os
    if not os.path.exists(O.LIB):
        O.build(ref=False)
    return O.Oracle()


@functools.lru_cache(maxsize=8)
def golden_hmatrix(name):
    """(golden dict, oracle tree, oracle H) for a golden case."""
    g = load_golden(name)
    o = oracle_api()
    pts = golden_points(g)
    T = o.tree(pts, int(g["bmin"]), float(g["eta2"]), int(g["kind"]))
    H = o.assemble(T, pts, int(g["kernel"]), float(g["kappa"]), float(g["eps"]), weight=float(g["weight"]))
    return g, T, H


@functools.lru_cache(maxsize=8)
def sphere_hmatrix(N, bmin, eps, seed=5, kernel=O.KERNEL_LAPLACE, kappa=0.0, weighted=True):
    o = oracle_api()
    pts = geometry.sphere(N, seed)
    T = o.tree(pts, bmin, 0.64)
    H = o.assemble(T, pts, kernel, kappa, eps, weight=(4 * np.pi / N) if weighted else 1.0)
    return pts, T, H


def leaves_of(H):
    from ahmed_amd import _lib
    L = H.layout()
    return _lib.leaf_array(L["b1"], L["n1"], L["b2"], L["n2"], L["rank"], L["offset"]), L["data"]


def relerr(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.max(np.abs(a - b)) / max(np.max(np.abs(b)), 1e-300))
