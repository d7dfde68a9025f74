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


def block_matrix(sizes, seed, rank_of=None):
    """Leaves of a block matrix on row/column segments of the given sizes:
    dense diagonal blocks, low-rank off-diagonal blocks of ranks 1..6 (or
    rank_of(i, j)), with the dense N x N matrix they represent (column-major
    leaf values as in mblock: dense n1 x n2, low-rank U n1 x k then V n2 x k)."""
    from ahmed_amd import _lib
    rng = np.random.default_rng(seed)
    b = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int64)
    N = int(b[-1])
    L = {f: [] for f in ("b1", "n1", "b2", "n2", "rank", "offset")}
    chunks, off = [], 0
    A = np.zeros((N, N))
    for i, mi in enumerate(sizes):
        for j, mj in enumerate(sizes):
            if i == j:
                k, vals = -1, rng.standard_normal(mi * mj)
                A[b[i]:b[i + 1], b[j]:b[j + 1]] = vals.reshape(mj, mi).T
            else:
                k = int(1 + (i * 7 + j * 3) % 6) if rank_of is None else int(rank_of(i, j))
                vals = rng.standard_normal(k * (mi + mj))
                U = vals[:mi * k].reshape(k, mi).T
                V = vals[mi * k:].reshape(k, mj).T
                A[b[i]:b[i + 1], b[j]:b[j + 1]] = U @ V.T
            for f, v in zip(("b1", "n1", "b2", "n2", "rank", "offset"), (b[i], mi, b[j], mj, k, off)):
                L[f].append(v)
            chunks.append(vals)
            off += vals.size
    leaves = _lib.leaf_array(L["b1"], L["n1"], L["b2"], L["n2"], L["rank"], L["offset"])
    return N, leaves, np.concatenate(chunks), A


def bf16_rne(a):
    """FP32 values rounded to BF16 (nearest, ties to even), returned as FP32."""
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def bf16_vtx_emulation(N, L, X):
    """The AHM_BF16 multi-RHS product restated in FP64 arithmetic: dense
    blocks and U in FP32, V and X in BF16 (from their FP32 values) for
    t = V^T X, y = D X + U t.  Returns (y, |U| (|V|^T |X|)) with the second
    from the FP64 values (the scale of the BF16 rounding error)."""
    X32 = X.astype(np.float32).astype(np.float64)
    Xb = bf16_rne(X.astype(np.float32)).astype(np.float64)
    Xa = np.abs(X)
    y = np.zeros_like(X)
    lr = np.zeros_like(X)
    data = L["data"]
    for b1, n1, b2, n2, k, o in zip(*(np.asarray(L[f], np.int64) for f in ("b1", "n1", "b2", "n2", "rank",
                                                                           "offset"))):
        if k < 0:
            D = data[o:o + n1 * n2].reshape(n2, n1).T
            y[b1:b1 + n1] += D.astype(np.float32).astype(np.float64) @ X32[b2:b2 + n2]
        elif k > 0:
            U = data[o:o + n1 * k].reshape(k, n1).T
            V = data[o + n1 * k:o + (n1 + n2) * k].reshape(k, n2).T
            Vb = bf16_rne(V.astype(np.float32)).astype(np.float64)
            y[b1:b1 + n1] += U.astype(np.float32).astype(np.float64) @ (Vb.T @ Xb[b2:b2 + n2])
            lr[b1:b1 + n1] += np.abs(U) @ (np.abs(V).T @ Xa[b2:b2 + n2])
    return y, lr
