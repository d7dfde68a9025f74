"""Triangular H-matrix factors for tests and benchmarks (input generator).

The reference obtains its triangular factors from H-arithmetic (HLU,
H/HLU.cpp:20-66), which this package does not implement.  To exercise and
time the triangular solves at sizes the reference's HLU cannot produce here,
``block_factors`` builds a pair (L, U) on the block structure of an assembled
H-matrix A in the reference's storage (mblock layout of the L and U arrays,
Include/H.h:1383-1388 for how HLU_solve uses them):

* diagonal leaves: the LU decomposition with partial pivoting of A's dense
  diagonal leaf, packed exactly as mblock::decomp_LU does
  (H/mblock_D.cpp:1078-1107): L column j holds the permutation index on the
  diagonal and L's entries below it, U column j holds rows 0..j;
* off-diagonal leaves: A's own leaves strictly below (L) and above (U) the
  diagonal, unchanged.

(L, U) is then a block Gauss-Seidel-like approximation of A with exactly
the leaf sizes and ranks of A; it is a valid input of the solves, not the
factorisation HLU would compute.
"""
from __future__ import annotations

import numpy as np
import scipy.linalg as sla

KIND_ABSENT, KIND_DENSE, KIND_LTM, KIND_UTM = -2, -1, -3, -4


def pack_lu(D: np.ndarray):
    """(packed L, packed U) of a square dense block, mblock::decomp_LU layout."""
    n = D.shape[0]
    lu, piv = sla.lu_factor(np.array(D, order="F", copy=True), check_finite=False)
    perm = np.arange(n)
    Lp = np.empty(n * (n + 1) // 2)
    Up = np.empty(n * (n + 1) // 2)
    oL = 0
    for j in range(n):
        p = int(piv[j])
        perm[j], perm[p] = perm[p], perm[j]
        Up[j * (j + 1) // 2: (j + 1) * (j + 2) // 2] = lu[: j + 1, j]
        Lp[oL] = perm[j]
        Lp[oL + 1: oL + n - j] = lu[j + 1:, j]
        oL += n - j
    return Lp, Up


def block_factors(layout: dict):
    """(L layout, U layout) from an H-matrix layout (b1, n1, b2, n2, rank,
    offset, data; rank < 0 dense) whose diagonal leaves are dense and square.
    Returned layouts use the factor kind codes (-2 absent, -1 dense, k >= 0
    low rank, -3 / -4 packed diagonal) with values in leaf order."""
    b1, n1, b2, n2 = (np.asarray(layout[f]).astype(np.int64) for f in ("b1", "n1", "b2", "n2"))
    rank = np.asarray(layout["rank"]).astype(np.int64)
    off = np.asarray(layout["offset"]).astype(np.int64)
    data = layout["data"]
    diag = (b1 == b2) & (n1 == n2)
    if np.any(diag & (rank >= 0)):
        raise ValueError("diagonal leaves must be dense")
    nv = np.where(rank < 0, n1 * n2, np.maximum(rank, 0) * (n1 + n2))
    out = {}
    for which in "LU":
        side = (b1 > b2) if which == "L" else (b1 < b2)
        kind = np.full(len(b1), KIND_ABSENT, np.int64)
        kind[side] = np.where(rank[side] < 0, KIND_DENSE, rank[side])
        kind[diag] = KIND_LTM if which == "L" else KIND_UTM
        size = np.where(diag, n1 * (n1 + 1) // 2, np.where(side, nv, 0))
        offs = np.concatenate([[0], np.cumsum(size)[:-1]]).astype(np.uint64)
        buf = np.empty(int(size.sum()))
        # off-diagonal leaves: contiguous runs copied in one go
        sel = np.flatnonzero(side & (size > 0))
        for l in sel:
            o, d, s = int(off[l]), int(offs[l]), int(size[l])
            buf[d:d + s] = data[o:o + s]
        out[which] = dict(b1=b1.astype(np.uint32), n1=n1.astype(np.uint32), b2=b2.astype(np.uint32),
                          n2=n2.astype(np.uint32), kind=kind.astype(np.int32), offset=offs, data=buf)
    for l in np.flatnonzero(diag):
        n, o = int(n1[l]), int(off[l])
        D = np.asarray(data[o:o + n * n]).reshape(n, n).T
        Lp, Up = pack_lu(D)
        dL, dU = int(out["L"]["offset"][l]), int(out["U"]["offset"][l])
        out["L"]["data"][dL:dL + len(Lp)] = Lp
        out["U"]["data"][dU:dU + len(Up)] = Up
    return out["L"], out["U"]


def dense_factor(layout: dict, N: int, which: str) -> np.ndarray:
    """The N x N matrix a factor layout represents: 'L' gives P^T L with the
    unit lower triangles of the diagonal leaves row-permuted as
    mblock::ltr_solve inverts them (H/mblock_D.cpp:1183-1205), 'U' the upper
    factor.  For tests on small N."""
    A = np.zeros((N, N))
    data = layout["data"]
    for l in range(len(layout["kind"])):
        k = int(layout["kind"][l])
        b1, n1, b2, n2, o = (int(layout[f][l]) for f in ("b1", "n1", "b2", "n2", "offset"))
        if k == KIND_ABSENT:
            continue
        if k == KIND_DENSE:
            A[b1:b1 + n1, b2:b2 + n2] = data[o:o + n1 * n2].reshape(n2, n1).T
        elif k >= 0:
            U = data[o:o + n1 * k].reshape(k, n1).T
            V = data[o + n1 * k:o + (n1 + n2) * k].reshape(k, n2).T
            A[b1:b1 + n1, b2:b2 + n2] = U @ V.T
        elif k == KIND_UTM:
            T = np.zeros((n1, n1))
            p = o
            for j in range(n1):
                T[: j + 1, j] = data[p:p + j + 1]
                p += j + 1
            A[b1:b1 + n1, b1:b1 + n1] = T
        else:
            T = np.eye(n1)
            ip = np.empty(n1, np.int64)
            p = o
            for j in range(n1):
                ip[j] = int(data[p])
                T[j + 1:, j] = data[p + 1:p + n1 - j]
                p += n1 - j
            # ltr_solve: z_i = b[ip_i], z = T^{-1} z  =>  b = P^T T z with (P b)_i = b[ip_i]
            M = np.zeros((n1, n1))
            M[ip, :] = T
            A[b1:b1 + n1, b1:b1 + n1] = M
    return A
