"""CPU oracle for the AHMED H-matrix hot path -- TEST INFRASTRUCTURE ONLY.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import this package; it is the checker, never the product.

``Oracle`` wraps ``liboracle.so`` (clean-room C restatement, oracle.c).
``Reference`` wraps ``_ref/libahmedref.so`` (the reference's own sources,
built in place by ``make -C oracle ref``) where that library exists.
Both expose the same calls so tests can run one against the other.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "liboracle.so")
REFLIB = os.path.join(HERE, "_ref", "libahmedref.so")
REFSRC = "/root/reference"

TREE_BBX, TREE_PCA = 0, 1
TREE_SYM = 0x100   # or'ed into a tree kind: symmetric block tree (subdivide_sym)
HEM = -5           # rank code of a packed Hermitian diagonal leaf (mblock::isHeM)
KERNEL_LAPLACE, KERNEL_YUKAWA, KERNEL_EXPR = 0, 1, 2

_u = C.c_uint
_d = C.c_double
_p = C.c_void_p


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and the reference library where its sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "oracle"], check=True)
    if ref and os.path.isdir(REFSRC):
        subprocess.run(["make", "-s", "-C", HERE, "ref"], check=True)


def _ptr(a):
    return a.ctypes.data_as(_p)


class _Tree:
    def __init__(self, api, handle, N, sym=False):
        self.api, self.h, self.N, self.sym = api, handle, N, sym

    def __del__(self):
        try:
            self.api._tree_free(self.h)
        except Exception:
            pass

    def perm(self):
        op = np.empty(self.N, np.uint32)
        po = np.empty(self.N, np.uint32)
        self.api._tree_perm(self.h, _ptr(op), _ptr(po))
        return op, po

    def leaves(self):
        nl = self.api._tree_nleaves(self.h)
        b1, n1, b2, n2, i0 = (np.empty(nl, np.uint32) for _ in range(5))
        adm = np.empty(nl, np.uint8)
        self.api._tree_leaves(self.h, _ptr(b1), _ptr(n1), _ptr(b2), _ptr(n2), _ptr(adm), _ptr(i0))
        return dict(b1=b1, n1=n1, b2=b2, n2=n2, adm=adm, i0=i0)


class _HMat:
    """H-matrix handle: leaves in idx order, mblock data layout."""

    def __init__(self, api, handle, N, tree=None, sym=False):
        self.api, self.h, self.N, self.tree, self.sym = api, handle, N, tree, sym

    def __del__(self):
        try:
            self.api._hmat_free(self.h)
        except Exception:
            pass

    def layout(self):
        return self.api._layout(self)

    def matvec(self, x, d=1.0, y=None):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.N) if y is None else np.ascontiguousarray(y, np.float64).copy()
        f = self.api._matvec_sym if self.sym else self.api._matvec
        f(self.h, _d(d), _ptr(x), _ptr(y))
        return y

    def matvec_t(self, x, d=1.0, y=None):
        """y += d A^T x (mltaGeHhVec); x has the row length, y the column length."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(self.N) if y is None else np.ascontiguousarray(y, np.float64).copy()
        f = self.api._matvec_sym if self.sym else self.api._matvec_t   # symmetric: A^T = A
        f(self.h, _d(d), _ptr(x), _ptr(y))
        return y

    def matmat(self, X, d=1.0, Y=None):
        X = np.asfortranarray(X, np.float64)
        p = X.shape[1]
        Y = np.zeros((self.N, p), order="F") if Y is None else np.array(Y, np.float64, order="F", copy=True)
        f = self.api._matmat_sym if self.sym else self.api._matmat
        f(self.h, _d(d), _u(p), _ptr(X), _u(self.N), _ptr(Y), _u(self.N))
        return Y

    def gmres(self, b, x0=None, eps=1e-8, m=30, nsteps=10000):
        b = np.ascontiguousarray(b, np.float64).copy()
        x = np.zeros(self.N) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        e = _d(eps)
        ns = _u(nsteps)
        f = self.api._gmres_sym if self.sym else self.api._gmres
        rc = f(self.h, _ptr(b), _ptr(x), C.byref(e), _u(m), C.byref(ns))
        return x, int(rc), e.value, ns.value

    def cg(self, b, x0=None, eps=1e-8, nsteps=10000):
        b = np.ascontiguousarray(b, np.float64).copy()
        x = np.zeros(self.N) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        e = _d(eps)
        ns = _u(nsteps)
        f = self.api._cg_sym if self.sym else self.api._cg
        rc = f(self.h, _ptr(b), _ptr(x), C.byref(e), C.byref(ns))
        return x, int(rc), e.value, ns.value

    def to_dense(self):
        """Expand every leaf into the N x N matrix (permuted ordering)."""
        L = self.layout()
        A = np.zeros((self.N, self.N))
        data = L["data"]
        for l in range(len(L["b1"])):
            b1, n1, b2, n2, k, o = (int(L[f][l]) for f in ("b1", "n1", "b2", "n2", "rank", "offset"))
            if k == HEM:
                B = np.zeros((n1, n1))
                iu = np.triu_indices(n1)
                # packed upper triangle, column by column: (i, j) at i + j (j + 1) / 2
                B[iu] = data[o + iu[0] + iu[1] * (iu[1] + 1) // 2]
                A[b1:b1 + n1, b2:b2 + n2] = B + np.triu(B, 1).T
            elif k < 0:
                A[b1:b1 + n1, b2:b2 + n2] = data[o:o + n1 * n2].reshape(n2, n1).T
            elif k > 0:
                U = data[o:o + n1 * k].reshape(k, n1).T
                V = data[o + n1 * k:o + (n1 + n2) * k].reshape(k, n2).T
                A[b1:b1 + n1, b2:b2 + n2] = U @ V.T
            if self.sym and b1 != b2 and k != 0:
                A[b2:b2 + n2, b1:b1 + n1] = A[b1:b1 + n1, b2:b2 + n2].T
        return A


# factor leaf kinds (oracle/ref_harness.cpp factor_kind): >= 0 low rank
KIND_ABSENT, KIND_DENSE, KIND_LTM, KIND_UTM = -2, -1, -3, -4


class _Factors:
    """L and U of the reference's HLU on a tree (leaves in idx order)."""

    def __init__(self, api, handle, tree):
        self.api, self.h, self.tree, self.N = api, handle, tree, tree.N

    def __del__(self):
        try:
            self.api.lib.ref_factors_free(self.h)
        except Exception:
            pass

    def layout(self, which):
        """which 'L' or 'U': dict(b1, n1, b2, n2, kind, offset, data); kind per
        leaf: -2 absent, -1 dense, -3 packed unit lower with the row permutation
        on its diagonal, -4 packed upper, k >= 0 low rank."""
        lv = self.tree.leaves()
        nl = len(lv["b1"])
        kind = np.empty(nl, np.int32)
        off = np.empty(nl, np.uint64)
        w = 0 if which == "L" else 1
        nv = self.api.lib.ref_factor_layout(self.h, w, _ptr(kind), _ptr(off))
        data = np.empty(nv)
        self.api.lib.ref_factor_data(self.h, w, _ptr(data))
        return dict(b1=lv["b1"], n1=lv["n1"], b2=lv["b2"], n2=lv["n2"], kind=kind, offset=off, data=data)

    def solve(self, which, B):
        """which 'L' (LtHGeM_solve), 'U' (UtHGeM_solve), 'LU' (HLU_solve per
        column), 'Lt' (LtHhGeM_solve), 'Ut' (UtHhGeM_solve), 'UtU'
        (HCholesky_solve with U) or 'LUt' (HLUh_solve): returns X with
        op X = B; B is a vector or N x p."""
        B = np.asarray(B, np.float64)
        vec = B.ndim == 1
        X = np.array(B.reshape(self.N, -1), order="F", copy=True)
        w = {"L": 0, "U": 1, "LU": 2, "Lt": 3, "Ut": 4, "UtU": 5, "LUt": 6}[which]
        self.api.lib.ref_trisolve(self.h, w, _u(X.shape[1]), _ptr(X), _u(self.N))
        return X[:, 0].copy() if vec else X

    def gmres(self, H, b, x0=None, eps=1e-8, m=30, nsteps=10000):
        """GMRes with HLU_solve as right preconditioner (solvers/GMRES.cpp)."""
        b = np.ascontiguousarray(b, np.float64).copy()
        x = np.zeros(self.N) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        e = _d(eps)
        ns = _u(nsteps)
        rc = self.api.lib.ref_gmres_hlu(H.h, self.h, _ptr(b), _ptr(x), C.byref(e), _u(m), C.byref(ns))
        return x, int(rc), e.value, ns.value


class Oracle:
    """ctypes binding of liboracle.so."""

    def __init__(self, path: str = LIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle`")
        L = self.lib = C.CDLL(path, mode=C.RTLD_GLOBAL)  # MKL kernels resolve against libmkl_core
        L.or_tree_build.restype = _p
        L.or_tree_build.argtypes = [_p, _u, _u, _d, C.c_int]
        L.or_tree_free.argtypes = [_p]
        L.or_tree_nleaves.restype = _u
        L.or_tree_nleaves.argtypes = [_p]
        L.or_tree_nclusters.restype = _u
        L.or_tree_nclusters.argtypes = [_p]
        L.or_tree_perm.argtypes = [_p, _p, _p]
        L.or_tree_leaves.argtypes = [_p] * 7
        L.or_tree_clusters.argtypes = [_p] * 6
        L.or_kernel_block.argtypes = [C.c_int, _d, _d, _p, _p, _u, _u, _u, _u, _p]
        L.or_kernel_scale.restype = _d
        L.or_kernel_scale.argtypes = [C.c_int, _d, _d, _p, _p, _u, _u, _u, _u]
        L.or_aca_dense.restype = C.c_int
        L.or_aca_dense.argtypes = [_p, _u, _u, _d, _u, _u, _d, C.POINTER(_u), _p, _p]
        L.or_acar_dense.restype = C.c_int
        L.or_acar_dense.argtypes = [_p, _u, _u, _d, _u, _u, _d, C.POINTER(_u), _p, _p]
        L.or_hmat_assemble_method.restype = _p
        L.or_hmat_assemble_method.argtypes = [_p, _p, C.c_int, _d, _d, _d, _u, C.c_int, C.c_int]
        L.or_addtrll.argtypes = [_u, _u, _u, _p, _p, _d, _u, C.POINTER(_u), _p, _p]
        L.or_hmat_assemble.restype = _p
        L.or_hmat_assemble.argtypes = [_p, _p, C.c_int, _d, _d, _d, _u, C.c_int]
        L.or_hmat_from_leaves.restype = _p
        L.or_hmat_from_leaves.argtypes = [_u, _u, _p, _p, _p, _p, _p, _p, _p]
        L.or_hmat_free.argtypes = [_p]
        L.or_hmat_nleaves.restype = _u
        L.or_hmat_nleaves.argtypes = [_p]
        L.or_hmat_nvals.restype = C.c_size_t
        L.or_hmat_nvals.argtypes = [_p]
        L.or_hmat_leaves.argtypes = [_p] * 7
        L.or_hmat_data.restype = C.POINTER(C.c_double)
        L.or_hmat_data.argtypes = [_p]
        L.or_hmat_matvec.restype = C.c_int
        L.or_hmat_matvec.argtypes = [_p, _d, _p, _p]
        L.or_hmat_matvec_t.restype = C.c_int
        L.or_hmat_matvec_t.argtypes = [_p, _d, _p, _p]
        L.or_hmat_matvec_omp.restype = C.c_int
        L.or_hmat_matvec_omp.argtypes = [_p, _d, _p, _p, C.c_int]
        L.or_hmat_matmat.restype = C.c_int
        L.or_hmat_matmat.argtypes = [_p, _d, _u, _p, _u, _p, _u]
        L.or_gmres.restype = _u
        L.or_gmres.argtypes = [_p, _p, _p, C.POINTER(_d), _u, C.POINTER(_u), _p, _u]
        L.or_cg.restype = _u
        L.or_cg.argtypes = [_p, _p, _p, C.POINTER(_d), C.POINTER(_u)]
        L.or_gmres_dense.restype = _u
        L.or_gmres_dense.argtypes = [_u, _p, _p, _p, C.POINTER(_d), _u, C.POINTER(_u)]
        L.or_cg_dense.restype = _u
        L.or_cg_dense.argtypes = [_u, _p, _p, _p, C.POINTER(_d), C.POINTER(_u)]
        L.or_blas_version.argtypes = [C.c_char_p, C.c_int]
        L.or_trih_solve.restype = C.c_int
        L.or_trih_solve.argtypes = [_u, C.c_char, _u] + [_p] * 7 + [_u, _p, _u]
        self._tree_free = L.or_tree_free
        self._tree_perm = L.or_tree_perm
        self._tree_nleaves = L.or_tree_nleaves
        self._tree_leaves = L.or_tree_leaves
        self._hmat_free = L.or_hmat_free
        self._matvec = L.or_hmat_matvec
        self._matvec_t = L.or_hmat_matvec_t
        self._matmat = L.or_hmat_matmat
        self._cg = L.or_cg

    def _gmres(self, h, b, x, e, m, ns):
        return self.lib.or_gmres(h, b, x, e, m, ns, None, _u(0))

    def trih_solve(self, F, uplo, B):
        """op^{-1} B for a factor layout F (dict b1, n1, b2, n2, kind, offset,
        data as _Factors.layout), uplo 'L' or 'U'; B a vector or N x p."""
        B = np.asarray(B, np.float64)
        vec = B.ndim == 1
        N = B.shape[0]
        X = np.array(B.reshape(N, -1), order="F", copy=True)
        a = [np.ascontiguousarray(F[f], np.uint32) for f in ("b1", "n1", "b2", "n2")]
        kind = np.ascontiguousarray(F["kind"], np.int32)
        off = np.ascontiguousarray(F["offset"], np.uint64)
        data = np.ascontiguousarray(F["data"], np.float64)
        rc = self.lib.or_trih_solve(_u(N), uplo.encode(), _u(len(kind)), *(_ptr(v) for v in a), _ptr(kind),
                                    _ptr(off), _ptr(data), _u(X.shape[1]), _ptr(X), _u(N))
        if rc != 0:
            raise ValueError("malformed triangular factor")
        return X[:, 0].copy() if vec else X

    def blas_version(self) -> str:
        buf = C.create_string_buffer(256)
        self.lib.or_blas_version(buf, 255)
        return buf.value.decode()

    # -- trees ---------------------------------------------------------------
    def tree(self, pts, bmin, eta2, kind=TREE_BBX):
        pts = np.ascontiguousarray(pts, np.float64)
        h = self.lib.or_tree_build(_ptr(pts), _u(len(pts)), _u(bmin), _d(eta2), kind)
        return _Tree(self, h, len(pts), sym=bool(kind & TREE_SYM))

    def clusters(self, tree):
        nc = self.lib.or_tree_nclusters(tree.h)
        nb, ne, ic = (np.empty(nc, np.uint32) for _ in range(3))
        s0, s1 = np.empty(nc, np.int32), np.empty(nc, np.int32)
        self.lib.or_tree_clusters(tree.h, _ptr(nb), _ptr(ne), _ptr(s0), _ptr(s1), _ptr(ic))
        return dict(nbeg=nb, nend=ne, son0=s0, son1=s1, icom=ic)

    # -- kernels -------------------------------------------------------------
    def kernel_block(self, kernel, kappa, pts, op, b1, n1, b2, n2, weight=1.0):
        pts = np.ascontiguousarray(pts, np.float64)
        op = np.ascontiguousarray(op, np.uint32)
        out = np.empty(n1 * n2)
        self.lib.or_kernel_block(kernel, _d(kappa), _d(weight), _ptr(pts), _ptr(op), b1, n1, b2, n2, _ptr(out))
        return out.reshape(n2, n1).T

    # -- ACA / recompression on explicit data ----------------------------------
    def aca_dense(self, A, eps, kmax, i0=0, scale=None, rows_only=False):
        """ACA (or ACAr with rows_only) on an explicit block: (ok, U, V)."""
        A = np.asfortranarray(A, np.float64)
        n1, n2 = A.shape
        if scale is None:
            scale = abs(A[n1 // 2, n2 // 2])
        U = np.zeros((kmax + 1) * n1)
        V = np.zeros((kmax + 1) * n2)
        k = _u(0)
        f = self.lib.or_acar_dense if rows_only else self.lib.or_aca_dense
        ok = f(_ptr(A), n1, n2, _d(eps), kmax, i0, _d(scale), C.byref(k), _ptr(U), _ptr(V))
        k = k.value
        return bool(ok), U[:k * n1].reshape(k, n1).T.copy(), V[:k * n2].reshape(k, n2).T.copy()

    def addtrll(self, U, V, delta, kgoal=None):
        U = np.asfortranarray(U, np.float64)
        V = np.asfortranarray(V, np.float64)
        n1, k = U.shape
        n2 = V.shape[0]
        kgoal = k if kgoal is None else kgoal
        Uo = np.zeros(n1 * max(k, 1))
        Vo = np.zeros(n2 * max(k, 1))
        kt = _u(0)
        self.lib.or_addtrll(n1, n2, k, _ptr(U), _ptr(V), _d(delta), kgoal, C.byref(kt), _ptr(Uo), _ptr(Vo))
        kt = kt.value
        return Uo[:n1 * kt].reshape(kt, n1).T.copy(), Vo[:n2 * kt].reshape(kt, n2).T.copy()

    # -- H-matrices ----------------------------------------------------------
    def assemble(self, tree, pts, kernel=KERNEL_LAPLACE, kappa=0.0, eps=1e-6, rankmax=1 << 30, nthreads=0,
                 weight=1.0, rows_only=False):
        """apprx_unsym on every leaf (rows_only: ACAr instead of ACA)."""
        pts = np.ascontiguousarray(pts, np.float64)
        h = self.lib.or_hmat_assemble_method(tree.h, _ptr(pts), kernel, _d(kappa), _d(weight), _d(eps),
                                             _u(rankmax), nthreads, 1 if rows_only else 0)
        return _HMat(self, h, tree.N, tree)

    def from_leaves(self, N, b1, n1, b2, n2, rank, offset, data):
        arrs = [np.ascontiguousarray(a, np.uint32) for a in (b1, n1, b2, n2)]
        rank = np.ascontiguousarray(rank, np.int32)
        offset = np.ascontiguousarray(offset, np.uint64)
        data = np.ascontiguousarray(data, np.float64)
        h = self.lib.or_hmat_from_leaves(N, len(rank), *[_ptr(a) for a in arrs], _ptr(rank), _ptr(offset), _ptr(data))
        return _HMat(self, h, N)

    def save(self, H, path, eta, bmin):
        """The reference's savembls (Include/H.h:90-103) of H to path."""
        if not self.lib.ref_hmat_save(H.h, str(path).encode(), _d(eta), _u(bmin)):
            raise OSError(f"ref_hmat_save({path}) failed")

    def load(self, tree, path):
        """The reference's loadmbls (Include/H.h:60-75) onto tree's leaves; (H, eta, bmin)."""
        eta, bmin = _d(0.0), _u(0)
        h = self.lib.ref_hmat_load(tree.h, str(path).encode(), C.byref(eta), C.byref(bmin))
        if not h:
            raise OSError(f"ref_hmat_load({path}) failed")
        return _HMat(self, h, tree.N, tree), eta.value, bmin.value

    def _layout(self, H):
        nl = self.lib.or_hmat_nleaves(H.h)
        b1, n1, b2, n2 = (np.empty(nl, np.uint32) for _ in range(4))
        rank = np.empty(nl, np.int32)
        off = np.empty(nl, np.uint64)
        self.lib.or_hmat_leaves(H.h, _ptr(b1), _ptr(n1), _ptr(b2), _ptr(n2), _ptr(rank), _ptr(off))
        nv = self.lib.or_hmat_nvals(H.h)
        ptr = self.lib.or_hmat_data(H.h)
        data = np.ctypeslib.as_array(ptr, shape=(nv,)).copy() if nv else np.zeros(0)
        return dict(b1=b1, n1=n1, b2=b2, n2=n2, rank=rank, offset=off, data=data)

    def matvec_omp(self, H, x, d=1.0, nthreads=0):
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(H.N)
        self.lib.or_hmat_matvec_omp(H.h, _d(d), _ptr(x), _ptr(y), nthreads)
        return y

    def gmres_dense(self, A, b, eps=1e-8, m=30, nsteps=10000):
        A = np.asfortranarray(A, np.float64)
        n = A.shape[0]
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(n)
        e, ns = _d(eps), _u(nsteps)
        rc = self.lib.or_gmres_dense(n, _ptr(A), _ptr(b), _ptr(x), C.byref(e), m, C.byref(ns))
        return x, int(rc), e.value, ns.value

    def cg_dense(self, A, b, eps=1e-8, nsteps=10000):
        A = np.asfortranarray(A, np.float64)
        n = A.shape[0]
        b = np.ascontiguousarray(b, np.float64)
        x = np.zeros(n)
        e, ns = _d(eps), _u(nsteps)
        rc = self.lib.or_cg_dense(n, _ptr(A), _ptr(b), _ptr(x), C.byref(e), C.byref(ns))
        return x, int(rc), e.value, ns.value


class Reference:
    """ctypes binding of the compiled reference (oracle/_ref/libahmedref.so)."""

    def __init__(self, path: str = REFLIB):
        if not os.path.exists(path):
            raise FileNotFoundError(f"{path} missing: run `make -C oracle ref`")
        L = self.lib = C.CDLL(path, mode=C.RTLD_GLOBAL)  # MKL kernels resolve against libmkl_core
        L.ref_tree_build.restype = _p
        L.ref_tree_build.argtypes = [_p, _u, _u, _d, C.c_int]
        L.ref_tree_free.argtypes = [_p]
        L.ref_tree_nleaves.restype = _u
        L.ref_tree_nleaves.argtypes = [_p]
        L.ref_tree_perm.argtypes = [_p, _p, _p]
        L.ref_tree_leaves.argtypes = [_p] * 7
        L.ref_hmat_assemble.restype = _p
        L.ref_hmat_assemble.argtypes = [_p, C.c_int, _d, _d, _d, _u]
        L.ref_hmat_free.argtypes = [_p]
        L.ref_hmat_from_data.restype = _p
        L.ref_hmat_from_data.argtypes = [_p, _p, _p, _p]
        L.ref_hmat_layout.restype = C.c_size_t
        L.ref_hmat_layout.argtypes = [_p, _p, _p]
        L.ref_hmat_data.argtypes = [_p, _p]
        L.ref_hmat_matvec.restype = C.c_int
        L.ref_hmat_matvec.argtypes = [_p, _d, _p, _p]
        L.ref_hmat_matvec_t.restype = C.c_int
        L.ref_hmat_matvec_t.argtypes = [_p, _d, _p, _p]
        L.ref_hmat_matmat.restype = C.c_int
        L.ref_hmat_matmat.argtypes = [_p, _d, _u, _p, _u, _p, _u]
        L.ref_gmres.restype = _u
        L.ref_gmres.argtypes = [_p, _p, _p, C.POINTER(_d), _u, C.POINTER(_u)]
        L.ref_hmat_matvec_omp.restype = C.c_int
        L.ref_hmat_matvec_omp.argtypes = [_p, _d, _p, _p, C.c_int]
        L.ref_gmres_omp.restype = _u
        L.ref_gmres_omp.argtypes = [_p, _p, _p, C.POINTER(_d), _u, C.POINTER(_u), C.c_int]
        L.ref_cg.restype = _u
        L.ref_cg.argtypes = [_p, _p, _p, C.POINTER(_d), C.POINTER(_u)]
        L.ref_aca_dense.restype = C.c_int
        L.ref_aca_dense.argtypes = [_p, _u, _u, _d, _u, _u, _d, C.POINTER(_u), _p, _p]
        L.ref_acar_dense.restype = C.c_int
        L.ref_acar_dense.argtypes = [_p, _u, _u, _d, _u, _u, _d, C.POINTER(_u), _p, _p]
        L.ref_hmat_assemble_omp.restype = _p
        L.ref_hmat_assemble_omp.argtypes = [_p, C.c_int, _d, _d, _d, _u, _u, C.c_int, C.POINTER(_d), _p]
        L.ref_hmat_assemble_acar.restype = _p
        L.ref_hmat_assemble_acar.argtypes = [_p, C.c_int, _d, _d, _d, _u]
        L.ref_addtrll.argtypes = [_u, _u, _u, _p, _p, _d, _u, C.POINTER(_u), _p, _p]
        L.ref_hmat_save.restype = C.c_int
        L.ref_hmat_save.argtypes = [_p, C.c_char_p, _d, _u]
        L.ref_hmat_load.restype = _p
        L.ref_hmat_load.argtypes = [_p, C.c_char_p, C.POINTER(_d), C.POINTER(_u)]
        L.ref_hlu.restype = _p
        L.ref_hlu.argtypes = [_p, _d, _u]
        L.ref_factors_free.argtypes = [_p]
        L.ref_factor_layout.restype = C.c_size_t
        L.ref_factor_layout.argtypes = [_p, C.c_int, _p, _p]
        L.ref_factor_data.argtypes = [_p, C.c_int, _p]
        L.ref_factors_from_data.restype = _p
        L.ref_factors_from_data.argtypes = [_p] * 7
        L.ref_trisolve.argtypes = [_p, C.c_int, _u, _p, _u]
        L.ref_gmres_hlu.restype = _u
        L.ref_gmres_hlu.argtypes = [_p, _p, _p, _p, C.POINTER(_d), _u, C.POINTER(_u)]
        L.ref_hmat_assemble_sym.restype = _p
        L.ref_hmat_assemble_sym.argtypes = [_p, C.c_int, _d, _d, _d, _u]
        L.ref_hmat_matvec_sym.restype = C.c_int
        L.ref_hmat_matvec_sym.argtypes = [_p, _d, _p, _p]
        L.ref_cg_sym.restype = _u
        L.ref_cg_sym.argtypes = [_p, _p, _p, C.POINTER(_d), C.POINTER(_u)]
        L.ref_gmres_sym.restype = _u
        L.ref_gmres_sym.argtypes = [_p, _p, _p, C.POINTER(_d), _u, C.POINTER(_u)]
        L.ref_hmat_matmat_sym.restype = C.c_int
        L.ref_hmat_matmat_sym.argtypes = [_p, _d, _u, _p, _u, _p, _u]
        self._matmat_sym = L.ref_hmat_matmat_sym
        self._matvec_sym = L.ref_hmat_matvec_sym
        self._cg_sym = L.ref_cg_sym
        self._gmres_sym = L.ref_gmres_sym
        self._tree_free = L.ref_tree_free
        self._tree_perm = L.ref_tree_perm
        self._tree_nleaves = L.ref_tree_nleaves
        self._tree_leaves = L.ref_tree_leaves
        self._hmat_free = L.ref_hmat_free
        self._matvec = L.ref_hmat_matvec
        self._matvec_t = L.ref_hmat_matvec_t
        self._matmat = L.ref_hmat_matmat
        self._gmres = L.ref_gmres
        self._cg = L.ref_cg

    def tree(self, pts, bmin, eta2, kind=TREE_BBX):
        pts = np.ascontiguousarray(pts, np.float64)
        h = self.lib.ref_tree_build(_ptr(pts), _u(len(pts)), _u(bmin), _d(eta2), kind)
        return _Tree(self, h, len(pts), sym=bool(kind & TREE_SYM))

    def assemble(self, tree, pts=None, kernel=KERNEL_LAPLACE, kappa=0.0, eps=1e-6, rankmax=1 << 30, nthreads=0,
                 weight=1.0, rows_only=False):
        """matgenGeH_omp's leaf rule (rows_only: ACAr); on a symmetric tree
        (kind | TREE_SYM) the symmetric assembly matgenHeH_omp / apprx_sym."""
        if tree.sym:
            h = self.lib.ref_hmat_assemble_sym(tree.h, kernel, _d(kappa), _d(weight), _d(eps), _u(rankmax))
            return _HMat(self, h, tree.N, tree, sym=True)
        f = self.lib.ref_hmat_assemble_acar if rows_only else self.lib.ref_hmat_assemble
        h = f(tree.h, kernel, _d(kappa), _d(weight), _d(eps), _u(rankmax))
        return _HMat(self, h, tree.N, tree)

    def assemble_omp(self, tree, kernel=KERNEL_LAPLACE, kappa=0.0, eps=1e-6, weight=1.0, stride=1, nthreads=0,
                     rankmax=1 << 30):
        """matgenGeH_omp's OpenMP leaf loop (Include/matgen_omp.h:59-75) over
        every stride-th leaf of the leaf list: (H, seconds of the loop, sampled
        mask by leaf idx).  Leaves outside the sample are rank 0."""
        nthreads = nthreads or len(os.sched_getaffinity(0))
        sec = _d(0.0)
        mask = np.zeros(self.lib.ref_tree_nleaves(tree.h), np.uint8)
        h = self.lib.ref_hmat_assemble_omp(tree.h, kernel, _d(kappa), _d(weight), _d(eps), _u(rankmax),
                                           _u(stride), nthreads, C.byref(sec), _ptr(mask))
        return _HMat(self, h, tree.N, tree), sec.value, mask.astype(bool)

    def matvec_omp(self, H, x, d=1.0, nthreads=0):
        """y = d A x with the reference's leaf products on nthreads cores
        (private y per thread, summed)."""
        x = np.ascontiguousarray(x, np.float64)
        y = np.zeros(H.N)
        self.lib.ref_hmat_matvec_omp(H.h, _d(d), _ptr(x), _ptr(y), nthreads or len(os.sched_getaffinity(0)))
        return y

    def gmres_omp(self, H, b, x0=None, eps=1e-8, m=30, nsteps=10000, nthreads=0):
        """The reference's GMRes with the leaf-parallel operator on nthreads cores."""
        b = np.ascontiguousarray(b, np.float64).copy()
        x = np.zeros(H.N) if x0 is None else np.ascontiguousarray(x0, np.float64).copy()
        e, ns = _d(eps), _u(nsteps)
        rc = self.lib.ref_gmres_omp(H.h, _ptr(b), _ptr(x), C.byref(e), _u(m), C.byref(ns),
                                    nthreads or len(os.sched_getaffinity(0)))
        return x, int(rc), e.value, ns.value

    def from_data(self, tree, rank, offset, data):
        """The reference's mblock array filled with given leaf values (idx order)."""
        rank = np.ascontiguousarray(rank, np.int32)
        offset = np.ascontiguousarray(offset, np.uint64)
        data = np.ascontiguousarray(data, np.float64)
        h = self.lib.ref_hmat_from_data(tree.h, _ptr(rank), _ptr(offset), _ptr(data))
        return _HMat(self, h, tree.N, tree, sym=tree.sym)

    def save(self, H, path, eta, bmin):
        """The reference's savembls (Include/H.h:90-103) of H to path."""
        if not self.lib.ref_hmat_save(H.h, str(path).encode(), _d(eta), _u(bmin)):
            raise OSError(f"ref_hmat_save({path}) failed")

    def load(self, tree, path):
        """The reference's loadmbls (Include/H.h:60-75) onto tree's leaves; (H, eta, bmin)."""
        eta, bmin = _d(0.0), _u(0)
        h = self.lib.ref_hmat_load(tree.h, str(path).encode(), C.byref(eta), C.byref(bmin))
        if not h:
            raise OSError(f"ref_hmat_load({path}) failed")
        return _HMat(self, h, tree.N, tree), eta.value, bmin.value

    def hlu(self, H, eps, rankmax=1 << 30):
        """The reference's HLU (H/HLU.cpp:389-395) of a copy of H: _Factors."""
        h = self.lib.ref_hlu(H.h, _d(eps), _u(rankmax))
        if not h:
            raise RuntimeError("reference HLU failed")
        return _Factors(self, h, H.tree)

    def factors_from_data(self, tree, Lf, Uf):
        """L/U mblock arrays rebuilt from factor layouts (as _Factors.layout)."""
        keep = []

        def arr(a, t):
            a = np.ascontiguousarray(a, t)
            keep.append(a)
            return _ptr(a)
        h = self.lib.ref_factors_from_data(
            tree.h, arr(Lf["kind"], np.int32), arr(Lf["offset"], np.uint64), arr(Lf["data"], np.float64),
            arr(Uf["kind"], np.int32), arr(Uf["offset"], np.uint64), arr(Uf["data"], np.float64))
        return _Factors(self, h, tree)

    def _layout(self, H):
        lv = H.tree.leaves()
        nl = len(lv["b1"])
        rank = np.empty(nl, np.int32)
        off = np.empty(nl, np.uint64)
        nv = self.lib.ref_hmat_layout(H.h, _ptr(rank), _ptr(off))
        data = np.empty(nv)
        self.lib.ref_hmat_data(H.h, _ptr(data))
        return dict(b1=lv["b1"], n1=lv["n1"], b2=lv["b2"], n2=lv["n2"], rank=rank, offset=off, data=data)

    def aca_dense(self, A, eps, kmax, i0=0, scale=None, rows_only=False):
        A = np.asfortranarray(A, np.float64)
        n1, n2 = A.shape
        if scale is None:
            scale = abs(A[n1 // 2, n2 // 2])
        U = np.zeros((kmax + 1) * n1)
        V = np.zeros((kmax + 1) * n2)
        k = _u(0)
        f = self.lib.ref_acar_dense if rows_only else self.lib.ref_aca_dense
        ok = f(_ptr(A), n1, n2, _d(eps), kmax, i0, _d(scale), C.byref(k), _ptr(U), _ptr(V))
        k = k.value
        return bool(ok), U[:k * n1].reshape(k, n1).T.copy(), V[:k * n2].reshape(k, n2).T.copy()

    def addtrll(self, U, V, delta, kgoal=None):
        U = np.asfortranarray(U, np.float64)
        V = np.asfortranarray(V, np.float64)
        n1, k = U.shape
        n2 = V.shape[0]
        kgoal = k if kgoal is None else kgoal
        Uo = np.zeros(n1 * max(k, 1))
        Vo = np.zeros(n2 * max(k, 1))
        kt = _u(0)
        self.lib.ref_addtrll(n1, n2, k, _ptr(U), _ptr(V), _d(delta), kgoal, C.byref(kt), _ptr(Uo), _ptr(Vo))
        kt = kt.value
        return Uo[:n1 * kt].reshape(kt, n1).T.copy(), Vo[:n2 * kt].reshape(kt, n2).T.copy()
