"""ctypes binding of libahmed_amd.so (the C ABI declared in include/ahm.h).

The HIP library is the product; there is no CPU fallback.  ``load()`` raises
if the shared library is missing.  Device vectors are torch tensors on the
library context's device; every call runs on torch's current HIP stream.
"""
from __future__ import annotations

import ctypes as C
import os
import weakref

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libahmed_amd.so")

F64, F32, BF16 = 0, 1, 2
OK = 0
STATUS = {0: "AHM_OK", 1: "AHM_ERR_INVALID", 2: "AHM_ERR_HIP", 3: "AHM_ERR_NOMEM",
          4: "AHM_ERR_UNSUPPORTED", 5: "AHM_ERR_NOT_CONVERGED"}


class LeafDesc(C.Structure):
    _fields_ = [("b1", C.c_uint32), ("n1", C.c_uint32), ("b2", C.c_uint32), ("n2", C.c_uint32),
                ("rank", C.c_int32), ("flags", C.c_uint32), ("offset", C.c_uint64)]


LEAF_DTYPE = np.dtype([("b1", "<u4"), ("n1", "<u4"), ("b2", "<u4"), ("n2", "<u4"),
                       ("rank", "<i4"), ("flags", "<u4"), ("offset", "<u8")])
assert LEAF_DTYPE.itemsize == C.sizeof(LeafDesc)


class HMatStats(C.Structure):
    _fields_ = [("n_leaves", C.c_uint64), ("n_dense", C.c_uint64), ("n_lowrank", C.c_uint64),
                ("nvals", C.c_uint64), ("nvals_dense", C.c_uint64), ("nvals_u", C.c_uint64),
                ("nvals_v", C.c_uint64), ("t_len", C.c_uint64), ("bytes_per_value", C.c_uint64),
                ("n_tasks_a", C.c_uint64), ("n_tasks_b", C.c_uint64), ("max_rank", C.c_uint32),
                ("n_rows", C.c_uint32), ("n_cols", C.c_uint32), ("symmetric", C.c_uint32)]


class Kernel(C.Structure):
    _fields_ = [("kind", C.c_int32), ("flags", C.c_uint32), ("kappa", C.c_double), ("weight", C.c_double)]


class AssemblyStats(C.Structure):
    _fields_ = [("n_candidates", C.c_uint64), ("n_retries", C.c_uint64), ("n_lowrank", C.c_uint64),
                ("n_dense", C.c_uint64), ("sum_rank", C.c_uint64), ("sum_aca_rank", C.c_uint64),
                ("slab_bytes", C.c_uint64), ("peak_slab_bytes", C.c_uint64),
                ("aca_ms", C.c_double)]


KERNEL_LAPLACE, KERNEL_YUKAWA, KERNEL_EXPR = 0, 1, 2
ASSEMBLE_ACAR = 1   # Kernel.flags: ACAr instead of ACA on the admissible leaves
ASSEMBLE_SYM = 2    # Kernel.flags: symmetric storage (set by assemble on AHM_TREE_SYM trees)


class AhmError(RuntimeError):
    pass


_lib = None

# every symbol declared in include/ahm.h, with (restype, argtypes)
_p, _u32, _i, _d = C.c_void_p, C.c_uint32, C.c_int, C.c_double
SIGNATURES = {
    "ahm_last_error": (C.c_char_p, []),
    "ahm_ctx_create": (_i, [_i, C.POINTER(_p)]),
    "ahm_ctx_destroy": (_i, [_p]),
    "ahm_ctx_release_cached": (_i, [_p]),
    "ahm_ctx_cached_bytes": (_i, [_p, C.POINTER(C.c_uint64)]),
    "ahm_ctx_set_stream": (_i, [_p, _p]),
    "ahm_ctx_sync": (_i, [_p]),
    "ahm_tree_build": (_i, [_p, _u32, _u32, _d, _i, C.POINTER(_p)]),
    "ahm_tree_destroy": (_i, [_p]),
    "ahm_tree_info": (_i, [_p, C.POINTER(_u32), C.POINTER(_u32)]),
    "ahm_tree_perm": (_i, [_p, _p, _p]),
    "ahm_tree_leaves": (_i, [_p, _p, _p, _p, _p, _p, _p]),
    "ahm_hmat_upload": (_i, [_p, _u32, _u32, _u32, _p, _p, _i, _i, _i, C.POINTER(_p)]),
    "ahm_hmat_convert": (_i, [_p, _i, C.POINTER(_p)]),
    "ahm_hmat_destroy": (_i, [_p]),
    "ahm_hmat_stats_get": (_i, [_p, C.POINTER(HMatStats)]),
    "ahm_hmat_download": (_i, [_p, _p, _p]),
    "ahm_mbls_save": (_i, [C.c_char_p, _d, _u32, _u32, _p, _p, _i]),
    "ahm_mbls_info": (_i, [C.c_char_p, _i, C.POINTER(_d), C.POINTER(_u32), C.POINTER(_u32),
                           C.POINTER(C.c_uint64)]),
    "ahm_mbls_read": (_i, [C.c_char_p, _i, _u32, _p, _p]),
    "ahm_assemble": (_i, [_p, _p, _p, _u32, _u32, _p, _p, _p, _p, _p, _p, C.POINTER(Kernel), _d, _u32, _p,
                          _i, _i, C.POINTER(AssemblyStats), _p, C.POINTER(_p)]),
    "ahm_assemble_tree": (_i, [_p, _p, C.POINTER(Kernel), _d, _u32, _p, _i, _i, C.POINTER(AssemblyStats),
                               _p, C.POINTER(_p)]),
    "ahm_gmres": (_i, [_p, _p, _p, C.POINTER(_d), _u32, C.POINTER(_u32), _p, _u32, C.POINTER(_u32)]),
    "ahm_cg": (_i, [_p, _p, _p, C.POINTER(_d), C.POINTER(_u32)]),
    "ahm_gmres_ir": (_i, [_p, _p, _p, _p, C.POINTER(_d), _u32, _d, _u32, _u32, C.POINTER(_u32),
                          C.POINTER(_u32), _p, _u32]),
    "ahm_assemble_batch": (_i, [_p, _u32, _p, _p, _d, _u32, _i, _i, C.POINTER(AssemblyStats), C.POINTER(_p)]),
    "ahm_hmat_segments": (_i, [_p, C.POINTER(_u32), _p]),
    "ahm_gmres_batch": (_i, [_p, _p, _p, _d, _u32, _u32, _i, _p, _p, _p]),
    "ahm_hmat_set_graphs": (_i, [_p, _i]),
    "ahm_hmat_set_arnoldi": (_i, [_p, _i]),
    "ahm_rank_predict": (_i, [_p, _p, _u32, _p, C.c_float, _u32, _p, _p, _p]),
    "ahm_hmat_matvec": (_i, [_p, _d, _p, _u32, _p, _u32, _u32, _i]),
    "ahm_hmat_matvec_noalloc": (_i, [_p, _d, _p, _u32, _p, _u32, _u32, _i]),
    "ahm_hmat_autotune": (_i, [_p, _u32, _i, _u32, C.POINTER(C.c_uint64), C.POINTER(_d), C.POINTER(_i)]),
    "ahm_hmat_set_fused": (_i, [_p, _i]),
    "ahm_hmat_fused_status": (_i, [_p, C.POINTER(_i)]),
    "ahm_hmat_matvec_host": (_i, [_p, _d, _p, _u32, _p, _u32, _u32]),
    "ahm_hmat_matvec_t": (_i, [_p, _d, _p, _u32, _p, _u32, _u32, _i]),
    "ahm_hmat_matvec_t_host": (_i, [_p, _d, _p, _u32, _p, _u32, _u32]),
    "ahm_hmat_matvec_profile": (_i, [_p, _d, _p, _u32, _p, _u32, _u32, _i, _u32, C.POINTER(_d),
                                     C.POINTER(_d)]),
    "ahm_trih_upload": (_i, [_p, _u32, C.c_char, _u32, _p, _p, _i, C.POINTER(_p)]),
    "ahm_trih_destroy": (_i, [_p]),
    "ahm_trih_solve": (_i, [_p, _p, _u32, _u32]),
    "ahm_trih_solve_host": (_i, [_p, _p, _u32, _u32]),
    "ahm_trih_info": (_i, [_p, C.POINTER(_u32), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    "ahm_gmres_prec": (_i, [_p, _p, _p, _p, _p, C.POINTER(_d), _u32, C.POINTER(_u32), _p, _u32,
                            C.POINTER(_u32)]),
    "ahm_cg_prec": (_i, [_p, _p, _p, _p, _p, C.POINTER(_d), C.POINTER(_u32)]),
}

LEAF_LTM, LEAF_UTM, LEAF_HEM = 1, 2, 4
HEM = -5   # rank code of packed Hermitian diagonal leaves in layouts (oracle convention)


def factor_leaves(layout) -> tuple:
    """(leaves, data) for ahm_trih_upload from a factor layout: per leaf b1,
    n1, b2, n2, offset and kind (-2 absent, -1 dense, k >= 0 low rank, -3
    packed unit lower with the row permutation on its diagonal, -4 packed
    upper), the mblock storage of the reference's L and U arrays."""
    kind = np.asarray(layout["kind"])
    keep = kind != -2
    L = np.zeros(int(keep.sum()), LEAF_DTYPE)
    for f in ("b1", "n1", "b2", "n2", "offset"):
        L[f] = np.asarray(layout[f])[keep]
    k = kind[keep]
    L["rank"] = np.where(k >= 0, k, -1)
    L["flags"] = np.where(k == -3, LEAF_LTM, np.where(k == -4, LEAF_UTM, 0))
    return L, np.ascontiguousarray(layout["data"], np.float64)


def load():
    """Load libahmed_amd.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise AhmError(f"{LIB_PATH} is missing: run __graft_entry__.build() "
                           "(make -C ahmed_amd/csrc)")
        # torch ships its own HIP runtime; loading it first makes this library
        # bind to the same runtime instance (same SONAME), so device pointers
        # and streams from torch are valid here
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        lib = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(lib, name)
            f.restype = res
            f.argtypes = args
        _lib = lib
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != OK:
        msg = load().ahm_last_error().decode(errors="replace")
        raise AhmError(f"{what}: {STATUS.get(rc, rc)}: {msg}")


def _torch():
    import torch
    return torch


class Context:
    """One device + the torch stream current at each call."""

    def __init__(self, device: int = 0):
        lib = load()
        self.device = device
        h = _p()
        check(lib.ahm_ctx_create(device, C.byref(h)), "ahm_ctx_create")
        self.h = h
        # handles made on this context: closed before the context itself,
        # since every handle refers to it
        self._children = weakref.WeakSet()

    def bind_stream(self):
        torch = _torch()
        s = torch.cuda.current_stream(self.device).cuda_stream
        check(load().ahm_ctx_set_stream(self.h, _p(s)), "ahm_ctx_set_stream")

    def sync(self):
        check(load().ahm_ctx_sync(self.h), "ahm_ctx_sync")

    def release_cached(self):
        """Free the cached device blocks (>= 1 GiB) of this device."""
        check(load().ahm_ctx_release_cached(self.h), "ahm_ctx_release_cached")

    def cached_bytes(self) -> int:
        """Bytes of released device blocks the allocation cache holds."""
        v = C.c_uint64(0)
        check(load().ahm_ctx_cached_bytes(self.h, C.byref(v)), "ahm_ctx_cached_bytes")
        return int(v.value)

    def close(self):
        if self.h:
            for c in list(getattr(self, "_children", ())):
                c.close()
            load().ahm_ctx_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


TREE_BBX, TREE_PCA = 0, 1
TREE_SYM = 0x100   # or'ed into a kind: symmetric block tree (subdivide_sym)


def mbls_save(path, eta: float, bmin: int, leaves: np.ndarray, data: np.ndarray) -> None:
    """Write leaves (LEAF_DTYPE, idx order) and their mblock-layout values in
    the reference's savembls format (Include/H.h:90-103)."""
    leaves = np.ascontiguousarray(leaves, LEAF_DTYPE)
    dt = F64 if data.dtype == np.float64 else F32
    data = np.ascontiguousarray(data)
    check(load().ahm_mbls_save(os.fsencode(path), eta, bmin, len(leaves), leaves.ctypes.data_as(_p),
                               data.ctypes.data_as(_p), dt), "ahm_mbls_save")


def mbls_read(path, dtype=np.float64):
    """Read a savembls file: (eta, bmin, leaves with n1/n2/rank/offset, values)."""
    dt = F64 if np.dtype(dtype) == np.float64 else F32
    eta, bmin, nl, nv = _d(), _u32(), _u32(), C.c_uint64()
    check(load().ahm_mbls_info(os.fsencode(path), dt, C.byref(eta), C.byref(bmin), C.byref(nl),
                               C.byref(nv)), "ahm_mbls_info")
    leaves = np.zeros(nl.value, LEAF_DTYPE)
    data = np.empty(nv.value, dtype)
    check(load().ahm_mbls_read(os.fsencode(path), dt, nl.value, leaves.ctypes.data_as(_p),
                               data.ctypes.data_as(_p)), "ahm_mbls_read")
    return eta.value, bmin.value, leaves, data


class Tree:
    """Cluster tree + block cluster tree built by the C++ surface (host)."""

    def __init__(self, pts, bmin: int, eta2: float, kind: int = TREE_BBX):
        pts = np.ascontiguousarray(pts, np.float64)
        self.N = len(pts)
        self.sym = bool(kind & TREE_SYM)
        h = _p()
        check(load().ahm_tree_build(pts.ctypes.data_as(_p), self.N, bmin, eta2, kind, C.byref(h)),
              "ahm_tree_build")
        self.h = h
        nl, nc = _u32(), _u32()
        check(load().ahm_tree_info(h, C.byref(nl), C.byref(nc)), "ahm_tree_info")
        self.n_leaves, self.n_clusters = nl.value, nc.value

    def perm(self):
        op = np.empty(self.N, np.uint32)
        po = np.empty(self.N, np.uint32)
        check(load().ahm_tree_perm(self.h, op.ctypes.data_as(_p), po.ctypes.data_as(_p)), "perm")
        return op, po

    def leaves(self):
        n = self.n_leaves
        b1, n1, b2, n2, i0 = (np.empty(n, np.uint32) for _ in range(5))
        adm = np.empty(n, np.uint8)
        check(load().ahm_tree_leaves(self.h, *[a.ctypes.data_as(_p) for a in (b1, n1, b2, n2, adm, i0)]),
              "leaves")
        return dict(b1=b1, n1=n1, b2=b2, n2=n2, adm=adm, i0=i0)

    def close(self):
        if self.h:
            load().ahm_tree_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def leaf_array(b1, n1, b2, n2, rank, offset) -> np.ndarray:
    """Leaf descriptors from a layout; rank code HEM (-5) marks packed
    Hermitian diagonal leaves (rank -1 with flag LEAF_HEM)."""
    a = np.zeros(len(rank), LEAF_DTYPE)
    rank = np.asarray(rank)
    a["b1"], a["n1"], a["b2"], a["n2"] = b1, n1, b2, n2
    a["rank"], a["offset"] = np.where(rank == HEM, -1, rank), offset
    a["flags"] = np.where(rank == HEM, LEAF_HEM, 0)
    return a


class TriFactor:
    """Triangular H-matrix factor on the device (ahm_trih): uplo 'L' (unit
    lower, LtHGeM_solve), 'U' (upper, UtHGeM_solve), or 'l' / 'u' for solves
    with the transpose of that factor (LtHhGeM_solve / UtHhGeM_solve)."""

    def __init__(self, ctx: Context, handle, n: int, uplo: str):
        self.ctx, self.h, self.n, self.uplo = ctx, handle, n, uplo
        ctx._children.add(self)

    @classmethod
    def upload(cls, ctx: Context, n: int, uplo: str, leaves: np.ndarray, data: np.ndarray) -> "TriFactor":
        leaves = np.ascontiguousarray(leaves, LEAF_DTYPE)
        data = np.ascontiguousarray(data, np.float64)
        h = _p()
        ctx.bind_stream()
        check(load().ahm_trih_upload(ctx.h, n, uplo.encode(), len(leaves), leaves.ctypes.data_as(_p),
                                     data.ctypes.data_as(_p), F64, C.byref(h)), "ahm_trih_upload")
        return cls(ctx, h, n, uplo)

    @classmethod
    def from_layout(cls, ctx: Context, n: int, uplo: str, layout) -> "TriFactor":
        leaves, data = factor_leaves(layout)
        return cls.upload(ctx, n, uplo, leaves, data)

    def close(self):
        # a handle whose context is already destroyed (finalisation order at
        # interpreter exit) is dropped without touching the freed context
        if self.h:
            if self.ctx.h:
                load().ahm_trih_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def info(self) -> dict:
        nd, ni, nv = _u32(), C.c_uint64(), C.c_uint64()
        check(load().ahm_trih_info(self.h, C.byref(nd), C.byref(ni), C.byref(nv)), "ahm_trih_info")
        return {"n_diag": nd.value, "n_items": ni.value, "nvals": nv.value}

    def solve(self, B):
        """B <- op^{-1} B in place: a float64 device tensor of n values, or
        (p, n) for p right-hand sides (row j = right-hand side j)."""
        if B.dim() == 1:
            p, ldb = 1, B.shape[0]
        else:
            p, ldb = B.shape[0], B.stride(0)
        self.ctx.bind_stream()
        check(load().ahm_trih_solve(self.h, _p(B.data_ptr()), ldb, p), "ahm_trih_solve")
        return B


class HMatrix:
    """Device H-matrix (ahm_hmat)."""

    def __init__(self, ctx: Context, handle, n_rows: int, n_cols: int, storage: int):
        self.ctx, self.h, self.n_rows, self.n_cols, self.storage = ctx, handle, n_rows, n_cols, storage
        ctx._children.add(self)

    @classmethod
    def upload(cls, ctx: Context, n_rows, n_cols, leaves: np.ndarray, data: np.ndarray,
               storage: int = F64, keep_leaves: bool = False) -> "HMatrix":
        leaves = np.ascontiguousarray(leaves, LEAF_DTYPE)
        ddt = F64 if data.dtype == np.float64 else F32
        data = np.ascontiguousarray(data)
        h = _p()
        ctx.bind_stream()
        check(load().ahm_hmat_upload(ctx.h, n_rows, n_cols, len(leaves), leaves.ctypes.data_as(_p),
                                     data.ctypes.data_as(_p), ddt, storage, int(keep_leaves),
                                     C.byref(h)), "ahm_hmat_upload")
        H = cls(ctx, h, n_rows, n_cols, storage)
        H.leaf_dtype = ddt
        return H

    @classmethod
    def assemble(cls, ctx: Context, tree: "Tree", kernel: int = KERNEL_LAPLACE, kappa: float = 0.0,
                 weight: float = 1.0, eps: float = 1e-6, rankmax: int = 0xFFFFFFFF, rank_hint=None,
                 storage: int = F64, keep_leaves: bool = False, aca_ranks: bool = False,
                 rows_only: bool = False):
        """GPU assembly (ACA + recompression + dense leaves) on a tree; returns
        (HMatrix, stats dict); stats["aca_rank"] holds per-leaf ACA ranks when
        aca_ranks is set; rows_only uses ACAr (apprx_sym's variant)."""
        k = Kernel(kernel, ASSEMBLE_ACAR if rows_only else 0, kappa, weight)
        st = AssemblyStats()
        h = _p()
        hint = None
        if rank_hint is not None:
            hint = np.ascontiguousarray(rank_hint, np.uint32)
        ar = np.zeros(tree.n_leaves, np.uint32) if aca_ranks else None
        ctx.bind_stream()
        check(load().ahm_assemble_tree(ctx.h, tree.h, C.byref(k), eps, rankmax,
                                       hint.ctypes.data_as(_p) if hint is not None else None, storage,
                                       int(keep_leaves), C.byref(st),
                                       ar.ctypes.data_as(_p) if ar is not None else None, C.byref(h)),
              "ahm_assemble_tree")
        out = {f: getattr(st, f) for f, _ in AssemblyStats._fields_}
        if ar is not None:
            out["aca_rank"] = ar
        return cls(ctx, h, tree.N, tree.N, storage), out

    @classmethod
    def assemble_batch(cls, ctx: Context, trees, kernels, eps: float = 1e-6, rankmax: int = 0xFFFFFFFF,
                       storage: int = F64, keep_leaves: bool = False):
        """Independent H-matrices (trees[s] with kernels[s] = (kind, kappa,
        weight)) assembled in one batch and stacked block-diagonally."""
        n = len(trees)
        th = (_p * n)(*[t.h for t in trees])
        ks = (Kernel * n)(*[Kernel(k[0], 0, k[1], k[2]) for k in kernels])
        st = AssemblyStats()
        h = _p()
        ctx.bind_stream()
        check(load().ahm_assemble_batch(ctx.h, n, C.cast(th, _p), C.cast(ks, _p), eps, rankmax, storage,
                                        int(keep_leaves), C.byref(st), C.byref(h)), "ahm_assemble_batch")
        N = sum(t.N for t in trees)
        return cls(ctx, h, N, N, storage), {f: getattr(st, f) for f, _ in AssemblyStats._fields_}

    def segments(self):
        ns = _u32()
        check(load().ahm_hmat_segments(self.h, C.byref(ns), None), "segments")
        off = np.zeros(ns.value + 1, np.uint32)
        check(load().ahm_hmat_segments(self.h, C.byref(ns), off.ctypes.data_as(_p)), "segments")
        return off

    def gmres_batch(self, b, x, eps: float = 1e-8, m: int = 50, nsteps: int = 10000, use_graph: bool = False):
        S = len(self.segments()) - 1
        e = np.zeros(S)
        st = np.zeros(S, np.uint32)
        cv = np.zeros(S, np.int32)
        self.ctx.bind_stream()
        rc = load().ahm_gmres_batch(self.h, _p(b.data_ptr()), _p(x.data_ptr()), eps, m, nsteps, int(use_graph),
                                    e.ctypes.data_as(_p), st.ctypes.data_as(_p), cv.ctypes.data_as(_p))
        if rc not in (OK, 5):
            check(rc, "ahm_gmres_batch")
        return dict(converged=cv.astype(bool), eps=e, nsteps=st)

    def convert(self, storage: int = F32) -> "HMatrix":
        h = _p()
        self.ctx.bind_stream()
        check(load().ahm_hmat_convert(self.h, storage, C.byref(h)), "ahm_hmat_convert")
        return HMatrix(self.ctx, h, self.n_rows, self.n_cols, storage)

    def stats(self) -> dict:
        st = HMatStats()
        check(load().ahm_hmat_stats_get(self.h, C.byref(st)), "ahm_hmat_stats_get")
        return {f: getattr(st, f) for f, _ in HMatStats._fields_}

    def save(self, path, eta: float, bmin: int) -> None:
        """savembls of this matrix (needs keep_leaves): the reference's format."""
        leaves, data = self.download()
        mbls_save(path, eta, bmin, leaves, data)

    @classmethod
    def load(cls, ctx: Context, path, pts, kind: int = TREE_BBX, storage: int = F64,
             keep_leaves: bool = False):
        """loadmbls (Include/H.h:60-75) onto the block cluster tree rebuilt from
        pts with the file's bmin and eta; returns (HMatrix, Tree)."""
        eta, bmin, leaves, data = mbls_read(path)
        tree = Tree(pts, bmin, eta * eta, kind)
        lv = tree.leaves()
        if tree.n_leaves != len(leaves) or not (np.array_equal(lv["n1"], leaves["n1"])
                                                and np.array_equal(lv["n2"], leaves["n2"])):
            tree.close()
            raise AhmError(f"{path}: leaves do not match the tree of these points "
                           f"(bmin {bmin}, eta {eta})")
        leaves["b1"], leaves["b2"] = lv["b1"], lv["b2"]
        return cls.upload(ctx, tree.N, tree.N, leaves, data, storage, keep_leaves), tree

    def leaves(self) -> np.ndarray:
        """Leaf descriptors (idx order; rank -1 = dense), without the values."""
        leaves = np.zeros(self.stats()["n_leaves"], LEAF_DTYPE)
        check(load().ahm_hmat_download(self.h, leaves.ctypes.data_as(_p), None), "download")
        return leaves

    def download(self):
        leaves = self.leaves()
        n1, n2 = leaves["n1"].astype(np.int64), leaves["n2"].astype(np.int64)
        k = leaves["rank"].astype(np.int64)
        hem = (leaves["flags"] & LEAF_HEM) != 0
        nv = np.where(k >= 0, k * (n1 + n2), np.where(hem, n1 * (n1 + 1) // 2, n1 * n2)) + \
            leaves["offset"].astype(np.int64)
        # kept values are in the element type they were uploaded in (FP64 for
        # matrices assembled on the GPU)
        dt = np.float32 if getattr(self, "leaf_dtype", F64) == F32 else np.float64
        data = np.empty(int(nv.max()) if len(nv) else 0, dt)
        check(load().ahm_hmat_download(self.h, None, data.ctypes.data_as(_p)), "download")
        return leaves, data

    def matvec(self, x, y, alpha: float = 1.0):
        """y += alpha * H x for torch tensors (1-D, or 2-D column-major as (p, n))."""
        torch = _torch()
        vec = F64 if x.dtype == torch.float64 else F32
        if x.dim() == 1:
            p, ldx, ldy = 1, x.shape[0], y.shape[0]
        else:  # (p, n) row-major storage == n x p column-major
            p, ldx, ldy = x.shape[0], x.stride(0), y.stride(0)
        self.ctx.bind_stream()
        check(load().ahm_hmat_matvec(self.h, alpha, _p(x.data_ptr()), ldx, _p(y.data_ptr()), ldy, p,
                                     vec), "ahm_hmat_matvec")
        return y

    def matvec_noalloc(self, x, y, alpha: float = 1.0):
        """matvec through ahm_hmat_matvec_noalloc: the workspace is the
        library's static per-device array, nothing is allocated."""
        torch = _torch()
        vec = F64 if x.dtype == torch.float64 else F32
        if x.dim() == 1:
            p, ldx, ldy = 1, x.shape[0], y.shape[0]
        else:
            p, ldx, ldy = x.shape[0], x.stride(0), y.stride(0)
        self.ctx.bind_stream()
        check(load().ahm_hmat_matvec_noalloc(self.h, alpha, _p(x.data_ptr()), ldx, _p(y.data_ptr()), ldy, p,
                                             vec), "ahm_hmat_matvec_noalloc")
        return y

    def matvec_t(self, x, y, alpha: float = 1.0):
        """y += alpha * H^T x (mltaGeHhVec, real case): x has n_rows values per
        right-hand side, y n_cols; 1-D tensors or (p, n) as in matvec."""
        torch = _torch()
        vec = F64 if x.dtype == torch.float64 else F32
        if x.dim() == 1:
            p, ldx, ldy = 1, x.shape[0], y.shape[0]
        else:
            p, ldx, ldy = x.shape[0], x.stride(0), y.stride(0)
        self.ctx.bind_stream()
        check(load().ahm_hmat_matvec_t(self.h, alpha, _p(x.data_ptr()), ldx, _p(y.data_ptr()), ldy, p,
                                       vec), "ahm_hmat_matvec_t")
        return y

    def profile(self, x, y, reps: int = 10, alpha: float = 1.0):
        """Mean (ms_phase_a, ms_phase_b) of `reps` products timed with HIP events."""
        torch = _torch()
        vec = F64 if x.dtype == torch.float64 else F32
        if x.dim() == 1:
            p, ldx, ldy = 1, x.shape[0], y.shape[0]
        else:
            p, ldx, ldy = x.shape[0], x.stride(0), y.stride(0)
        a, b = _d(), _d()
        self.ctx.bind_stream()
        check(load().ahm_hmat_matvec_profile(self.h, alpha, _p(x.data_ptr()), ldx, _p(y.data_ptr()), ldy,
                                             p, vec, reps, C.byref(a), C.byref(b)), "matvec_profile")
        return a.value, b.value

    def gmres(self, b, x, eps: float = 1e-8, m: int = 30, nsteps: int = 10000, nhist: int = 0,
              prec=None):
        """Device GMRES(m) (solvers/GMRES.cpp semantics); b, x FP64 torch
        tensors, x updated in place; prec = (L, U) TriFactors (either may be
        None) makes HLU_solve the right preconditioner.  Returns
        dict(converged, eps, nsteps, nmatvec, hist)."""
        e, ns, mv = _d(eps), _u32(nsteps), _u32(0)
        hist = np.zeros(max(nhist, 1))
        self.ctx.bind_stream()
        if prec is None:
            rc = load().ahm_gmres(self.h, _p(b.data_ptr()), _p(x.data_ptr()), C.byref(e), m, C.byref(ns),
                                  hist.ctypes.data_as(_p), nhist, C.byref(mv))
        else:
            L, U = prec
            rc = load().ahm_gmres_prec(self.h, L.h if L is not None else None, U.h if U is not None else None,
                                       _p(b.data_ptr()), _p(x.data_ptr()), C.byref(e), m, C.byref(ns),
                                       hist.ctypes.data_as(_p), nhist, C.byref(mv))
        if rc not in (OK, 5):
            check(rc, "ahm_gmres")
        return dict(converged=rc == OK, eps=e.value, nsteps=ns.value, nmatvec=mv.value,
                    hist=hist[:nhist].copy())

    def set_graphs(self, on: bool = True):
        """GMRES steps of this handle replayed from captured hipGraphs, or
        launched one by one (the default)."""
        check(load().ahm_hmat_set_graphs(self.h, int(on)), "ahm_hmat_set_graphs")

    def set_arnoldi(self, mode: int = 0):
        """Form of the GMRES orthogonalisation and rotations: 1 chunked
        launches, 2 one workgroup per system (systems up to 16384 rows),
        0 automatic.  Both forms give bit-identical iterates."""
        check(load().ahm_hmat_set_arnoldi(self.h, int(mode)), "ahm_hmat_set_arnoldi")

    def cg(self, b, x, eps: float = 1e-8, nsteps: int = 10000, prec=None):
        e, ns = _d(eps), _u32(nsteps)
        self.ctx.bind_stream()
        if prec is None:
            rc = load().ahm_cg(self.h, _p(b.data_ptr()), _p(x.data_ptr()), C.byref(e), C.byref(ns))
        else:
            L, U = prec
            rc = load().ahm_cg_prec(self.h, L.h if L is not None else None, U.h if U is not None else None,
                                    _p(b.data_ptr()), _p(x.data_ptr()), C.byref(e), C.byref(ns))
        if rc not in (OK, 5):
            check(rc, "ahm_cg")
        return dict(converged=rc == OK, eps=e.value, nsteps=ns.value)

    def gmres_ir(self, lo: "HMatrix", b, x, eps: float = 1e-8, m: int = 50, inner_eps: float = 1e-5,
                 inner_steps: int = 1000, max_outer: int = 20, nhist: int = 32):
        """FP64 iterative refinement with inner GMRES on `lo` (e.g. FP32 copy)."""
        e, outer, inner = _d(eps), _u32(0), _u32(0)
        hist = np.zeros(max(nhist, 1))
        self.ctx.bind_stream()
        rc = load().ahm_gmres_ir(self.h, lo.h, _p(b.data_ptr()), _p(x.data_ptr()), C.byref(e), m, inner_eps,
                                 inner_steps, max_outer, C.byref(outer), C.byref(inner),
                                 hist.ctypes.data_as(_p), nhist)
        if rc not in (OK, 5):
            check(rc, "ahm_gmres_ir")
        return dict(converged=rc == OK, eps=e.value, outer=outer.value, inner=inner.value,
                    hist=hist[:min(nhist, outer.value + 1)].copy())

    def autotune(self, p: int = 1, reps: int = 5):
        """Pick the fastest tile size (row-panel / V^T chunk values) and, for
        p <= 8, launch form for p right-hand sides: (chunk values, ms, fused)."""
        cv, ms, fu = C.c_uint64(), _d(), _i()
        vec = F64 if self.storage == F64 else F32
        self.ctx.bind_stream()
        check(load().ahm_hmat_autotune(self.h, p, vec, reps, C.byref(cv), C.byref(ms), C.byref(fu)),
              "autotune")
        return cv.value, ms.value, int(fu.value)

    def set_fused(self, on: int = 1):
        """Launch form of products of 1-8 right-hand sides: 0 three kernels,
        1 (True) one launch whose row panels form the V^T x of their own
        low-rank leaves, 2 one persistent cooperative launch of the
        three-kernel tasks with grid barriers between the phases."""
        check(load().ahm_hmat_set_fused(self.h, int(on)), "ahm_hmat_set_fused")

    def fused_status(self) -> bool:
        """Always False: the single launch has no waits to give up (kept for
        the C-ABI's status query)."""
        a = _i()
        check(load().ahm_hmat_fused_status(self.h, C.byref(a)), "ahm_hmat_fused_status")
        return bool(a.value)

    def matvec_host(self, x: np.ndarray, y: np.ndarray, alpha: float = 1.0):
        x = np.asfortranarray(x, np.float64)
        y2 = np.asfortranarray(y, np.float64).copy()
        p = 1 if x.ndim == 1 else x.shape[1]
        self.ctx.bind_stream()
        check(load().ahm_hmat_matvec_host(self.h, alpha, x.ctypes.data_as(_p), self.n_cols,
                                          y2.ctypes.data_as(_p), self.n_rows, p), "matvec_host")
        return y2

    def close(self):
        # as TriFactor.close: never destroy through a freed context
        if self.h:
            if self.ctx.h:
                load().ahm_hmat_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
