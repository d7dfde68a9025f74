"""GPU transposed H-matvec y += d A^T x (the real case of mltaGeHhVec /
mltaGeHhGeM, H/mltaGeHVec.cpp:61-104) against the compiled reference's outputs
(tests/golden/golden_transposed.npz) and the oracle, on the identical
H-matrix.  Several right-hand sides take the MFMA column dots (32 at a time)
on panels of every height.  Bar: FP64 within 1e-12 relative (max-norm);
FP32 storage within 5e-6 of the FP64 product; repeated products
bit-identical."""
import ctypes as C
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_names
from helpers import golden_hmatrix, leaves_of, relerr, sphere_hmatrix

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from ahmed_amd import _lib
    assert torch.cuda.is_available()
    return _lib.Context(0)


def upload(ctx, H, storage=0):
    from ahmed_amd import _lib
    leaves, data = leaves_of(H)
    return _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data, storage=storage)


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


@pytest.mark.parametrize("name", golden_names())
def test_transposed_matches_reference_golden(ctx, name):
    _, _, H = golden_hmatrix(name)
    t = np.load(os.path.join(GOLDEN, "golden_transposed.npz"))
    print(f"[gpu] transposed golden {name} N={H.N}", flush=True)
    D = upload(ctx, H)
    y = dev(t[f"{name}_y0"])
    D.matvec_t(dev(t[f"{name}_x"]), y, float(t[f"{name}_d"]))
    err = relerr(y.cpu().numpy(), t[f"{name}_y"])
    print(f"[gpu] {name}: A^T x vs reference {err:.2e}")
    assert err <= 1e-12
    # deterministic: the same product twice gives the same bits
    y2 = dev(t[f"{name}_y0"])
    D.matvec_t(dev(t[f"{name}_x"]), y2, float(t[f"{name}_d"]))
    assert torch.equal(y, y2)
    D.close()


def test_transposed_multi_rhs_fp32_and_host(ctx):
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(16384, 64, 1e-8)
    print("[gpu] transposed p=3, FP32, host entry N=16384", flush=True)
    D = upload(ctx, H)
    rng = np.random.default_rng(9)
    X = rng.standard_normal((3, H.N))
    ref = np.stack([H.matvec_t(X[j], d=1.25) for j in range(3)])
    Y = dev(np.zeros((3, H.N)))
    D.matvec_t(dev(X), Y, 1.25)
    assert relerr(Y.cpu().numpy(), ref) <= 1e-12
    # host pointers (the surface's mltaGeHhGeM path)
    Yh = np.zeros((3, H.N))
    Xh = np.ascontiguousarray(X)
    _lib.check(_lib.load().ahm_hmat_matvec_t_host(D.h, 1.25, Xh.ctypes.data_as(C.c_void_p), H.N,
                                                  Yh.ctypes.data_as(C.c_void_p), H.N, 3), "host")
    assert relerr(Yh, ref) <= 1e-12
    # FP32 storage
    D32 = D.convert()
    Y32 = dev(np.zeros((3, H.N)), torch.float32)
    D32.matvec_t(dev(X, torch.float32), Y32, 1.25)
    err = relerr(Y32.cpu().numpy().astype(np.float64), ref)
    print(f"[gpu] transposed FP32 storage: rel.err vs FP64 {err:.2e}")
    assert err <= 5e-6
    D32.close()
    D.close()


def test_transposed_is_the_adjoint(ctx):
    """<A^T x, z> = <x, A z> for the GPU's own two products."""
    _, _, H = sphere_hmatrix(16384, 64, 1e-8)
    D = upload(ctx, H)
    rng = np.random.default_rng(4)
    x, z = rng.standard_normal(H.N), rng.standard_normal(H.N)
    at = dev(np.zeros(H.N))
    az = dev(np.zeros(H.N))
    D.matvec_t(dev(x), at)
    D.matvec(dev(z), az)
    lhs = float(torch.dot(at, dev(z)))
    rhs = float(torch.dot(dev(x), az))
    assert abs(lhs - rhs) <= 1e-12 * abs(rhs) * 10
    D.close()


@pytest.mark.parametrize("p", [2, 17, 32, 40])
def test_transposed_multi_rhs_mfma_every_panel_height(ctx, p):
    """Several right-hand sides run the column dots 32 at a time on MFMA:
    row panels of every height class, V^T panels of more than 128 rows, a
    second group of right-hand sides (p = 40), FP64 and FP32 storage, against
    the dense transpose."""
    from ahmed_amd import _lib
    from test_gpu_matvec import MFMA_SIZES, VT_HEIGHTS
    from helpers import block_matrix
    nseg = len(VT_HEIGHTS)

    def rank_of(i, j):
        q, r = divmod(VT_HEIGHTS[j], nseg - 1)
        return q + (1 if (i - (i > j)) < r else 0)

    for sizes, seed, rk in ((MFMA_SIZES, 11, None), ([64] * nseg, 5, rank_of)):
        N, leaves, data, A = block_matrix(sizes, seed, rk)
        print(f"[gpu] transposed MFMA N={N} p={p}", flush=True)
        D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
        rng = np.random.default_rng(p)
        X = rng.standard_normal((p, N))
        Y0 = rng.standard_normal((p, N))
        Y = dev(Y0)
        D.matvec_t(dev(X), Y, -0.75)
        ref = Y0 - 0.75 * (A.T @ X.T).T
        assert relerr(Y.cpu().numpy(), ref) <= 1e-12
        # the same columns one at a time (the single-RHS kernels)
        for j in (0, p - 1):
            y = dev(Y0[j])
            D.matvec_t(dev(X[j]), y, -0.75)
            assert relerr(y.cpu().numpy(), ref[j]) <= 1e-12
        D32 = D.convert()
        Y32 = dev(np.zeros((p, N)), torch.float32)
        D32.matvec_t(dev(X, torch.float32), Y32, 1.0)
        assert relerr(Y32.cpu().numpy().astype(np.float64), (A.T @ X.T).T) <= 5e-6
