"""Multi-RHS products with V^T X on BF16 MFMA (AHM_BF16 storage: FP32
values, a BF16 copy of V; X rounded to BF16 for V^T X, FP32 accumulation).

Checks, for 17 to 64 right-hand sides:
* against an FP64 restatement of the same arithmetic (V and X rounded to
  BF16 to nearest-even, everything else FP32-rounded): within 1e-5 relative
  (max-norm).  What remains is FP32 accumulation order;
* against the FP64 product, element by element: |y - y64| <=
  2^-8 (1 + 1e-3) (|U| (|V|^T |X|)) + 5e-6 max|y64|.  The first term is the
  BF16 rounding of V and X (each <= 2^-9 relative); the second is the FP32
  storage and accumulation bar of the FP32 products.  The normwise error
  against FP64 is printed;
* fewer than 9 right-hand sides multiply exactly as the FP32 matrix (bit for
  bit).
The rounding used by the restatement is pinned against torch's bfloat16
conversion (CPU test)."""
import numpy as np
import pytest

from helpers import bf16_rne, bf16_vtx_emulation, block_matrix, leaves_of, relerr, sphere_hmatrix

torch = pytest.importorskip("torch")


def test_bf16_restatement_rounding_matches_torch():
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.standard_normal(20000).astype(np.float32) * 10.0 ** rng.integers(-8, 8, 20000),
                        # exact ties: 1 + 2^-8 (to even: 1), 1 + 3 * 2^-8 (to even: 1 + 2^-6)
                        np.array([1 + 2 ** -8, 1 + 3 * 2 ** -8, -(1 + 2 ** -8), 3.0, 0.0], np.float32)])
    ref = torch.tensor(x).to(torch.bfloat16).float().numpy()
    assert np.array_equal(bf16_rne(x), ref)
    assert bf16_rne(np.float32(1 + 2 ** -8)) == 1.0
    assert bf16_rne(np.float32(1 + 3 * 2 ** -8)) == np.float32(1 + 2 ** -6)


@pytest.fixture(scope="module")
def ctx():
    from ahmed_amd import _lib
    assert torch.cuda.is_available()
    return _lib.Context(0)


def check_bf16(D, N, L, X, y64, what):
    from ahmed_amd import _lib
    Db = D.convert(_lib.BF16)
    p = X.shape[1]
    Y = torch.zeros((p, N), dtype=torch.float32, device="cuda")
    Db.matvec(torch.as_tensor(X.T.astype(np.float32).copy(), device="cuda"), Y)
    yb = Y.cpu().numpy().T.astype(np.float64)
    ye, lr = bf16_vtx_emulation(N, L, X)
    e_em = relerr(yb, ye)
    e64 = relerr(yb, y64)
    bound = 2.0 ** -8 * (1 + 1e-3) * lr + 5e-6 * np.max(np.abs(y64))
    ratio = float(np.max(np.abs(yb - y64) / bound))
    print(f"[gpu] BF16 V^T X {what} p={p}: vs FP64 {e64:.2e}, vs restatement {e_em:.2e}, "
          f"worst error / bound {ratio:.3f}", flush=True)
    assert e_em <= 1e-5
    assert ratio <= 1.0
    return Db


@pytest.mark.gpu
@pytest.mark.parametrize("p", [17, 32, 33, 64])
def test_bf16_vtx_sphere(ctx, p):
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(16384, 64, 1e-8)
    leaves, data = leaves_of(H)
    D = _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)
    X = np.random.default_rng(p).standard_normal((H.N, p))
    check_bf16(D, H.N, H.layout(), X, H.matmat(X), "sphere N=16384")


@pytest.mark.gpu
@pytest.mark.parametrize("p", [17, 32, 48])
def test_bf16_vtx_every_panel_height(ctx, p):
    from ahmed_amd import _lib
    from test_gpu_matvec import MFMA_SIZES
    N, leaves, data, A = block_matrix(MFMA_SIZES, 11)
    D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    X = np.random.default_rng(p).standard_normal((N, p))
    L = {f: leaves[f] for f in ("b1", "n1", "b2", "n2", "rank", "offset")}
    L["data"] = data
    check_bf16(D, N, L, X, A @ X, f"panel heights N={N}")


@pytest.mark.gpu
@pytest.mark.parametrize("p", [1, 8])
def test_bf16_storage_below_nine_rhs_is_the_fp32_product(ctx, p):
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(16384, 64, 1e-8)
    leaves, data = leaves_of(H)
    D = _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)
    D32, Db = D.convert(_lib.F32), D.convert(_lib.BF16)
    shape = (H.N,) if p == 1 else (p, H.N)
    X = torch.randn(shape, dtype=torch.float32, device="cuda")
    Ya, Yb = torch.zeros_like(X), torch.zeros_like(X)
    D32.matvec(X, Ya)
    Db.matvec(X, Yb)
    torch.cuda.synchronize()
    assert torch.equal(Ya, Yb)
