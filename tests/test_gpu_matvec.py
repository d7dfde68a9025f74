"""GPU H-matvec parity: the HIP library against the reference's own outputs
(golden fixtures) and against the oracle, on the identical H-matrix.

Bar: FP64 products within 1e-12 relative (max-norm) of the reference/oracle;
FP32 storage within 2e-6 relative of the FP64 product (the stated
precision of FP32 leaves with FP32 accumulation).
"""
import numpy as np
import pytest

from conftest import golden_names
from helpers import block_matrix, golden_hmatrix, leaves_of, relerr, sphere_hmatrix

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from ahmed_amd import _lib
    assert torch.cuda.is_available()
    return _lib.Context(0)


def upload(ctx, H, storage=0, keep=False):
    from ahmed_amd import _lib
    leaves, data = leaves_of(H)
    return _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data, storage=storage, keep_leaves=keep)


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


@pytest.mark.parametrize("name", golden_names())
def test_matvec_matches_reference_golden(ctx, name):
    g, T, H = golden_hmatrix(name)
    print(f"[gpu] golden {name} N={H.N}", flush=True)
    D = upload(ctx, H)
    y = dev(np.zeros(H.N))
    D.matvec(dev(g["x"]), y, 1.0)
    assert relerr(y.cpu().numpy(), g["y"]) <= 1e-12
    # p = 2, alpha = -0.5, accumulating into zero (mltaGeHGeM)
    X = dev(g["X"].T.copy())            # (p, n) row-major == n x p column-major
    Y = dev(np.zeros((2, H.N)))
    D.matvec(X, Y, -0.5)
    assert relerr(Y.cpu().numpy().T, g["Y"]) <= 1e-12


def test_matvec_accumulates_and_scales(ctx):
    _, T, H = sphere_hmatrix(6000, 32, 1e-8)
    D = upload(ctx, H)
    rng = np.random.default_rng(0)
    x, y0 = rng.standard_normal(H.N), rng.standard_normal(H.N)
    ref = H.matvec(x, d=-2.5, y=y0)
    y = dev(y0)
    D.matvec(dev(x), y, -2.5)
    assert relerr(y.cpu().numpy(), ref) <= 1e-12


def test_matvec_cfg2_shape_single_and_multi_rhs(ctx):
    # sphere, leaf size 64, eta 0.8, ACA eps 1e-8 (the one-GPU FP64 config, at N=32768)
    _, T, H = sphere_hmatrix(32768, 64, 1e-8, weighted=False)
    print(f"[gpu] cfg2-shape N={H.N}", flush=True)
    D = upload(ctx, H)
    st = D.stats()
    assert st["nvals"] == len(H.layout()["data"])
    rng = np.random.default_rng(1)
    x = rng.standard_normal(H.N)
    y = dev(np.zeros(H.N))
    D.matvec(dev(x), y)
    assert relerr(y.cpu().numpy(), H.matvec(x)) <= 1e-12
    X = rng.standard_normal((H.N, 32))
    Y = dev(np.zeros((32, H.N)))
    D.matvec(dev(X.T.copy()), Y)
    assert relerr(Y.cpu().numpy().T, H.matmat(X)) <= 1e-12


def test_matvec_fp32_storage_error_is_stated(ctx):
    _, T, H = sphere_hmatrix(16384, 64, 1e-8)
    D = upload(ctx, H)
    D32 = D.convert()
    rng = np.random.default_rng(2)
    x = rng.standard_normal(H.N)
    y64 = H.matvec(x)
    y = dev(np.zeros(H.N), torch.float32)
    D32.matvec(dev(x, torch.float32), y)
    err = relerr(y.cpu().numpy(), y64)
    print(f"[gpu] FP32 leaves + FP32 accumulation vs FP64 H-matvec: {err:.2e}")
    assert err <= 2e-6


def test_matvec_host_drop_in(ctx):
    g, T, H = golden_hmatrix("pca_yukawa1500")
    D = upload(ctx, H)
    y = D.matvec_host(g["x"], np.zeros(H.N))
    assert relerr(y, g["y"]) <= 1e-12


def test_keep_leaves_download_roundtrip(ctx):
    g, T, H = golden_hmatrix("cube_expr2048")
    D = upload(ctx, H, keep=True)
    leaves, data = D.download()
    L = H.layout()
    assert np.array_equal(leaves["rank"], L["rank"])
    assert np.array_equal(data, L["data"])


def test_rank_zero_and_all_dense_edge_cases(ctx):
    from ahmed_amd import _lib
    rng = np.random.default_rng(3)
    N = 300
    # two dense diagonal leaves, one rank-0 leaf, one rank-2 leaf
    b1 = [0, 150, 0, 150]
    n1 = [150, 150, 150, 150]
    b2 = [0, 150, 150, 0]
    n2 = [150, 150, 150, 150]
    rank = [-1, -1, 0, 2]
    sizes = [150 * 150, 150 * 150, 0, 2 * 300]
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    data = rng.standard_normal(int(sum(sizes)))
    leaves = _lib.leaf_array(b1, n1, b2, n2, rank, off)
    D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    A = np.zeros((N, N))
    A[:150, :150] = data[:22500].reshape(150, 150).T
    A[150:, 150:] = data[22500:45000].reshape(150, 150).T
    U = data[45000:45300].reshape(2, 150).T
    V = data[45300:45600].reshape(2, 150).T
    A[150:, :150] = U @ V.T
    x = rng.standard_normal(N)
    y = dev(np.zeros(N))
    D.matvec(dev(x), y)
    assert relerr(y.cpu().numpy(), A @ x) <= 1e-13
    with pytest.raises(_lib.AhmError):
        bad = leaves.copy()
        bad["n1"][0] = 400
        _lib.HMatrix.upload(ctx, N, N, bad, data)


# 3, 8: 8-wide streaming kernel; 12, 16: MFMA panel products with one
# 16-column tile; 17, 40: MFMA with two tiles / a second group
@pytest.mark.parametrize("p", [3, 8, 12, 16, 17, 40])
def test_multi_rhs_odd_counts_fp64_and_fp32(ctx, p):
    _, T, H = sphere_hmatrix(16384, 64, 1e-8)
    D = upload(ctx, H)
    rng = np.random.default_rng(p)
    X = rng.standard_normal((H.N, p))
    Y0 = rng.standard_normal((H.N, p))
    ref = H.matmat(X, d=0.75, Y=Y0)
    Y = dev(Y0.T.copy())
    D.matvec(dev(X.T.copy()), Y, 0.75)
    assert relerr(Y.cpu().numpy().T, ref) <= 1e-12
    D32 = D.convert()
    Y32 = dev(np.zeros((p, H.N)), torch.float32)
    D32.matvec(dev(X.T.copy(), torch.float32), Y32, 1.0)
    err = relerr(Y32.cpu().numpy().T.astype(np.float64), H.matmat(X))
    print(f"[gpu] FP32 p={p}: rel.err vs FP64 {err:.2e}")
    # FP32 values, right-hand sides and accumulation: 1e-6..2e-6 (max-norm)
    # measured; the sum order follows the task shapes, hence the margin
    assert err <= 5e-6


def test_autotuned_chunking_keeps_results(ctx):
    _, T, H = sphere_hmatrix(32768, 64, 1e-8, weighted=False)
    D = upload(ctx, H)
    rng = np.random.default_rng(9)
    x = dev(rng.standard_normal(H.N))
    y0 = dev(np.zeros(H.N))
    D.matvec(x, y0)
    X = dev(rng.standard_normal((8, H.N)))
    Y0 = dev(np.zeros((8, H.N)))
    D.matvec(X, Y0)
    for p in (1, 8):
        cv, ms, fused = D.autotune(p=p, reps=3)
        print(f"[gpu] autotune p={p}: chunk {cv} values, fused {fused}, {ms:.3f} ms")
        assert not D.fused_status()
        y = dev(np.zeros(H.N))
        D.matvec(x, y)
        Y = dev(np.zeros((8, H.N)))
        D.matvec(X, Y)
        # chunking changes the summation order across chunks only
        assert relerr(y.cpu().numpy(), y0.cpu().numpy()) <= 1e-14
        assert relerr(Y.cpu().numpy(), Y0.cpu().numpy()) <= 1e-14
    assert relerr(y0.cpu().numpy(), H.matvec(x.cpu().numpy())) <= 1e-12


# Row panels of every height class of the MFMA product's wavefront shapes:
# 1-4 row tiles (one whole tile or less per wavefront), 5-8 (16-column
# rounds, split tiles), 9-16 (several whole tiles per wavefront), and a
# segment of 300 rows cut into 256 + 44.  The V^T panels take the summed
# ranks of a column segment (up to about 90 rows).  p = 17 and 32: one group
# of 32 right-hand sides (two column tiles); 33 and 48: a second group with
# one column tile; 64: two full groups.
MFMA_SIZES = [1, 5, 16, 17, 31, 33, 48, 63, 65, 80, 96, 100, 112, 127, 128, 129, 144, 176, 191,
              200, 224, 240, 255, 256, 300]


@pytest.mark.parametrize("p", [17, 32, 33, 48, 64])
def test_mfma_product_every_panel_height(ctx, p):
    from ahmed_amd import _lib
    N, leaves, data, A = block_matrix(MFMA_SIZES, 11)
    print(f"[gpu] MFMA panel heights N={N} p={p}", flush=True)
    D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    rng = np.random.default_rng(p)
    X = rng.standard_normal((N, p))
    Y0 = rng.standard_normal((N, p))
    Y = dev(Y0.T.copy())
    D.matvec(dev(X.T.copy()), Y, -1.25)
    ref = Y0 - 1.25 * (A @ X)
    assert relerr(Y.cpu().numpy().T, ref) <= 1e-12
    if p == 32:
        D32 = D.convert()
        Y32 = dev(np.zeros((p, N)), torch.float32)
        D32.matvec(dev(X.T.copy(), torch.float32), Y32, 1.0)
        assert relerr(Y32.cpu().numpy().T.astype(np.float64), A @ X) <= 5e-6


# V^T panels (the summed ranks of one column segment's low-rank leaves) of
# every height from one past 8 row tiles to 256 rows, and one of 300 rows cut
# into 256 + 44: the MFMA launch runs V^T panels of more than 8 row tiles as
# two workgroups of half the row tiles each.
VT_HEIGHTS = [129, 136, 144, 152, 160, 175, 192, 200, 224, 241, 256, 300]


@pytest.mark.parametrize("p", [17, 32, 40])
def test_mfma_product_tall_vt_panels(ctx, p):
    from ahmed_amd import _lib
    nseg = len(VT_HEIGHTS)

    def rank_of(i, j):   # leaf (i, j), i != j: the column's 11 ranks sum to VT_HEIGHTS[j]
        q, r = divmod(VT_HEIGHTS[j], nseg - 1)
        return q + (1 if (i - (i > j)) < r else 0)

    N, leaves, data, A = block_matrix([64] * nseg, 5, rank_of)
    print(f"[gpu] MFMA tall V^T panels N={N} p={p}", flush=True)
    D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    rng = np.random.default_rng(p + 100)
    X = rng.standard_normal((N, p))
    Y = dev(np.zeros((p, N)))
    D.matvec(dev(X.T.copy()), Y, 0.75)
    assert relerr(Y.cpu().numpy().T, 0.75 * (A @ X)) <= 1e-12
    D32 = D.convert()
    Y32 = dev(np.zeros((p, N)), torch.float32)
    D32.matvec(dev(X.T.copy(), torch.float32), Y32, 1.0)
    assert relerr(Y32.cpu().numpy().T.astype(np.float64), A @ X) <= 5e-6


def cache_test_matrix(N, rng, poison):
    """Leaves [0, h) x [0, N) and [h, N) x [0, h) dense, [h, N) x [h, N) of rank
    6 (so products use the V^T x workspace as well as the arena); poison:
    values NaN / +Inf / -Inf instead of random numbers."""
    from ahmed_amd import _lib
    h, k = N // 2, 6
    sizes = [h * N, (N - h) * h, k * 2 * (N - h)]
    off = np.concatenate([[0], np.cumsum(sizes)[:-1]])
    leaves = _lib.leaf_array([0, h, h], [h, N - h, N - h], [0, 0, h], [N, h, N - h], [-1, -1, k], off)
    if poison:
        data = np.tile(np.array([np.nan, np.inf, -np.inf]), sum(sizes) // 3 + 1)[:sum(sizes)]
        return leaves, data, None
    data = rng.standard_normal(sum(sizes))
    A = np.empty((N, N))
    A[:h] = data[:h * N].reshape(N, h).T
    A[h:, :h] = data[off[1]:off[2]].reshape(h, N - h).T
    U = data[off[2]:off[2] + k * (N - h)].reshape(k, N - h).T
    V = data[off[2] + k * (N - h):].reshape(k, N - h).T
    A[h:, h:] = U @ V.T
    return leaves, data, A


def test_large_blocks_reused_from_the_allocation_cache(ctx):
    """Buffers of >= 1 GiB go to the allocation cache when released and are
    handed to the next large request (best fit).  A matrix of NaN / Inf
    values is multiplied (1 and 32 right-hand sides, so every workspace holds
    them too) and released: the cache grows by its arena.  The next, smaller
    matrix takes that block (the cache shrinks) and its products are finite
    and correct, so no path relies on fresh device memory being zero.
    Releasing the cache empties it and leaves the context usable."""
    from ahmed_amd import _lib
    rng = np.random.default_rng(7)
    ctx.release_cached()
    assert ctx.cached_bytes() == 0
    print("[gpu] allocation cache: poisoned matrix N=16000", flush=True)
    leaves, data, _ = cache_test_matrix(16000, rng, poison=True)
    H = _lib.HMatrix.upload(ctx, 16000, 16000, leaves, data)
    for p in (1, 32):
        X = torch.ones((p, 16000), dtype=torch.float64, device="cuda").squeeze(0)
        Y = torch.zeros_like(X)
        H.matvec(X, Y)
    torch.cuda.synchronize()
    before = ctx.cached_bytes()
    del H
    grown = ctx.cached_bytes()
    print(f"[gpu] cached bytes {before} -> {grown} after release", flush=True)
    assert grown >= before + 8 * len(data)
    del data
    N = 15500
    leaves, data, A = cache_test_matrix(N, rng, poison=False)
    H = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    after = ctx.cached_bytes()
    print(f"[gpu] cached bytes {grown} -> {after} after the second upload", flush=True)
    assert after < grown
    x = rng.standard_normal(N)
    y = torch.zeros(N, dtype=torch.float64, device="cuda")
    H.matvec(dev(x), y)
    yh = y.cpu().numpy()
    assert np.all(np.isfinite(yh))
    assert relerr(yh, A @ x) <= 1e-13
    X = rng.standard_normal((N, 32))
    Y = torch.zeros((32, N), dtype=torch.float64, device="cuda")
    H.matvec(dev(X.T.copy()), Y)
    Yh = Y.cpu().numpy().T
    assert np.all(np.isfinite(Yh))
    assert relerr(Yh, A @ X) <= 1e-12
    del H
    ctx.release_cached()
    assert ctx.cached_bytes() == 0
    H = _lib.HMatrix.upload(ctx, 64, 64, _lib.leaf_array([0], [64], [0], [64], [-1], [0]),
                            np.arange(64 * 64, dtype=np.float64))
    y = torch.zeros(64, dtype=torch.float64, device="cuda")
    H.matvec(dev(np.ones(64)), y)
    ref = np.arange(64 * 64, dtype=np.float64).reshape(64, 64).T @ np.ones(64)
    assert relerr(y.cpu().numpy(), ref) <= 1e-15
