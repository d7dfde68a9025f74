"""Products of 1-8 right-hand sides in one launch (ahm_hmat_set_fused): one
workgroup per row panel that forms the V^T x values of its own low-rank
leaves, so that no workgroup waits for another and no chunk sums are needed.

Checks: against the three-launch form and the oracle / dense reference for
FP64 (<= 1e-12 relative; the V^T x sums run in another order) and FP32, 1, 3
and 8 right-hand sides, on a sphere matrix, on the matrix with row panels of
every height and on a batched matrix; repeated launches bit-identical; replays
of a captured graph, each checked, with eager work between them; the
autotuner's choice between the two forms leaves the product correct."""
import numpy as np
import pytest

from helpers import block_matrix, leaves_of, relerr, sphere_hmatrix

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from ahmed_amd import _lib
    assert torch.cuda.is_available()
    return _lib.Context(0)


def both_forms(D, X, alpha=0.5):
    Y0 = torch.zeros_like(X)
    Y1 = torch.zeros_like(X)
    D.set_fused(False)
    D.matvec(X, Y0, alpha)
    D.set_fused(True)
    D.matvec(X, Y1, alpha)
    torch.cuda.synchronize()
    return Y0, Y1


@pytest.mark.parametrize("p", [1, 3, 8])
def test_single_launch_sphere(ctx, p):
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(32768, 64, 1e-8, weighted=False)
    print(f"[gpu] single-launch product sphere N=32768 p={p}", flush=True)
    leaves, data = leaves_of(H)
    D = _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)
    shape = (H.N,) if p == 1 else (p, H.N)
    X = torch.randn(shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(p))
    Y0, Y1 = both_forms(D, X)
    assert relerr(Y1.cpu().numpy(), Y0.cpu().numpy()) <= 1e-13
    ref = H.matvec(X.cpu().numpy(), d=0.5) if p == 1 else H.matmat(X.cpu().numpy().T, d=0.5).T
    assert relerr(Y1.cpu().numpy(), ref) <= 1e-12
    D32 = D.convert()
    Y0, Y1 = both_forms(D32, X.float())
    assert relerr(Y1.cpu().numpy(), Y0.cpu().numpy()) <= 1e-5
    assert relerr(Y1.double().cpu().numpy(), ref) <= 2e-5


@pytest.mark.parametrize("p", [1, 8])
def test_single_launch_every_panel_height_and_repeats(ctx, p):
    from ahmed_amd import _lib
    from test_gpu_matvec import MFMA_SIZES
    N, leaves, data, A = block_matrix(MFMA_SIZES, 11)
    print(f"[gpu] single-launch product panel heights N={N} p={p}", flush=True)
    D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    shape = (N,) if p == 1 else (p, N)
    X = torch.randn(shape, dtype=torch.float64, device="cuda")
    Y0, Y1 = both_forms(D, X, -1.0)
    ref = -(A @ X.cpu().numpy().T).T if p > 1 else -(A @ X.cpu().numpy())
    assert relerr(Y1.cpu().numpy(), ref) <= 1e-12
    assert relerr(Y0.cpu().numpy(), ref) <= 1e-12
    for _ in range(10):
        Y2 = torch.zeros_like(X)
        D.matvec(X, Y2, -1.0)
        assert torch.equal(Y2, Y1)


def test_single_launch_batched_and_autotune(ctx):
    from ahmed_amd import _lib, geometry
    trees, kernels = [], []
    rng = np.random.default_rng(5)
    for s in range(5):
        n = int(rng.integers(1500, 5000))
        pts = geometry.ellipsoid(n, (1.0, float(rng.uniform(1, 3)), float(rng.uniform(1, 3))), seed=40 + s)
        trees.append(_lib.Tree(pts, 32, 0.64))
        kernels.append((_lib.KERNEL_YUKAWA, float(rng.uniform(0, 4)), 1.0 / n))
    HB, _ = _lib.HMatrix.assemble_batch(ctx, trees, kernels, eps=1e-6)
    N = HB.n_rows
    print(f"[gpu] single-launch product batched N={N}", flush=True)
    X = torch.randn((8, N), dtype=torch.float64, device="cuda")
    Y0, Y1 = both_forms(HB, X, 1.0)
    assert relerr(Y1.cpu().numpy(), Y0.cpu().numpy()) <= 1e-12
    chunk, ms, fused = HB.autotune(p=8)
    print(f"[gpu] autotune: chunk {chunk}, {ms:.3f} ms, single launch {fused}", flush=True)
    Y2 = torch.zeros_like(X)
    HB.matvec(X, Y2, 1.0)
    assert relerr(Y2.cpu().numpy(), Y0.cpu().numpy()) <= 1e-12


def test_single_launch_replays_from_a_captured_graph(ctx):
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(16384, 64, 1e-8)
    print("[gpu] single-launch product in a captured graph, N=16384", flush=True)
    leaves, data = leaves_of(H)
    D = _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)
    D.set_fused(True)
    X = torch.randn((8, H.N), dtype=torch.float64, device="cuda")
    Yref = torch.zeros_like(X)
    D.matvec(X, Yref)
    Y = torch.zeros_like(X)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        D.matvec(X, Y)
    for r in range(4):
        Y.zero_()
        junk = torch.randn(1 << 20, device="cuda").sum()   # eager work between replays
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(Y, Yref), f"replay {r}"
        del junk
