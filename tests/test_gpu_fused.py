"""Products of 1-8 right-hand sides in one persistent launch
(ahm_hmat_set_fused): a work queue over the interleaving of X, the phase-A
tasks, their chunk sums, the row panels and their chunk sums, with waits on
the count of completed items.  The queue resets itself at the end of every
launch, so launches repeat (and replay from a captured graph) without a
memset.

Checks: bit-identical to the three-launch form (same tasks, same sums) for
FP64 and FP32, 1, 3 and 8 right-hand sides, on a sphere matrix and on the
matrix with row panels of every height; ten launches in a row; replays of a
captured graph, each checked, with eager work between them; no wait gave up
(ahm_hmat_fused_status)."""
import numpy as np
import pytest

from helpers import block_matrix, leaves_of, relerr, sphere_hmatrix

pytestmark = [pytest.mark.gpu,
              pytest.mark.skip(reason="tests of the removed work-queue form of the single launch, which did "
                                      "not finish in its first GPU run (torch.cuda.synchronize blocked > 120 s in "
                                      "test_fused_equals_three_launches_sphere[1]; cause not found); its "
                                      "successor, without waits between workgroups, is tested in "
                                      "test_gpu_single_launch.py")]
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
def test_fused_equals_three_launches_sphere(ctx, p):
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(32768, 64, 1e-8, weighted=False)
    print(f"[gpu] fused product sphere N=32768 p={p}", flush=True)
    leaves, data = leaves_of(H)
    D = _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)
    shape = (H.N,) if p == 1 else (p, H.N)
    X = torch.randn(shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(p))
    Y0, Y1 = both_forms(D, X)
    assert torch.equal(Y0, Y1)
    ref = H.matvec(X.cpu().numpy(), d=0.5) if p == 1 else H.matmat(X.cpu().numpy().T, d=0.5).T
    assert relerr(Y1.cpu().numpy(), ref) <= 1e-12
    D32 = D.convert()
    Y0, Y1 = both_forms(D32, X.float())
    assert torch.equal(Y0, Y1)
    assert not D.fused_status() and not D32.fused_status()


@pytest.mark.parametrize("p", [1, 8])
def test_fused_every_panel_height_and_repeats(ctx, p):
    from ahmed_amd import _lib
    from test_gpu_matvec import MFMA_SIZES
    N, leaves, data, A = block_matrix(MFMA_SIZES, 11)
    print(f"[gpu] fused product panel heights N={N} p={p}", flush=True)
    D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    shape = (N,) if p == 1 else (p, N)
    X = torch.randn(shape, dtype=torch.float64, device="cuda")
    Y0, Y1 = both_forms(D, X, -1.0)
    assert torch.equal(Y0, Y1)
    ref = -(A @ X.cpu().numpy().T).T if p > 1 else -(A @ X.cpu().numpy())
    assert relerr(Y1.cpu().numpy(), ref) <= 1e-12
    for _ in range(10):
        Y2 = torch.zeros_like(X)
        D.matvec(X, Y2, -1.0)
        assert torch.equal(Y2, Y1)
    assert not D.fused_status()


def test_fused_product_replays_from_a_captured_graph(ctx):
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(16384, 64, 1e-8)
    print("[gpu] fused product in a captured graph, N=16384", flush=True)
    leaves, data = leaves_of(H)
    D = _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)
    D.set_fused(True)
    X = torch.randn((8, H.N), dtype=torch.float64, device="cuda")
    Yref = torch.zeros_like(X)
    D.matvec(X, Yref)                  # sizes the workspace before capture
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
    assert not D.fused_status()
