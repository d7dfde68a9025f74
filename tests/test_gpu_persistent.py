"""Products of 1-8 right-hand sides in ONE persistent cooperative launch
(ahm_hmat_set_fused(h, 2)): as many workgroups as stay resident walk the
task lists of the three-kernel form phase after phase, with grid barriers
between the phases.

The arithmetic and its order are those of the separate launches, so every
check is bit-identity with the three-kernel form (form 0), plus the oracle /
dense reference to 1e-12: FP64 and FP32, 1, 3 and 8 right-hand sides, a sphere
matrix, the matrix with row panels of every height (ten launches in a row), a
batched matrix; GMRES on top of it (the solver's gated single-system
products) with the same iterates; a product recorded into a hipGraph falls
back to the separate launches and replays correctly."""
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


def forms_0_and_2(D, X, alpha=0.5):
    Y0 = torch.zeros_like(X)
    Y2 = torch.zeros_like(X)
    D.set_fused(0)
    D.matvec(X, Y0, alpha)
    D.set_fused(2)
    D.matvec(X, Y2, alpha)
    torch.cuda.synchronize()
    D.set_fused(0)
    return Y0, Y2


@pytest.mark.parametrize("p", [1, 3, 8])
def test_persistent_sphere(ctx, p):
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(32768, 64, 1e-8, weighted=False)
    print(f"[gpu] persistent product sphere N=32768 p={p}", flush=True)
    leaves, data = leaves_of(H)
    D = _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)
    shape = (H.N,) if p == 1 else (p, H.N)
    X = torch.randn(shape, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(p))
    Y0, Y2 = forms_0_and_2(D, X)
    assert torch.equal(Y0, Y2)
    ref = H.matvec(X.cpu().numpy(), d=0.5) if p == 1 else H.matmat(X.cpu().numpy().T, d=0.5).T
    assert relerr(Y2.cpu().numpy(), ref) <= 1e-12
    D32 = D.convert()
    Y0, Y2 = forms_0_and_2(D32, X.float())
    assert torch.equal(Y0, Y2)


@pytest.mark.parametrize("p", [1, 8])
def test_persistent_every_panel_height_and_repeats(ctx, p):
    from ahmed_amd import _lib
    from test_gpu_matvec import MFMA_SIZES
    N, leaves, data, A = block_matrix(MFMA_SIZES, 11)
    print(f"[gpu] persistent product panel heights N={N} p={p}", flush=True)
    D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    shape = (N,) if p == 1 else (p, N)
    X = torch.randn(shape, dtype=torch.float64, device="cuda")
    Y0, Y2 = forms_0_and_2(D, X, -1.0)
    assert torch.equal(Y0, Y2)
    ref = -(A @ X.cpu().numpy().T).T if p > 1 else -(A @ X.cpu().numpy())
    assert relerr(Y2.cpu().numpy(), ref) <= 1e-12
    D.set_fused(2)
    for _ in range(10):
        Y3 = torch.zeros_like(X)
        D.matvec(X, Y3, -1.0)
        assert torch.equal(Y3, Y2)
    D.set_fused(0)


def test_persistent_batched_and_autotune(ctx):
    from ahmed_amd import _lib, geometry
    trees, kernels = [], []
    rng = np.random.default_rng(6)
    for s in range(5):
        n = int(rng.integers(1500, 5000))
        pts = geometry.ellipsoid(n, (1.0, float(rng.uniform(1, 3)), float(rng.uniform(1, 3))), seed=50 + s)
        trees.append(_lib.Tree(pts, 32, 0.64))
        kernels.append((_lib.KERNEL_YUKAWA, float(rng.uniform(0, 4)), 1.0 / n))
    HB, _ = _lib.HMatrix.assemble_batch(ctx, trees, kernels, eps=1e-6)
    N = HB.n_rows
    print(f"[gpu] persistent product batched N={N}", flush=True)
    X = torch.randn((8, N), dtype=torch.float64, device="cuda")
    Y0, Y2 = forms_0_and_2(HB, X, 1.0)
    assert torch.equal(Y0, Y2)
    chunk, ms, form = HB.autotune(p=8)
    print(f"[gpu] autotune: chunk {chunk}, {ms:.3f} ms, form {form}", flush=True)
    assert form in (0, 1, 2)
    Y3 = torch.zeros_like(X)
    HB.matvec(X, Y3, 1.0)
    assert relerr(Y3.cpu().numpy(), Y0.cpu().numpy()) <= 1e-12


def test_persistent_under_gmres_and_graphs(ctx):
    """GMRES's single-system products (gated once the system converges) in
    the persistent form give the three-kernel iterates bit for bit; with
    captured steps the products are recorded as separate launches."""
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(16384, 64, 1e-8)
    print("[gpu] persistent product under GMRES, N=16384", flush=True)
    leaves, data = leaves_of(H)
    D = _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)
    b = torch.as_tensor(H.matvec(np.ones(H.N)), device="cuda")
    xs = {}
    for form, graphs in ((0, False), (2, False), (2, True)):
        D.set_fused(form)
        D.set_graphs(graphs)
        for rep in range(2):
            x = torch.zeros(H.N, dtype=torch.float64, device="cuda")
            r = D.gmres(b, x, eps=1e-10, m=30, nsteps=300)
            assert r["converged"]
        xs[(form, graphs)] = x.clone()
    D.set_graphs(False)
    D.set_fused(0)
    assert torch.equal(xs[(0, False)], xs[(2, False)])
    assert torch.equal(xs[(0, False)], xs[(2, True)])
    # a product captured with form 2 set replays correctly (separate launches)
    D.set_fused(2)
    X = torch.randn(H.N, dtype=torch.float64, device="cuda")
    Yref = torch.zeros_like(X)
    D.matvec(X, Yref)
    Y = torch.zeros_like(X)
    s = torch.cuda.Stream()
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g, stream=s):
        D.matvec(X, Y)
    for rep in range(3):
        Y.zero_()
        junk = torch.randn(1 << 20, device="cuda").sum()
        g.replay()
        torch.cuda.synchronize()
        assert torch.equal(Y, Yref), f"replay {rep}"
        del junk
    D.set_fused(0)
