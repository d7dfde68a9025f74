"""Device GMRES / CG / FP64 iterative refinement against the reference's own
solver results (golden fixtures) and the oracle, on the identical H-matrix.

GMRES and CG run the reference's algorithm (MGS, Givens, restarts) with
different FP64 summation orders, so: same convergence flag and step count as
the reference, solution within 1e-10 (relative, max-norm) of the reference's,
achieved residual within 1e-3 relative of the reference's value.
"""
import numpy as np
import pytest

from conftest import golden_names, load_golden
from helpers import golden_hmatrix, leaves_of, relerr, sphere_hmatrix

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from ahmed_amd import _lib
    return _lib.Context(0)


def upload(ctx, H):
    from ahmed_amd import _lib
    leaves, data = leaves_of(H)
    return _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("name", golden_names())
def test_gmres_matches_reference(ctx, name):
    g, T, H = golden_hmatrix(name)
    print(f"[gpu] gmres {name}", flush=True)
    D = upload(ctx, H)
    x = dev(np.zeros(H.N))
    r = D.gmres(dev(g["b"]), x, eps=float(g["gmres_tol"]), m=int(g["gmres_m"]),
                nsteps=int(g["gmres_nsteps_in"]))
    assert r["converged"] == (int(g["gmres_rc"]) == 0)
    assert r["nsteps"] == int(g["gmres_nsteps"])
    assert abs(r["eps"] - float(g["gmres_eps"])) <= 1e-3 * float(g["gmres_eps"])
    if int(g["gmres_rc"]) == 0:
        assert relerr(x.cpu().numpy(), g["gmres_x"]) <= 1e-10


@pytest.mark.parametrize("name", [n for n in golden_names() if "cg_x" in load_golden(n)])
def test_cg_matches_reference(ctx, name):
    g, T, H = golden_hmatrix(name)
    D = upload(ctx, H)
    x = dev(np.zeros(H.N))
    r = D.cg(dev(g["b"]), x, eps=float(g["cg_tol"]), nsteps=int(g["cg_nsteps_in"]))
    assert r["converged"] == (int(g["cg_rc"]) == 0)
    assert r["nsteps"] == int(g["cg_nsteps"])
    assert relerr(x.cpu().numpy(), g["cg_x"]) <= 1e-10


def test_gmres_edge_cases(ctx):
    _, T, H = sphere_hmatrix(6000, 32, 1e-8)
    D = upload(ctx, H)
    # b = 0: x = 0, no steps (solvers/GMRES.cpp:95-102)
    x = dev(np.ones(H.N))
    r = D.gmres(dev(np.zeros(H.N)), x, eps=1e-8, m=10, nsteps=10)
    assert r["converged"] and r["nsteps"] == 0 and not x.any()
    # too few steps: not converged, eps = achieved residual, nsteps unchanged
    b = H.matvec(np.ones(H.N))
    x = dev(np.zeros(H.N))
    r = D.gmres(dev(b), x, eps=1e-14, m=5, nsteps=3)
    assert not r["converged"] and r["eps"] > 1e-14 and r["nsteps"] == 3
    # the true residual of the returned x agrees with the reported one
    xr = x.cpu().numpy()
    true = np.linalg.norm(b - H.matvec(xr)) / np.linalg.norm(b)
    assert abs(true - r["eps"]) <= 1e-6 * true


def test_gmres_graph_replay_matches_eager(ctx):
    """Arnoldi steps replayed from captured hipGraphs give the one-by-one
    launches' iterates bit for bit, on the capturing call, on later calls,
    and after the product's workspace moves (graphs captured again)."""
    _, T, H = sphere_hmatrix(32768, 64, 1e-8)
    D = upload(ctx, H)
    b = dev(H.matvec(np.random.default_rng(5).standard_normal(H.N)))

    def solve(graph):
        D.set_graphs(graph)
        x = dev(np.zeros(H.N))
        r = D.gmres(b, x, eps=1e-10, m=8, nsteps=500, nhist=200)
        return r, x

    re, xe = solve(False)
    print(f"[gpu] GMRES(8) eager: {re['nsteps']} steps, {re['nmatvec']} products", flush=True)
    assert re["converged"] and re["nsteps"] > 8   # restarted
    # products: one per step and one per restart ((nsteps - 1) // m restarts)
    assert re["nmatvec"] == re["nsteps"] + (re["nsteps"] - 1) // 8
    # the implicit residuals decrease (Givens on the device) and end below eps
    h = re["hist"][:re["nsteps"]]
    assert np.all(np.diff(h) <= 1e-15) and h[-1] < 1e-10
    for tag in ("capture", "replay"):
        rg, xg = solve(True)
        print(f"[gpu] GMRES(8) graph ({tag}): {rg['nsteps']} steps", flush=True)
        assert rg["converged"] and rg["nsteps"] == re["nsteps"] and rg["eps"] == re["eps"]
        assert rg["nmatvec"] == re["nmatvec"]
        assert np.array_equal(rg["hist"], re["hist"])
        assert torch.equal(xg, xe)
    # a 32-RHS product grows the shared workspace: the steps are captured again
    X = torch.randn(32, H.N, dtype=torch.float64, device="cuda")
    Y = torch.zeros(32, H.N, dtype=torch.float64, device="cuda")
    D.matvec(X, Y)
    rg, xg = solve(True)
    assert rg["nsteps"] == re["nsteps"] and torch.equal(xg, xe)
    true = np.linalg.norm(b.cpu().numpy() - H.matvec(xe.cpu().numpy())) / np.linalg.norm(b.cpu().numpy())
    assert true <= 1.01e-10


def test_fp32_gmres_and_refinement_reach_fp64_accuracy(ctx):
    _, T, H = sphere_hmatrix(32768, 64, 1e-8)
    D = upload(ctx, H)
    D32 = D.convert()
    b = H.matvec(np.random.default_rng(4).standard_normal(H.N))
    nb = np.linalg.norm(b)
    # FP64 GMRES
    x64 = dev(np.zeros(H.N))
    r64 = D.gmres(dev(b), x64, eps=1e-10, m=50, nsteps=500)
    assert r64["converged"]
    assert np.linalg.norm(b - H.matvec(x64.cpu().numpy())) / nb <= 1.01e-10
    # GMRES on the FP32 matrix alone stalls near FP32 accuracy
    x32 = dev(np.zeros(H.N))
    r32 = D32.gmres(dev(b), x32, eps=1e-10, m=50, nsteps=300)
    true32 = np.linalg.norm(b - H.matvec(x32.cpu().numpy())) / nb
    print(f"[gpu] FP32-only GMRES: reported {r32['eps']:.2e}, true FP64 residual {true32:.2e}")
    # FP64 refinement around FP32 GMRES reaches the FP64 tolerance
    xir = dev(np.zeros(H.N))
    rir = D.gmres_ir(D32, dev(b), xir, eps=1e-10, m=50, inner_eps=1e-5)
    assert rir["converged"], rir
    true = np.linalg.norm(b - H.matvec(xir.cpu().numpy())) / nb
    print(f"[gpu] IR: {rir['outer']} refinements, {rir['inner']} inner steps, residual {true:.2e}")
    assert true <= 1.01e-10
    assert relerr(xir.cpu().numpy(), x64.cpu().numpy()) <= 1e-8
