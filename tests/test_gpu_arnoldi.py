"""The two forms of the GMRES orthogonalisation and rotations
(ahm_hmat_set_arnoldi): chunked launches over all systems (mode 1) and one
workgroup per system doing the dot, the MGS passes, the normalisation and
the rotations in one launch (mode 2).  Mode 2 keeps the chunked form's
chunks, wavefront sums and summation order, so both give the same iterates
bit for bit: same steps, same residual history, same solution.

Cases: one system with a ragged last chunk (N = 5000), the N=8192 sphere,
the N=16384 sphere with restarts (16 chunks, the largest system of the
one-workgroup form), the N=32768 sphere (where mode 2 falls back to the
chunked launches), each eager and replayed from graphs; a batch of six
independent matrices; the preconditioned solve (to a tolerance: its
triangular solves round in a timing-dependent order)."""
import numpy as np
import pytest

from helpers import leaves_of, relerr, sphere_hmatrix

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from ahmed_amd import _lib
    return _lib.Context(0)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def solve(D, b, mode, graph, m, eps=1e-10):
    D.set_arnoldi(mode)
    D.set_graphs(graph)
    x = torch.zeros_like(b)
    r = D.gmres(b, x, eps=eps, m=m, nsteps=500, nhist=300)
    return r, x


@pytest.mark.parametrize("N,leaf,m", [(5000, 32, 30), (8192, 32, 30), (16384, 64, 8), (32768, 64, 8)])
def test_one_workgroup_form_equals_chunked(ctx, N, leaf, m):
    from ahmed_amd import _lib
    _, _, H = sphere_hmatrix(N, leaf, 1e-8)
    leaves, data = leaves_of(H)
    D = _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data)
    b = dev(H.matvec(np.random.default_rng(N).standard_normal(H.N)))
    print(f"[gpu] arnoldi forms N={N} m={m}", flush=True)
    r1, x1 = solve(D, b, 1, False, m)
    assert r1["converged"]
    if m == 8:
        assert r1["nsteps"] > m          # restarted
    # eager, then the capturing call, then a replaying call, for each form
    for mode, graph in ((2, False), (2, True), (2, True), (1, True), (1, True), (0, True)):
        r2, x2 = solve(D, b, mode, graph, m)
        assert r2["nsteps"] == r1["nsteps"] and r2["eps"] == r1["eps"], (mode, graph)
        assert r2["nmatvec"] == r1["nmatvec"]
        assert np.array_equal(r2["hist"], r1["hist"]), (mode, graph)
        assert torch.equal(x2, x1), (mode, graph)
    true = np.linalg.norm(b.cpu().numpy() - H.matvec(x1.cpu().numpy())) / np.linalg.norm(b.cpu().numpy())
    assert true <= 1.01e-10


def test_one_workgroup_form_batched(ctx):
    from ahmed_amd import _lib
    from test_gpu_batch import problems
    probs = problems()
    trees = [_lib.Tree(p, 32, 0.64) for p, _ in probs]
    HB, _ = _lib.HMatrix.assemble_batch(ctx, trees, [k for _, k in probs], eps=1e-6)
    N = HB.n_rows
    print(f"[gpu] arnoldi forms, batch of {len(probs)} (N={N})", flush=True)
    b = torch.zeros(N, dtype=torch.float64, device="cuda")
    HB.matvec(torch.ones(N, dtype=torch.float64, device="cuda"), b)
    res = []
    for mode, graph in ((1, False), (2, False), (2, True), (2, True), (1, True)):
        HB.set_arnoldi(mode)
        x = torch.zeros(N, dtype=torch.float64, device="cuda")
        r = HB.gmres_batch(b, x, eps=1e-10, m=30, nsteps=500, use_graph=graph)
        assert r["converged"].all()
        res.append((r, x))
    r1, x1 = res[0]
    for r, x in res[1:]:
        assert np.array_equal(r["nsteps"], r1["nsteps"])
        assert np.array_equal(r["eps"], r1["eps"])
        assert torch.equal(x, x1)


def test_one_workgroup_form_preconditioned(ctx, oracle_api):
    """Right-preconditioned GMRES (HLU_solve with the reference's factors),
    both forms.  The triangular solves subtract with FP64 atomics, so their
    rounding depends on timing: the forms agree in the reference's step
    count and to 1e-10 in the solution, not bit for bit."""
    from ahmed_amd import _lib
    from test_gpu_trih import NAMES, factor, golden_matrix, hlu_case
    g = hlu_case(NAMES[0])
    N = int(g["N"])
    _, leaves, data = golden_matrix(g, oracle_api)
    D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    F = [_lib.TriFactor.from_layout(ctx, N, w, factor(g, w)) for w in "LU"]
    out = []
    for mode in (1, 2):
        D.set_arnoldi(mode)
        x = dev(np.zeros(N))
        r = D.gmres(dev(g["b"]), x, eps=float(g["gmres_tol"]), m=int(g["gmres_m"]),
                    nsteps=int(g["gmres_nsteps_in"]), prec=F)
        assert r["converged"]
        out.append((r, x))
    print(f"[gpu] preconditioned GMRES, both forms: {out[0][0]['nsteps']} steps", flush=True)
    assert out[0][0]["nsteps"] == out[1][0]["nsteps"] == int(g["gmres_nsteps"])
    for _, x in out:
        assert relerr(x.cpu().numpy(), g["gmres_x"]) <= 1e-10
