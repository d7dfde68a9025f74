"""Batched independent H-matrices (the uneven multi-matrix configuration):
one batched GPU assembly of several matrices with different sizes, shapes
and kernels, stacked block-diagonally; products and GMRES for all systems
at once (lockstep Arnoldi steps, optionally hipGraph-captured).

Checks: every diagonal block equals the single-matrix assembly bit for bit;
the batched product equals the per-matrix products bit for bit; batched
GMRES gives every system the step count and, bit for bit, the solution of
its single solve (same reduction chunks, products and rotations; converged
systems frozen on the device); graph-replayed and eager batched GMRES agree
bit for bit (every replay is checked through the final result and steps)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


@pytest.fixture(scope="module")
def ctx():
    from ahmed_amd import _lib
    return _lib.Context(0)


def problems(n_mat=6, seed=3):
    from ahmed_amd import geometry
    rng = np.random.default_rng(seed)
    out = []
    for s in range(n_mat):
        N = int(np.exp(rng.uniform(np.log(1500), np.log(6000))))
        axes = (1.0, rng.uniform(1, 4), rng.uniform(1, 4))
        kappa = rng.uniform(0, 5)
        pts = geometry.ellipsoid(N, axes, seed=100 + s)
        area = 4 * np.pi * ((axes[0] * axes[1]) ** 1.6 + (axes[0] * axes[2]) ** 1.6 + (axes[1] * axes[2]) ** 1.6) ** (1 / 1.6) / 3 ** (1 / 1.6)
        out.append((pts, (1, kappa, area / N)))
    return out


def test_batched_assembly_matvec_and_gmres(ctx):
    from ahmed_amd import _lib
    probs = problems()
    trees = [_lib.Tree(p, 32, 0.64) for p, _ in probs]
    kernels = [k for _, k in probs]
    HB, st = _lib.HMatrix.assemble_batch(ctx, trees, kernels, eps=1e-6, keep_leaves=True)
    off = HB.segments()
    assert len(off) == len(probs) + 1 and off[-1] == HB.n_rows
    lb, db = HB.download()
    singles = []
    for s, (t, k) in enumerate(zip(trees, kernels)):
        Hs, _ = _lib.HMatrix.assemble(ctx, t, k[0], k[1], k[2], 1e-6, keep_leaves=True)
        singles.append(Hs)
        ls, ds = Hs.download()
        sel = (lb["b1"] >= off[s]) & (lb["b1"] < off[s + 1])
        assert sel.sum() == len(ls)
        assert np.array_equal(lb["rank"][sel], ls["rank"])
        for a, b in zip(np.flatnonzero(sel), range(len(ls))):
            n1, n2, k = int(ls["n1"][b]), int(ls["n2"][b]), int(ls["rank"][b])
            nv = n1 * n2 if k < 0 else k * (n1 + n2)
            oa, ob = int(lb["offset"][a]), int(ls["offset"][b])
            assert np.array_equal(db[oa:oa + nv], ds[ob:ob + nv])
    # one batched product == the per-matrix products, bit for bit
    N = HB.n_rows
    x = torch.randn(N, dtype=torch.float64, device="cuda")
    y = torch.zeros(N, dtype=torch.float64, device="cuda")
    HB.matvec(x, y)
    for s, Hs in enumerate(singles):
        ys = torch.zeros(int(off[s + 1] - off[s]), dtype=torch.float64, device="cuda")
        Hs.matvec(x[off[s]:off[s + 1]].contiguous(), ys)
        assert torch.equal(ys, y[off[s]:off[s + 1]])
    # 8 right-hand sides per system
    X = torch.randn(8, N, dtype=torch.float64, device="cuda")
    Y = torch.zeros(8, N, dtype=torch.float64, device="cuda")
    HB.matvec(X, Y)
    for s, Hs in enumerate(singles):
        Ys = torch.zeros(8, int(off[s + 1] - off[s]), dtype=torch.float64, device="cuda")
        Hs.matvec(X[:, off[s]:off[s + 1]].contiguous(), Ys)
        assert torch.equal(Ys, Y[:, off[s]:off[s + 1]])
    # batched GMRES vs single solves
    b = torch.zeros(N, dtype=torch.float64, device="cuda")
    HB.matvec(torch.ones(N, dtype=torch.float64, device="cuda"), b)
    xg = torch.zeros(N, dtype=torch.float64, device="cuda")
    rg = HB.gmres_batch(b, xg, eps=1e-10, m=30, nsteps=500, use_graph=True)
    xe = torch.zeros(N, dtype=torch.float64, device="cuda")
    re = HB.gmres_batch(b, xe, eps=1e-10, m=30, nsteps=500, use_graph=False)
    assert rg["converged"].all() and re["converged"].all()
    assert torch.equal(xg, xe) and np.array_equal(rg["nsteps"], re["nsteps"])
    for s, Hs in enumerate(singles):
        xs = torch.zeros(int(off[s + 1] - off[s]), dtype=torch.float64, device="cuda")
        r1 = Hs.gmres(b[off[s]:off[s + 1]].contiguous(), xs, eps=1e-10, m=30, nsteps=500)
        # same chunks, products and rotations as the system alone: bit for bit
        assert r1["converged"]
        assert int(r1["nsteps"]) == int(rg["nsteps"][s])
        assert torch.equal(xs, xg[off[s]:off[s + 1]])
    print(f"[gpu] batch of {len(probs)}: steps {rg['nsteps'].tolist()}")
