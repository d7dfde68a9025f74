"""GPU triangular H-solves (LtHGeM_solve / UtHGeM_solve, H/TH_solve.cpp:18-91;
HLU_solve, Include/H.h:1383-1388) through the C ABI.

* on the reference's own HLU factors (tests/golden/golden_hlu.npz): L\\b,
  U\\b, (LU)\\b, the transposed solves (LtHhGeM_solve, UtHhGeM_solve,
  HCholesky_solve with U, HLUh_solve) and three right-hand sides against the
  reference's outputs;
* on block factors of larger H-matrices (ahmed_amd.factors) against the
  oracle's restatement;
* malformed factors are rejected on the host before anything runs.
Bar: FP64 within 1e-12 relative (max-norm)."""
import numpy as np
import pytest

from helpers import relerr, sphere_hmatrix

pytestmark = pytest.mark.gpu

torch = pytest.importorskip("torch")

NAMES = ["cube640", "sphere700"]


@pytest.fixture(scope="module")
def ctx():
    from ahmed_amd import _lib
    assert torch.cuda.is_available()
    return _lib.Context(0)


def dev(a):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=torch.float64, device="cuda")


def hlu_case(name):
    import os
    from conftest import GOLDEN
    z = np.load(os.path.join(GOLDEN, "golden_hlu.npz"), allow_pickle=False)
    p = name + "_"
    return {k[len(p):]: z[k] for k in z.files if k.startswith(p)}


def factor(g, which):
    return {k: g[f"{which}_{k}"] for k in ("b1", "n1", "b2", "n2", "kind", "offset", "data")}


@pytest.mark.parametrize("name", NAMES)
def test_trih_matches_reference_golden(ctx, name):
    from ahmed_amd import _lib
    g = hlu_case(name)
    N = int(g["N"])
    print(f"[gpu] trih golden {name} N={N}", flush=True)
    F = {w: _lib.TriFactor.from_layout(ctx, N, w, factor(g, w)) for w in "LU"}
    for w in "LU":
        x = dev(g["b"])
        F[w].solve(x)
        assert relerr(x.cpu().numpy(), g[f"x{w}"]) <= 1e-12, w
        X = dev(g["B"].T)          # (p, N): row j = right-hand side j
        F[w].solve(X)
        assert relerr(X.cpu().numpy().T, g[f"X{w}"]) <= 1e-12, w
    x = dev(g["b"])
    F["L"].solve(x)
    F["U"].solve(x)
    assert relerr(x.cpu().numpy(), g["xLU"]) <= 1e-12
    info = F["L"].info()
    assert info["n_diag"] == int((g["L_kind"] == -3).sum())
    # transposed solves: the factors set up with 'l' / 'u'
    Ft = {w: _lib.TriFactor.from_layout(ctx, N, w.lower(), factor(g, w)) for w in "LU"}
    for w in "LU":
        x = dev(g["b"])
        Ft[w].solve(x)
        assert relerr(x.cpu().numpy(), g[f"x{w}t"]) <= 1e-12, w + "t"
        X = dev(g["B"].T)
        Ft[w].solve(X)
        assert relerr(X.cpu().numpy().T, g[f"X{w}t"]) <= 1e-12, w + "t"
    x = dev(g["b"])            # HCholesky_solve with U: U^T then U
    Ft["U"].solve(x)
    F["U"].solve(x)
    assert relerr(x.cpu().numpy(), g["xUtU"]) <= 1e-12
    x = dev(g["b"])            # HLUh_solve: U^T then L^T
    Ft["U"].solve(x)
    Ft["L"].solve(x)
    assert relerr(x.cpu().numpy(), g["xLUt"]) <= 1e-12


@pytest.mark.parametrize("N,eps", [(16384, 1e-8), (32768, 1e-6)])
def test_trih_block_factors_match_oracle(ctx, N, eps, oracle_api):
    from ahmed_amd import _lib
    from ahmed_amd.factors import block_factors
    _, _, H = sphere_hmatrix(N, 64, eps)
    Lf, Uf = block_factors(H.layout())
    rng = np.random.default_rng(N)
    b = rng.standard_normal(N)
    B = rng.standard_normal((N, 2))
    print(f"[gpu] trih block factors N={N}", flush=True)
    for w, F in (("L", Lf), ("U", Uf), ("l", Lf), ("u", Uf)):
        T = _lib.TriFactor.from_layout(ctx, N, w, F)
        ref = oracle_api.trih_solve(F, w, b)
        for rep in range(2):   # the counters and slots are reset per call
            x = dev(b)
            T.solve(x)
            err = relerr(x.cpu().numpy(), ref)
            print(f"[gpu]   {w} rep {rep}: rel.err {err:.2e}  {T.info()}")
            assert err <= 1e-12
        X = dev(B.T)
        T.solve(X)
        assert relerr(X.cpu().numpy().T, oracle_api.trih_solve(F, w, B)) <= 1e-12
    # more right-hand sides than one launch takes (16 chains): two launches
    B17 = rng.standard_normal((N, 17))
    T = _lib.TriFactor.from_layout(ctx, N, "L", Lf)
    X = dev(B17.T)
    T.solve(X)
    assert relerr(X.cpu().numpy().T, oracle_api.trih_solve(Lf, "L", B17)) <= 1e-12


def test_trih_rejects_malformed_factors(ctx):
    from ahmed_amd import _lib
    g = hlu_case("cube640")
    N = int(g["N"])
    L = factor(g, "L")
    leaves, data = _lib.factor_leaves(L)
    # a diagonal leaf missing: the diagonal no longer tiles [0, N)
    d = np.flatnonzero(leaves["flags"] == _lib.LEAF_LTM)
    with pytest.raises(_lib.AhmError, match="tile"):
        _lib.TriFactor.upload(ctx, N, "L", np.delete(leaves, d[3]), data)
    # the lower factor handed over as an upper one
    with pytest.raises(_lib.AhmError):
        _lib.TriFactor.upload(ctx, N, "U", leaves, data)
    with pytest.raises(_lib.AhmError):
        _lib.TriFactor.upload(ctx, N, "u", leaves, data)
    with pytest.raises(_lib.AhmError, match="uplo"):
        _lib.TriFactor.upload(ctx, N, "X", leaves, data)
    # a permutation entry out of range
    bad = data.copy()
    bad[int(leaves["offset"][d[0]])] = 1e6
    with pytest.raises(_lib.AhmError, match="permutation"):
        _lib.TriFactor.upload(ctx, N, "L", leaves, bad)
    # an off-diagonal leaf on the wrong side of the diagonal
    off = np.flatnonzero(leaves["flags"] == 0)[0]
    swapped = leaves.copy()
    swapped["b1"][off], swapped["b2"][off] = leaves["b2"][off], leaves["b1"][off]
    swapped["n1"][off], swapped["n2"][off] = leaves["n2"][off], leaves["n1"][off]
    with pytest.raises(_lib.AhmError, match="below"):
        _lib.TriFactor.upload(ctx, N, "L", swapped, data)
    # p = 0 is a no-op
    T = _lib.TriFactor.upload(ctx, N, "L", leaves, data)
    x = dev(np.ones((0, N)))
    T.solve(x)


def golden_matrix(g, oracle_api):
    """The reference's H-matrix of the fixture, assembled by the oracle (in
    the build container its values hash to the fixture's digest, asserted by
    tests/test_oracle_golden.py; another host's MKL may round differently)."""
    from ahmed_amd import _lib, geometry
    N, seed = int(g["N"]), int(g["seed"])
    pts = geometry.sphere(N, seed) if str(g["geometry"]) == "sphere" else geometry.cube_surface(N, seed)
    T = oracle_api.tree(pts, int(g["bmin"]), float(g["eta2"]))
    H = oracle_api.assemble(T, pts, 0, 0.0, float(g["eps"]), weight=float(g["weight"]))
    L = H.layout()
    leaves = _lib.leaf_array(L["b1"], L["n1"], L["b2"], L["n2"], L["rank"], L["offset"])
    return H, leaves, L["data"]


@pytest.mark.parametrize("name", NAMES)
def test_hlu_preconditioned_gmres_matches_reference(ctx, name, oracle_api):
    from ahmed_amd import _lib
    g = hlu_case(name)
    N = int(g["N"])
    H, leaves, data = golden_matrix(g, oracle_api)
    D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    F = [_lib.TriFactor.from_layout(ctx, N, w, factor(g, w)) for w in "LU"]
    x = dev(np.zeros(N))
    r = D.gmres(dev(g["b"]), x, eps=float(g["gmres_tol"]), m=int(g["gmres_m"]),
                nsteps=int(g["gmres_nsteps_in"]), prec=F)
    print(f"[gpu] HLU-preconditioned GMRes {name}: {r['nsteps']} steps, eps {r['eps']:.2e} "
          f"(reference {int(g['gmres_nsteps'])}, {float(g['gmres_eps']):.2e})")
    assert r["converged"] and r["nsteps"] == int(g["gmres_nsteps"])
    assert relerr(x.cpu().numpy(), g["gmres_x"]) <= 1e-10
    # without the preconditioner the same system takes several times the steps
    x0 = dev(np.zeros(N))
    r0 = D.gmres(dev(g["b"]), x0, eps=float(g["gmres_tol"]), m=int(g["gmres_m"]), nsteps=400)
    assert r0["converged"] and r0["nsteps"] > 2 * r["nsteps"]


def test_hlu_preconditioned_cg_converges(ctx, oracle_api):
    """CG with precond_apply = HLU_solve (solvers/CG.cpp:69-71).  The
    Laplace single-layer matrix is symmetric; its LU preconditioner is not
    exactly, so only convergence in far fewer steps is asserted."""
    from ahmed_amd import _lib
    g = hlu_case("cube640")
    N = int(g["N"])
    H, leaves, data = golden_matrix(g, oracle_api)
    D = _lib.HMatrix.upload(ctx, N, N, leaves, data)
    F = [_lib.TriFactor.from_layout(ctx, N, w, factor(g, w)) for w in "LU"]
    b = g["b"]
    x = dev(np.zeros(N))
    r = D.cg(dev(b), x, eps=1e-8, nsteps=200, prec=F)
    x0 = dev(np.zeros(N))
    r0 = D.cg(dev(b), x0, eps=1e-8, nsteps=400)
    print(f"[gpu] HLU-preconditioned CG: {r['nsteps']} steps (plain {r0['nsteps']})")
    assert r["converged"] and r0["converged"] and r["nsteps"] < r0["nsteps"]
    Ax = H.matvec(x.cpu().numpy())
    assert np.linalg.norm(Ax - b) <= 2e-8 * np.linalg.norm(b)
