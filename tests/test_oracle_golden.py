"""The oracle against golden vectors written by the compiled reference.

Fixtures: tests/golden/golden_*.npz (tests/golden/make_golden.py).  Trees,
leaves and ACA ranks must match exactly; stored leaf values bit-for-bit
(SHA-256); matvec, transposed matvec (golden_transposed.npz), GMRes and CG to 1e-13 (bit-equal in the build container,
tolerance covers a different MKL code path elsewhere).
"""
import hashlib
import os

import numpy as np
import pytest

from conftest import GOLDEN, golden_names, golden_points, load_golden

NAMES = golden_names()


@pytest.fixture(scope="module", params=NAMES)
def case(request, oracle_api):
    g = load_golden(request.param)
    g["name"] = request.param
    pts = golden_points(g)
    T = oracle_api.tree(pts, int(g["bmin"]), float(g["eta2"]), int(g["kind"]))
    H = oracle_api.assemble(T, pts, int(g["kernel"]), float(g["kappa"]), float(g["eps"]),
                            weight=float(g["weight"]))
    return g, T, H


def test_tree_matches_reference(case):
    g, T, _ = case
    op, po = T.perm()
    assert np.array_equal(op, g["op_perm"])
    assert np.array_equal(po, g["po_perm"])
    lv = T.leaves()
    for f in ("b1", "n1", "b2", "n2", "adm", "i0"):
        assert np.array_equal(lv[f], g[f]), f


def test_aca_ranks_and_leaf_values(case):
    g, _, H = case
    L = H.layout()
    assert np.array_equal(L["rank"], g["rank"])
    assert len(L["data"]) == int(g["nvals"])
    h = hashlib.sha256(np.ascontiguousarray(L["data"]).tobytes()).hexdigest()
    if h != str(g["data_sha256"]):
        # different BLAS code path: fall back to per-leaf Frobenius norms
        fro = np.zeros(len(L["rank"]))
        for l in range(len(L["rank"])):
            n1, n2, k, o = (int(L[f][l]) for f in ("n1", "n2", "rank", "offset"))
            if k < 0:
                fro[l] = np.linalg.norm(L["data"][o:o + n1 * n2])
            elif k > 0:
                U = L["data"][o:o + n1 * k].reshape(k, n1).T
                V = L["data"][o + n1 * k:o + (n1 + n2) * k].reshape(k, n2).T
                fro[l] = np.linalg.norm(U @ V.T)
        np.testing.assert_allclose(fro, g["leaf_fro"], rtol=1e-12, atol=0)


def test_matvec_matches_reference(case):
    g, _, H = case
    y = H.matvec(g["x"])
    assert np.max(np.abs(y - g["y"])) <= 1e-13 * np.max(np.abs(g["y"]))
    Y = H.matmat(g["X"], d=-0.5)
    assert np.max(np.abs(Y - g["Y"])) <= 1e-13 * np.max(np.abs(g["Y"]))


def test_transposed_matvec_matches_reference(case):
    g, _, H = case
    name = g["name"]
    t = np.load(os.path.join(GOLDEN, "golden_transposed.npz"))
    y = H.matvec_t(t[f"{name}_x"], d=float(t[f"{name}_d"]), y=t[f"{name}_y0"])
    assert np.max(np.abs(y - t[f"{name}_y"])) <= 1e-13 * np.max(np.abs(t[f"{name}_y"]))


def test_gmres_matches_reference(case):
    g, _, H = case
    x, rc, e, n = H.gmres(g["b"], eps=float(g["gmres_tol"]), m=int(g["gmres_m"]),
                          nsteps=int(g["gmres_nsteps_in"]))
    assert rc == int(g["gmres_rc"])
    assert n == int(g["gmres_nsteps"])
    assert abs(e - float(g["gmres_eps"])) <= 1e-6 * float(g["gmres_eps"])
    assert np.max(np.abs(x - g["gmres_x"])) <= 1e-10 * np.max(np.abs(g["gmres_x"]))


def test_cg_matches_reference(case):
    g, _, H = case
    if "cg_x" not in g:
        pytest.skip("no CG vector for this case")
    x, rc, e, n = H.cg(g["b"], eps=float(g["cg_tol"]), nsteps=int(g["cg_nsteps_in"]))
    assert rc == int(g["cg_rc"]) and n == int(g["cg_nsteps"])
    assert np.max(np.abs(x - g["cg_x"])) <= 1e-10 * np.max(np.abs(g["cg_x"]))


def test_aca_and_addtrll_on_explicit_blocks(oracle_api):
    g = load_golden("aca")
    for c in range(4):
        A = g[f"aca{c}_A"]
        eps, kmax, i0, scale = g[f"aca{c}_par"]
        ok, U, V = oracle_api.aca_dense(A, eps, int(kmax), int(i0), scale)
        assert ok == bool(g[f"aca{c}_ok"])
        assert U.shape == g[f"aca{c}_U"].shape
        np.testing.assert_allclose(U, g[f"aca{c}_U"], rtol=0, atol=1e-13 * np.abs(A).max())
        np.testing.assert_allclose(V, g[f"aca{c}_V"], rtol=0, atol=1e-13 * np.abs(A).max())
        Uo, Vo = oracle_api.addtrll(U, V, eps)
        assert Uo.shape == g[f"trl{c}_U"].shape
        P, Pg = Uo @ Vo.T, g[f"trl{c}_U"] @ g[f"trl{c}_V"].T
        assert np.linalg.norm(P - Pg) <= 1e-12 * np.linalg.norm(Pg)


def test_acar_on_explicit_blocks(oracle_api):
    """ACAr (Include/ACA.h:657-720) on the same explicit blocks: identical to
    the reference's crosses; where it succeeds, within 20 eps of the block."""
    g = load_golden("aca")
    for c in range(4):
        A = g[f"aca{c}_A"]
        eps, kmax, i0, scale = g[f"aca{c}_par"]
        ok, U, V = oracle_api.aca_dense(A, eps, int(kmax), int(i0), scale, rows_only=True)
        assert ok == bool(g[f"acar{c}_ok"])
        assert U.shape == g[f"acar{c}_U"].shape
        np.testing.assert_allclose(U, g[f"acar{c}_U"], rtol=0, atol=1e-13 * np.abs(A).max())
        np.testing.assert_allclose(V, g[f"acar{c}_V"], rtol=0, atol=1e-13 * np.abs(A).max())
        if ok:
            assert np.linalg.norm(A - U @ V.T) <= 20 * eps * np.linalg.norm(A)


def test_acar_assembly_matches_reference(case, oracle_api):
    """Every golden matrix assembled with ACAr on its admissible leaves:
    identical ranks and bit-identical leaf values (SHA-256) to the
    reference's (tests/golden/golden_acar.npz)."""
    g, T, _ = case
    z = np.load(os.path.join(GOLDEN, "golden_acar.npz"), allow_pickle=False)
    name = g["name"]
    pts = golden_points(g)
    H = oracle_api.assemble(T, pts, int(g["kernel"]), float(g["kappa"]), float(g["eps"]),
                            weight=float(g["weight"]), rows_only=True)
    L = H.layout()
    assert np.array_equal(L["rank"], z[f"{name}_rank"])
    assert hashlib.sha256(np.ascontiguousarray(L["data"]).tobytes()).hexdigest() == str(z[f"{name}_data_sha256"])
    y = H.matvec(z[f"{name}_x"])
    assert np.max(np.abs(y - z[f"{name}_y"])) <= 1e-13 * np.max(np.abs(z[f"{name}_y"]))


# ---- triangular H-solves (LtHGeM_solve / UtHGeM_solve / HLU_solve) on the
# reference's own HLU factors (tests/golden/golden_hlu.npz)
HLU_NAMES = ["cube640", "sphere700"]


def hlu_case(name):
    z = np.load(os.path.join(GOLDEN, "golden_hlu.npz"), allow_pickle=False)
    p = name + "_"
    return {k[len(p):]: z[k] for k in z.files if k.startswith(p)}


def factor(g, which):
    return {k: g[f"{which}_{k}"] for k in ("b1", "n1", "b2", "n2", "kind", "offset", "data")}


@pytest.mark.parametrize("name", HLU_NAMES)
def test_triangular_solves_match_reference(name, oracle_api):
    g = hlu_case(name)
    L, U = factor(g, "L"), factor(g, "U")
    for which, F in (("L", L), ("U", U)):
        x = oracle_api.trih_solve(F, which, g["b"])
        assert np.max(np.abs(x - g[f"x{which}"])) <= 1e-13 * np.max(np.abs(g[f"x{which}"]))
        X = oracle_api.trih_solve(F, which, g["B"])
        assert np.max(np.abs(X - g[f"X{which}"])) <= 1e-13 * np.max(np.abs(g[f"X{which}"]))
    xlu = oracle_api.trih_solve(U, "U", oracle_api.trih_solve(L, "L", g["b"]))
    assert np.max(np.abs(xlu - g["xLU"])) <= 1e-13 * np.max(np.abs(g["xLU"]))
    # transposed solves (LtHhGeM_solve / UtHhGeM_solve), HCholesky_solve with
    # U (U^T then U) and HLUh_solve (U^T then L^T)
    for which, F, op in (("L", L, "l"), ("U", U, "u")):
        x = oracle_api.trih_solve(F, op, g["b"])
        assert np.max(np.abs(x - g[f"x{which}t"])) <= 1e-13 * np.max(np.abs(g[f"x{which}t"]))
        X = oracle_api.trih_solve(F, op, g["B"])
        assert np.max(np.abs(X - g[f"X{which}t"])) <= 1e-13 * np.max(np.abs(g[f"X{which}t"]))
    xc = oracle_api.trih_solve(U, "U", oracle_api.trih_solve(U, "u", g["b"]))
    assert np.max(np.abs(xc - g["xUtU"])) <= 1e-13 * np.max(np.abs(g["xUtU"]))
    xh = oracle_api.trih_solve(L, "l", oracle_api.trih_solve(U, "u", g["b"]))
    assert np.max(np.abs(xh - g["xLUt"])) <= 1e-13 * np.max(np.abs(g["xLUt"]))


@pytest.mark.parametrize("name", HLU_NAMES)
def test_triangular_factors_closed_form(name, oracle_api):
    """Independent of both solvers: the dense matrices the factor layouts
    stand for (permutation rows of mblock::ltr_solve, H/mblock_D.cpp:
    1183-1205) map the reference's solutions back onto b, and a dense LAPACK
    solve of them gives the oracle's answer."""
    from ahmed_amd.factors import dense_factor
    g = hlu_case(name)
    N = int(g["N"])
    ML, MU = dense_factor(factor(g, "L"), N, "L"), dense_factor(factor(g, "U"), N, "U")
    b = g["b"]
    assert np.max(np.abs(ML @ g["xL"] - b)) <= 1e-13 * np.max(np.abs(b))
    assert np.max(np.abs(MU @ g["xU"] - b)) <= 1e-12 * np.max(np.abs(b))
    assert np.max(np.abs(ML @ (MU @ g["xLU"]) - b)) <= 1e-12 * np.max(np.abs(b))
    assert np.max(np.abs(ML.T @ g["xLt"] - b)) <= 1e-13 * np.max(np.abs(b))
    assert np.max(np.abs(MU.T @ g["xUt"] - b)) <= 1e-12 * np.max(np.abs(b))
    assert np.max(np.abs(MU.T @ (MU @ g["xUtU"]) - b)) <= 1e-12 * np.max(np.abs(b))
    assert np.max(np.abs(MU.T @ (ML.T @ g["xLUt"]) - b)) <= 1e-12 * np.max(np.abs(b))
    xo = oracle_api.trih_solve(factor(g, "U"), "U", b)
    assert np.max(np.abs(xo - np.linalg.solve(MU, b))) <= 1e-11 * np.max(np.abs(xo))


@pytest.mark.parametrize("name", HLU_NAMES)
def test_hlu_fixture_matrix_is_the_oracles(name, oracle_api):
    """The fixture's GMRes ran on the reference's H-matrix of these points;
    the oracle assembles the identical matrix (same SHA-256 of all values)."""
    from ahmed_amd import geometry
    g = hlu_case(name)
    N, seed = int(g["N"]), int(g["seed"])
    pts = geometry.sphere(N, seed) if str(g["geometry"]) == "sphere" else geometry.cube_surface(N, seed)
    T = oracle_api.tree(pts, int(g["bmin"]), float(g["eta2"]))
    H = oracle_api.assemble(T, pts, 0, 0.0, float(g["eps"]), weight=float(g["weight"]))
    L = H.layout()
    assert hashlib.sha256(np.ascontiguousarray(L["data"]).tobytes()).hexdigest() == str(g["data_sha256"])
    # and the fixture's factor leaves sit on this tree's blocks
    lv = T.leaves()
    for k in ("b1", "n1", "b2", "n2"):
        assert np.array_equal(lv[k], g[f"L_{k}"])
    assert int(g["gmres_rc"]) == 0 and int(g["gmres_nsteps"]) <= 5


def test_all_core_baselines_match_serial(case):
    """The CPU baselines bench.py times on all host cores: the oracle's
    leaf-parallel product (or_hmat_matvec_omp, private y per thread, summed)
    and, where the compiled reference exists, its own leaf products run the
    same way and its GMRes on that operator, against the serial products and
    solver (same values to rounding; GMRes steps within one, true residual
    within twice the tolerance)."""
    import oracle as O
    g, T, H = case
    o = O.Oracle()
    x = g["x"]
    y1 = H.matvec(x)
    for nt in (1, 3, 8):
        yo = o.matvec_omp(H, x, nthreads=nt)
        assert np.max(np.abs(yo - y1)) <= 1e-13 * np.max(np.abs(y1)), nt
    try:
        R = O.Reference()
    except FileNotFoundError:
        return
    from conftest import golden_points
    pts = golden_points(g)
    TR = R.tree(pts, int(g["bmin"]), float(g["eta2"]), int(g["kind"]))
    L = H.layout()
    HR = R.from_data(TR, L["rank"], L["offset"], L["data"])
    yr = R.matvec_omp(HR, x, nthreads=4)
    assert np.max(np.abs(yr - g["y"])) <= 1e-13 * np.max(np.abs(g["y"]))
    tol = float(g["gmres_tol"])
    xs, rc, e, n = R.gmres_omp(HR, g["b"], eps=tol, m=int(g["gmres_m"]), nsteps=int(g["gmres_nsteps_in"]),
                               nthreads=4)
    # the thread-summed products round differently: the iterates drift apart
    # by about tol; both reach it in about as many steps
    assert rc == int(g["gmres_rc"]) and abs(n - int(g["gmres_nsteps"])) <= 1
    if rc == 0:
        b = g["b"]
        assert np.linalg.norm(b - HR.matvec(xs)) <= 2 * tol * np.linalg.norm(b)
