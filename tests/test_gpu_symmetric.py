"""Symmetric H-matrices on the GPU: the path of the reference's BEM example
(Examples/PrecondBEM.cpp:360-375,450) -- a symmetric block tree
(subdivide_sym, cluster/blcluster.cpp:90-113), the symmetric assembly
(apprx_sym, Include/apprx.h:87-131: ACAr, packed Hermitian diagonal leaves),
the product mltaHeHVec / mltaHeHGeM (H/mltaGeHVec.cpp:420-472) and CG / GMRes
on HeHMatrix (Include/matrix.h:307-323).

Bars: identical H-matrix (leaves uploaded from the oracle, whose symmetric
assembly is bit-identical to the reference's) -> products within 1e-12 of the
reference's golden y / Y; CG and GMRES with the reference's step counts,
GMRES solutions within 1e-10 and CG solutions within 10 tol of the reference's
(CG carries the rounding of another summation order to about its tolerance)
with true residuals below 2 tol; the GPU's own symmetric assembly -> ranks equal to
the reference's on >= 97 % of the leaves (+-2 elsewhere), every leaf within
20 eps of the kernel block, the diagonal leaves exact, and its product within
1e-12 of the oracle's product of the downloaded leaves.  FP32 storage within
2e-6 of the FP64 product; graph replay equal to eager launches bit for bit.
"""
import os

import numpy as np
import pytest

import oracle as O
from conftest import GOLDEN
from helpers import leaves_of, oracle_api, relerr

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

G = np.load(os.path.join(GOLDEN, "golden_heh.npz"), allow_pickle=False)
NAMES = sorted({k[:-2] for k in G.files if k.endswith("_N")})


def field(name, f):
    return G[f"{name}_{f}"]


def points(name):
    from ahmed_amd import geometry
    geo, N, seed = str(field(name, "geometry")), int(field(name, "N")), int(field(name, "seed"))
    return geometry.sphere(N, seed) if geo == "sphere" else geometry.cube_surface(N, seed)


@pytest.fixture(scope="module")
def ctx():
    from ahmed_amd import _lib
    assert torch.cuda.is_available()
    return _lib.Context(0)


def dev(a, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device="cuda")


def oracle_sym(name):
    o = oracle_api()
    pts = points(name)
    T = o.tree(pts, int(field(name, "bmin")), float(field(name, "eta2")), int(field(name, "kind")) | O.TREE_SYM)
    H = o.assemble(T, pts, int(field(name, "kernel")), float(field(name, "kappa")), float(field(name, "eps")),
                   weight=float(field(name, "weight")))
    return pts, T, H


def upload(ctx, H, keep=False):
    from ahmed_amd import _lib
    leaves, data = leaves_of(H)
    return _lib.HMatrix.upload(ctx, H.N, H.N, leaves, data, keep_leaves=keep)


@pytest.mark.parametrize("name", NAMES)
def test_symmetric_products_match_reference(ctx, name):
    print(f"[gpu] symmetric product {name}", flush=True)
    _, _, H = oracle_sym(name)
    D = upload(ctx, H)
    assert D.stats()["symmetric"] == 1
    N = H.N
    y = dev(field(name, "y0"))
    D.matvec(dev(field(name, "x")), y, alpha=float(field(name, "d")))
    assert relerr(y.cpu().numpy(), field(name, "y")) <= 1e-12
    # the transposed product of a symmetric matrix is the product
    yt = dev(field(name, "y0"))
    D.matvec_t(dev(field(name, "x")), yt, alpha=float(field(name, "d")))
    assert torch.equal(yt, y)
    # three right-hand sides (mltaHeHGeM): rows of a (p, N) tensor
    X = dev(field(name, "X").T.copy())
    Y = torch.zeros(3, N, dtype=torch.float64, device="cuda")
    D.matvec(X, Y, alpha=0.5)
    assert relerr(Y.cpu().numpy().T, field(name, "Y")) <= 1e-12


@pytest.mark.parametrize("p", [12, 33])
def test_symmetric_multi_rhs_and_fp32(ctx, p):
    """Wider products (MFMA forward part) against the oracle, and FP32
    storage against the FP64 product."""
    name = NAMES[0]
    print(f"[gpu] symmetric p={p} {name}", flush=True)
    _, _, H = oracle_sym(name)
    D = upload(ctx, H)
    N = H.N
    X = np.random.default_rng(p).standard_normal((N, p))
    Y = torch.zeros(p, N, dtype=torch.float64, device="cuda")
    D.matvec(dev(X.T.copy()), Y)
    ref = H.matmat(X)
    assert relerr(Y.cpu().numpy().T, ref) <= 1e-12
    D32 = D.convert()
    assert D32.stats()["symmetric"] == 1
    Y32 = torch.zeros(p, N, dtype=torch.float32, device="cuda")
    D32.matvec(dev(X.T.copy(), torch.float32), Y32)
    err = relerr(Y32.double().cpu().numpy().T, ref)
    print(f"[gpu]   FP32 storage rel.err {err:.2e}")
    assert err <= 2e-6


@pytest.mark.parametrize("name", NAMES)
def test_symmetric_cg_and_gmres_match_reference(ctx, name):
    print(f"[gpu] symmetric CG/GMRES {name}", flush=True)
    _, _, H = oracle_sym(name)
    D = upload(ctx, H)
    b = dev(field(name, "b"))
    x = torch.zeros(H.N, dtype=torch.float64, device="cuda")
    tol = float(field(name, "cg_tol"))
    r = D.cg(b, x, eps=tol, nsteps=int(field(name, "cg_nsteps_in")))
    assert r["converged"] == (int(field(name, "cg_rc")) == 0)
    assert r["nsteps"] == int(field(name, "cg_nsteps"))
    # both solutions meet the tolerance; they differ by the rounding of the
    # iterates (another summation order), which CG carries to ~tol
    xs = x.cpu().numpy()
    assert relerr(xs, field(name, "cg_x")) <= 10 * tol
    bh = field(name, "b")
    assert np.linalg.norm(bh - H.matvec(xs)) <= 2 * tol * np.linalg.norm(bh)
    if f"{name}_gmres_x" in G.files:
        for graphs in (False, True):
            D.set_graphs(graphs)
            xs = []
            for rep in range(2):   # the second graph call replays the captured steps
                x = torch.zeros(H.N, dtype=torch.float64, device="cuda")
                r = D.gmres(b, x, eps=float(field(name, "gmres_tol")), m=int(field(name, "gmres_m")),
                            nsteps=int(field(name, "gmres_nsteps_in")))
                assert r["converged"] == (int(field(name, "gmres_rc")) == 0)
                assert r["nsteps"] == int(field(name, "gmres_nsteps"))
                assert relerr(x.cpu().numpy(), field(name, "gmres_x")) <= 1e-10
                xs.append(x.clone())
            assert torch.equal(xs[0], xs[1])
            if graphs:
                assert torch.equal(xs[1], x_eager)
            else:
                x_eager = xs[0]
        D.set_graphs(False)


@pytest.mark.parametrize("name", NAMES)
def test_gpu_symmetric_assembly(ctx, name):
    from ahmed_amd import _lib
    from test_gpu_assembly import leaf_blocks
    print(f"[gpu] symmetric assembly {name}", flush=True)
    pts = points(name)
    T = _lib.Tree(pts, int(field(name, "bmin")), float(field(name, "eta2")), int(field(name, "kind")) | _lib.TREE_SYM)
    D, st = _lib.HMatrix.assemble(ctx, T, int(field(name, "kernel")), float(field(name, "kappa")),
                                  float(field(name, "weight")), float(field(name, "eps")), keep_leaves=True)
    assert D.stats()["symmetric"] == 1
    leaves, data = D.download()
    hem = (leaves["flags"] & _lib.LEAF_HEM) != 0
    assert np.array_equal(hem, leaves["b1"] == leaves["b2"])
    code = np.where(hem, O.HEM, leaves["rank"])
    gr = field(name, "rank")
    same = np.mean(code == gr)
    print(f"[gpu]   ranks equal to the reference's on {same:.4f} of the leaves")
    assert same >= 0.97, same
    assert np.max(np.abs(code.astype(int) - gr.astype(int))) <= 2
    op, _ = T.perm()
    o = oracle_api()
    eps = float(field(name, "eps"))
    kern, kappa, w = int(field(name, "kernel")), float(field(name, "kappa")), float(field(name, "weight"))
    for d in leaves[hem]:
        b, n, off = int(d["b1"]), int(d["n1"]), int(d["offset"])
        K = o.kernel_block(kern, kappa, pts, op, b, n, b, n, weight=w)
        iu = np.triu_indices(n)
        packed = data[off + iu[0] + iu[1] * (iu[1] + 1) // 2]
        assert np.max(np.abs(packed - K[iu])) <= 4e-15 * np.max(np.abs(K))
    for d, blk in leaf_blocks(leaves[~hem], data):
        b1, n1, b2, n2 = (int(d[f]) for f in ("b1", "n1", "b2", "n2"))
        K = o.kernel_block(kern, kappa, pts, op, b1, n1, b2, n2, weight=w)
        if d["rank"] < 0:
            assert np.max(np.abs(blk - K)) <= 4e-15 * np.max(np.abs(K))
        else:
            assert np.linalg.norm(blk - K) <= 20 * eps * np.linalg.norm(K)
    # the identical (downloaded) matrix through the oracle's mltaHeHVec
    Hc = o.from_leaves(T.N, leaves["b1"], leaves["n1"], leaves["b2"], leaves["n2"], code, leaves["offset"], data)
    x = field(name, "x")
    y = torch.zeros(T.N, dtype=torch.float64, device="cuda")
    D.matvec(dev(x), y)
    assert relerr(y.cpu().numpy(), Hc.matvec(x)) <= 1e-12
    # and the reference's own product of its own assembly, to the ACA tolerance
    yr = field(name, "y") - field(name, "y0")
    assert relerr(float(field(name, "d")) * y.cpu().numpy(), yr) <= 50 * eps


def test_symmetric_upload_rejects_malformed_packed_leaves(ctx):
    from ahmed_amd import _lib
    L = np.zeros(1, _lib.LEAF_DTYPE)
    L["n1"], L["n2"], L["rank"], L["flags"] = 4, 5, -1, _lib.LEAF_HEM
    with pytest.raises(_lib.AhmError):
        _lib.HMatrix.upload(ctx, 4, 5, L, np.zeros(20))
    L["n2"], L["rank"] = 4, 2
    with pytest.raises(_lib.AhmError):
        _lib.HMatrix.upload(ctx, 4, 4, L, np.zeros(20))


def test_drop_in_symmetric_program():
    """tests/cpp/drop_in runs the reference's example sequence for the
    symmetric matrix too: subdivide_sym, matgenHeH_gpu, mltaHeHVec against
    direct kernel sums, CG through HeHMatrix (see test_gpu_surface)."""
    import subprocess
    from conftest import ROOT
    exe = os.path.join(ROOT, "tests", "cpp", "drop_in")
    assert os.path.exists(exe), "build first: make -C ahmed_amd/csrc"
    print("[gpu] tests/cpp/drop_in sym: bbx N=8000", flush=True)
    p = subprocess.run([exe, "sym"], capture_output=True, text=True, timeout=240)
    print(p.stdout, p.stderr)
    assert p.returncode == 0
    assert "drop_in sym OK" in p.stdout
