"""The product's C++ cluster / block-cluster trees (include/ahmed, via the C
ABI) against the oracle and the reference's golden trees: identical
permutations, leaves, admissibility flags and ACA start rows."""
import numpy as np
import pytest

import oracle as O
from ahmed_amd import geometry
from conftest import golden_names, golden_points, load_golden


@pytest.mark.parametrize("name", golden_names())
def test_tree_matches_reference_golden(name):
    from ahmed_amd import _lib
    g = load_golden(name)
    T = _lib.Tree(golden_points(g), int(g["bmin"]), float(g["eta2"]), int(g["kind"]))
    op, po = T.perm()
    assert np.array_equal(op, g["op_perm"]) and np.array_equal(po, g["po_perm"])
    lv = T.leaves()
    for f in ("b1", "n1", "b2", "n2", "adm", "i0"):
        assert np.array_equal(lv[f], g[f]), f


@pytest.mark.parametrize("kind", [O.TREE_BBX, O.TREE_PCA])
@pytest.mark.parametrize("geo,N,bmin", [("sphere", 20000, 64), ("cube", 12000, 32), ("ellipsoid", 9000, 16)])
def test_tree_matches_oracle(oracle_api, kind, geo, N, bmin):
    from ahmed_amd import _lib
    pts = {"sphere": lambda: geometry.sphere(N, 21), "cube": lambda: geometry.cube_surface(N, 22),
           "ellipsoid": lambda: geometry.ellipsoid(N, (1, 3, 2), 23)}[geo]()
    T = _lib.Tree(pts, bmin, 0.64, kind)
    R = oracle_api.tree(pts, bmin, 0.64, kind)
    assert np.array_equal(T.perm()[0], R.perm()[0])
    a, b = T.leaves(), R.leaves()
    for f in a:
        assert np.array_equal(a[f], b[f]), f


def test_single_point_and_tiny_trees(oracle_api):
    from ahmed_amd import _lib
    for N in (1, 2, 7):
        pts = geometry.sphere(N, 1)
        T = _lib.Tree(pts, 16, 0.64)
        lv, ref = T.leaves(), oracle_api.tree(pts, 16, 0.64).leaves()
        assert T.n_leaves == 1 and lv["n1"][0] == N
        # a one-point diagonal block is admissible through the reference's weak
        # admissibility (its box touches itself exactly)
        assert lv["adm"][0] == ref["adm"][0] == (1 if N == 1 else 0)


def test_symmetric_tree_matches_reference_golden():
    """subdivide_sym through the C ABI (kind | AHM_TREE_SYM) against the
    reference's symmetric trees (golden_heh.npz; cluster/blcluster.cpp:90-113)."""
    import os
    from ahmed_amd import _lib
    from conftest import GOLDEN
    G = np.load(os.path.join(GOLDEN, "golden_heh.npz"), allow_pickle=False)
    names = sorted({k[:-2] for k in G.files if k.endswith("_N")})
    assert names
    for name in names:
        geo, N, seed = str(G[f"{name}_geometry"]), int(G[f"{name}_N"]), int(G[f"{name}_seed"])
        pts = geometry.sphere(N, seed) if geo == "sphere" else geometry.cube_surface(N, seed)
        T = _lib.Tree(pts, int(G[f"{name}_bmin"]), float(G[f"{name}_eta2"]), int(G[f"{name}_kind"]) | _lib.TREE_SYM)
        assert T.sym
        op, po = T.perm()
        assert np.array_equal(op, G[f"{name}_op_perm"]) and np.array_equal(po, G[f"{name}_po_perm"])
        lv = T.leaves()
        for f in ("b1", "n1", "b2", "n2", "adm", "i0"):
            assert np.array_equal(lv[f], G[f"{name}_{f}"]), (name, f)
