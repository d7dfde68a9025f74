"""MLP rank-slab predictor: parameter packing, BF16 inference on the GPU
against a numpy restatement, and the invariance of the assembled matrix to
the slab sizes (ACA continues past under-predicted slabs)."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")


def bf16(a):
    u = np.asarray(a, np.float32).view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32)


def unpack(p):
    from ahmed_amd.rank_mlp import NF, NH
    o = 0
    W1 = p[o:o + NF * NH].reshape(NF, NH); o += NF * NH
    b1 = p[o:o + NH]; o += NH
    W2 = p[o:o + NH * NH].reshape(NH, NH); o += NH * NH
    b2 = p[o:o + NH]; o += NH
    W3 = p[o:o + NH]; o += NH
    b3 = p[o]; o += 1
    mu = p[o:o + NF]; o += NF
    sd = p[o:o + NF]
    return W1, b1, W2, b2, W3, b3, mu, sd


def forward_bf16(feats, p):
    W1, b1, W2, b2, W3, b3, mu, sd = unpack(p)
    x = bf16((feats - mu) / sd)
    h1 = bf16(np.maximum(x @ bf16(W1) + b1, 0))
    h2 = bf16(np.maximum(h1 @ bf16(W2) + b2, 0))
    return h2 @ bf16(W3) + b3


def test_packed_params_match_fp32_network():
    """The committed parameters evaluate (in FP32, via torch) to the BF16
    restatement within BF16 accuracy: pins the packing order."""
    from ahmed_amd.rank_mlp import load_params, NF, NH
    p = load_params()
    W1, b1, W2, b2, W3, b3, mu, sd = unpack(p)
    net = torch.nn.Sequential(torch.nn.Linear(NF, NH), torch.nn.ReLU(), torch.nn.Linear(NH, NH),
                              torch.nn.ReLU(), torch.nn.Linear(NH, 1))
    with torch.no_grad():
        net[0].weight.copy_(torch.tensor(W1.T)); net[0].bias.copy_(torch.tensor(b1))
        net[2].weight.copy_(torch.tensor(W2.T)); net[2].bias.copy_(torch.tensor(b2))
        net[4].weight.copy_(torch.tensor(W3[None])); net[4].bias.copy_(torch.tensor([b3]))
    rng = np.random.default_rng(0)
    feats = (mu + sd * rng.standard_normal((500, NF))).astype(np.float32)
    ref = net(torch.tensor((feats - mu) / sd)).detach().numpy()[:, 0]
    got = forward_bf16(feats, p)
    assert np.max(np.abs(got - ref)) < 0.15
    assert np.median(np.abs(got - ref)) < 0.03


@pytest.mark.gpu
def test_gpu_inference_matches_restatement():
    from ahmed_amd import _lib
    from ahmed_amd.rank_mlp import load_params, predict_hints, NF
    p = load_params()
    _, _, _, _, _, _, mu, sd = unpack(p)
    rng = np.random.default_rng(1)
    feats = (mu + sd * rng.standard_normal((3000, NF))).astype(np.float32)
    ctx = _lib.Context(0)
    maxk = np.full(3000, 60, np.uint32)
    hint, pred = predict_hints(ctx, feats, p, maxk=maxk, pad=1.25, margin=4)
    ref = forward_bf16(feats, p)
    assert np.max(np.abs(pred - ref)) < 1e-3
    want = np.clip(np.ceil(1.25 * np.maximum(np.exp(ref) - 1, 0)) + 4, 1, 60)
    assert np.mean(hint == want) > 0.99


@pytest.mark.gpu
def test_learned_slabs_do_not_change_the_matrix():
    from ahmed_amd import _lib, geometry
    from ahmed_amd.rank_mlp import leaf_features, load_params, predict_hints
    N, eps, kappa = 20000, 1e-6, 2.0
    pts = geometry.cube_surface(N, 5)
    T = _lib.Tree(pts, 64, 0.64)
    ctx = _lib.Context(0)
    D0, s0 = _lib.HMatrix.assemble(ctx, T, _lib.KERNEL_EXPR, kappa, 6.0 / N, eps, keep_leaves=True,
                                   aca_ranks=True)
    feats, adm, maxk = leaf_features(pts, T, eps, kappa)
    hint_adm, _ = predict_hints(ctx, feats, load_params(), maxk=maxk)
    hint = np.full(T.n_leaves, 48, np.uint32)
    hint[adm] = hint_adm
    D1, s1 = _lib.HMatrix.assemble(ctx, T, _lib.KERNEL_EXPR, kappa, 6.0 / N, eps, rank_hint=hint,
                                   keep_leaves=True, aca_ranks=True)
    l0, d0 = D0.download()
    l1, d1 = D1.download()
    assert np.array_equal(l0["rank"], l1["rank"])
    assert np.array_equal(s0["aca_rank"], s1["aca_rank"])
    o0 = np.argsort(l0["offset"])
    # same values leaf by leaf (layouts may place leaves differently)
    for a in range(0, T.n_leaves, 97):
        n1, n2, k = int(l0["n1"][a]), int(l0["n2"][a]), int(l0["rank"][a])
        nv = n1 * n2 if k < 0 else k * (n1 + n2)
        assert np.array_equal(d0[int(l0["offset"][a]):int(l0["offset"][a]) + nv],
                              d1[int(l1["offset"][a]):int(l1["offset"][a]) + nv])
    under = np.mean(hint_adm < s0["aca_rank"][adm])
    print(f"[gpu] learned slabs: under-predicted {under:.3%}, retries {s1['n_retries']}, "
          f"slab bytes {s1['slab_bytes'] / 1e6:.1f} MB vs default {s0['slab_bytes'] / 1e6:.1f} MB")
    assert s1["slab_bytes"] < s0["slab_bytes"]
