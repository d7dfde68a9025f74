"""Learned ACA rank-slab sizes (the 8-64-64-1 MLP of the batched GPU ACA).

Each admissible leaf is described by 8 geometric features; the network
predicts log(1 + ACA rank) and the GPU kernel ``ahm_rank_predict`` (BF16
inference, FP32 accumulation) turns the prediction into a rank slab for the
batched ACA.  ACA keeps going past an under-sized slab (the leaf is rerun
with a larger one), so the prediction only changes memory and time; plain ACA
stays both the fallback and the judge of the final ranks.

Training pairs (geometry, ACA rank) come from GPU assemblies of the three
single-matrix benchmark configurations (unit sphere N=8192 eps 1e-6 leaf 32;
unit sphere N=262144 eps 1e-8; unit-cube surface N=1048576 eps 1e-5, leaf 64;
all Laplace) and of decaying kernels with kappa across [0.5, 5]
(exp(-kappa r)/r on unit-cube surfaces of N=262144 and N=1048576,
exp(-kappa r)/(4 pi r) on a unit sphere of N=131072, eps 1e-6, other point
seeds than the benchmarks), so that the kappa feature varies.  The network is fitted to a
high quantile (0.98) of log(1 + rank) rather than its mean: the slab is
meant to hold the rank, and an under-sized slab costs a rerun of the leaf.

  python -m ahmed_amd.rank_mlp train --out ahmed_amd/rank_mlp_params.npy
"""
from __future__ import annotations

import argparse
import ctypes as C
import os
import time

import numpy as np

from . import _lib, geometry

HERE = os.path.dirname(os.path.abspath(__file__))
PARAMS_PATH = os.path.join(HERE, "rank_mlp_params.npy")
NF, NH = 8, 64
N_PARAMS = NF * NH + NH + NH * NH + NH + NH + 1 + 2 * NF


def leaf_features(pts: np.ndarray, tree, eps: float, kappa: float):
    """(features (n_adm x 8) float32, admissible leaf indices, maxk per leaf)."""
    op, _ = tree.perm()
    P = pts[op]
    lv = tree.leaves()
    adm = np.flatnonzero(lv["adm"])
    b1, n1 = lv["b1"][adm].astype(np.int64), lv["n1"][adm].astype(np.int64)
    b2, n2 = lv["b2"][adm].astype(np.int64), lv["n2"][adm].astype(np.int64)
    # bounding boxes of every distinct cluster range
    keys = np.concatenate([b1 * (1 << 32) + n1, b2 * (1 << 32) + n2])
    uk, inv = np.unique(keys, return_inverse=True)
    lo = np.empty((len(uk), 3))
    hi = np.empty((len(uk), 3))
    for u, k in enumerate(uk):
        b, n = int(k >> 32), int(k & 0xFFFFFFFF)
        seg = P[b:b + n]
        lo[u] = seg.min(0)
        hi[u] = seg.max(0)
    r, c = inv[:len(adm)], inv[len(adm):]
    d1 = np.linalg.norm(hi[r] - lo[r], axis=1)
    d2 = np.linalg.norm(hi[c] - lo[c], axis=1)
    gap = np.maximum(0.0, np.maximum(lo[c] - hi[r], lo[r] - hi[c]))
    dist = np.linalg.norm(gap, axis=1)
    dmax = np.maximum(np.maximum(d1, d2), 1e-300)
    N = len(pts)
    feats = np.stack([
        np.log2(N / n1), np.log(n1), np.log(n2),
        d1 / np.maximum(d2, 1e-300), dist / dmax,
        np.full(len(adm), np.log10(eps)), np.full(len(adm), kappa), np.log10(dmax),
    ], axis=1).astype(np.float32)
    maxk = (n1 * n2 // (n1 + n2) + 4).astype(np.uint32)
    return feats, adm, maxk


# (geometry, N, point seed, leaf size, eps, kernel, kappa, weight per point count)
TRAIN_CONFIGS = [
    ("sphere", 8192, 0, 32, 1e-6, _lib.KERNEL_LAPLACE, 0.0, 1.0),
    ("sphere", 262144, 0, 64, 1e-8, _lib.KERNEL_LAPLACE, 0.0, 1.0),
    ("cube", 1048576, 0, 64, 1e-5, _lib.KERNEL_LAPLACE, 0.0, 1.0),
    ("cube", 1048576, 1, 64, 1e-6, _lib.KERNEL_EXPR, 3.5, 6.0),
] + [("cube", 262144, 1, 64, 1e-6, _lib.KERNEL_EXPR, k, 6.0) for k in (0.5, 1.5, 3.0, 5.0)] \
  + [("sphere", 131072, 2, 64, 1e-6, _lib.KERNEL_YUKAWA, k, 4 * np.pi) for k in (1.0, 4.0)]


def collect(ctx, configs=None):
    """GPU assemblies of the training configurations -> (features, log1p(ACA rank))."""
    configs = TRAIN_CONFIGS if configs is None else configs
    X, Y = [], []
    for geo, N, seed, bmin, eps, kernel, kappa, wn in configs:
        pts = geometry.sphere(N, seed) if geo == "sphere" else geometry.cube_surface(N, seed)
        tree = _lib.Tree(pts, bmin, 0.64)
        H, st = _lib.HMatrix.assemble(_ctx(ctx), tree, kernel, kappa, wn / N, eps, aca_ranks=True)
        H.close()
        f, adm, _ = leaf_features(pts, tree, eps, kappa)
        X.append(f)
        Y.append(np.log1p(st["aca_rank"][adm].astype(np.float32)))
        print(f"collected {geo} N={N} kernel {kernel} kappa {kappa}: {len(adm)} admissible leaves, "
              f"mean ACA rank {st['aca_rank'][adm].mean():.2f}", flush=True)
    return np.concatenate(X), np.concatenate(Y)


def _ctx(ctx):
    return ctx if ctx is not None else _lib.Context(0)


def train(X: np.ndarray, Y: np.ndarray, epochs: int = 6, seed: int = 0, quantile: float = 0.98) -> np.ndarray:
    """FP32 training on the CPU with the pinball loss of the given quantile;
    returns the packed parameter vector."""
    import torch
    torch.manual_seed(seed)
    mu = X.mean(0)
    sd = X.std(0)
    sd[sd < 1e-6] = 1.0
    Xt = torch.tensor((X - mu) / sd, dtype=torch.float32)
    Yt = torch.tensor(Y, dtype=torch.float32)[:, None]
    net = torch.nn.Sequential(torch.nn.Linear(NF, NH), torch.nn.ReLU(), torch.nn.Linear(NH, NH),
                              torch.nn.ReLU(), torch.nn.Linear(NH, 1))
    opt = torch.optim.Adam(net.parameters(), lr=2e-3)
    n = len(Xt)
    for ep in range(epochs):
        perm = torch.randperm(n)
        tot = 0.0
        for i in range(0, n, 4096):
            idx = perm[i:i + 4096]
            opt.zero_grad()
            e = Yt[idx] - net(Xt[idx])
            loss = torch.mean(torch.maximum(quantile * e, (quantile - 1.0) * e))
            loss.backward()
            opt.step()
            tot += float(loss.detach()) * len(idx)
        print(f"epoch {ep}: pinball loss {tot / n:.4f}", flush=True)
    W1 = net[0].weight.detach().numpy().T    # (NF, NH): row = input
    W2 = net[2].weight.detach().numpy().T
    W3 = net[4].weight.detach().numpy()[0]
    return np.concatenate([W1.ravel(), net[0].bias.detach().numpy(), W2.ravel(),
                           net[2].bias.detach().numpy(), W3, net[4].bias.detach().numpy(),
                           mu, sd]).astype(np.float32)


def load_params(path: str = PARAMS_PATH) -> np.ndarray:
    p = np.load(path, allow_pickle=False).astype(np.float32)
    assert p.shape == (N_PARAMS,), p.shape
    return p


def predict_hints(ctx, feats: np.ndarray, params: np.ndarray, maxk=None, pad: float = 1.0,
                  margin: int = 2):
    """GPU inference: (rank-slab hints, predicted log1p(rank))."""
    feats = np.ascontiguousarray(feats, np.float32)
    params = np.ascontiguousarray(params, np.float32)
    n = len(feats)
    hint = np.zeros(n, np.uint32)
    pred = np.zeros(n, np.float32)
    mk = None if maxk is None else np.ascontiguousarray(maxk, np.uint32)
    _ctx(ctx).bind_stream()
    _lib.check(_lib.load().ahm_rank_predict(
        _ctx(ctx).h, feats.ctypes.data_as(C.c_void_p), n, params.ctypes.data_as(C.c_void_p), pad, margin,
        None if mk is None else mk.ctypes.data_as(C.c_void_p), hint.ctypes.data_as(C.c_void_p),
        pred.ctypes.data_as(C.c_void_p)), "ahm_rank_predict")
    return hint, pred


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("cmd", choices=["train"])
    ap.add_argument("--out", default=PARAMS_PATH)
    ap.add_argument("--epochs", type=int, default=8)
    args = ap.parse_args()
    ctx = _lib.Context(0)
    t0 = time.time()
    X, Y = collect(ctx)
    rng = np.random.default_rng(0)
    held = rng.random(len(X)) < 0.1
    params = train(X[~held], Y[~held], epochs=args.epochs)
    # held-out leaves: share under-sized and slab columns per rank column
    rank = np.expm1(Y[held]).round()
    for pad, margin in ((1.0, 1), (1.0, 2), (1.0, 4), (1.25, 4)):
        hint, _ = predict_hints(ctx, X[held], params, pad=pad, margin=margin)
        print(f"held out, pad {pad} margin {margin}: under-sized {np.mean(hint < rank):.3%}, "
              f"slab columns / rank columns {hint.sum() / max(rank.sum(), 1):.2f}", flush=True)
    np.save(args.out, params)
    print(f"saved {len(params)} parameters to {args.out} ({time.time() - t0:.1f} s)")


if __name__ == "__main__":
    main()
