#!/usr/bin/env python3
"""H-matrix hot-path benchmark on one MI355X.

Default workload (the N = 1M mixed-precision configuration, which fits one
GPU): N = 1048576 points uniform on the surface of the unit cube (face chosen
uniformly, uniform on the face; seed 0), bounding-box cluster tree with leaf
size 64, eta = 0.8, ACA eps = 1e-5, Laplace kernel 1/(4 pi r) with diagonal
1.0 and quadrature weight 6/N.  ``--config cfg2`` selects the one-GPU FP64
configuration instead (N = 262144 on the unit sphere, ACA eps 1e-8, GMRES to
1e-10, weight 4 pi/N).

The tree is built by the C++ surface, the H-matrix is assembled on the GPU
(batched ACA + recompression).  A "step" is one FP64 H-matvec y += H x with
one right-hand side; ``value`` is its algorithmic HBM throughput

    B_alg = 8 * sum_leaves nvals + 8 * (N_cols + 2 N_rows)   bytes per step

(leaf values read once, x read once, y read and written once).  The same
line carries the 32-RHS product, FP32 storage, GMRES(50) time-to-tolerance
in plain FP64 and with FP64 iterative refinement around FP32 GMRES, the
roofline of the dominant kernel (row-panel phase) and the reference's own
CPU mltaGeHVec timed on this host on the identical H-matrix.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--config cfg3|cfg2|cfg4|cfg5|trih] [--no-cpu]
"""
from __future__ import annotations

import argparse
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "H-matvec HBM GB/s; GMRES time-to-1e-8 mixed vs FP64 at N=1M, 1 GPU"
HBM_PEAK_GBS = 8000.0
FP64_PEAK_TFS = 78.6
FP32_PEAK_TFS = 157.3


def log(msg: str) -> None:
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def timed(fn, torch, steps: int, warmup: int):
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(steps):
        fn()
    e1.record()
    torch.cuda.synchronize()
    return e0.elapsed_time(e1) / steps  # ms per step


def host_cpu_info() -> dict:
    """CPUs of this host: online count, the ones this process may run on, the
    share the baseline uses, and the model name from /proc/cpuinfo."""
    model = None
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    affinity = len(os.sched_getaffinity(0))
    # one GPU's share of the host: the box runs at most 16 worker threads per
    # GPU (OMP_NUM_THREADS is set to that there)
    share = min(affinity, int(os.environ.get("OMP_NUM_THREADS", "16") or 16), 16)
    return {"online": os.cpu_count(), "affinity": affinity, "used": share, "model": model}


def cpu_baseline(pts, leaves, data, bmin, eta2, B_alg, y_gpu_of, budget_s=20.0, gmres=None, assembly=None):
    """The reference's compiled code (oracle/_ref) on this host's cores, on the
    identical leaves; falls back to the oracle restatement.

    y_gpu_of(x) -> the GPU product of host vector x (compared with the
    reference's product of the same x).  gmres = (b, tol, m): also time one
    GMRes solve of H x = b (x0 = 0), serial and with the leaf-parallel
    operator on all used cores.  assembly = (kernel, kappa, weight, eps,
    stride, gpu_ranks): time the reference's OpenMP leaf loop
    (matgenGeH_omp) on every stride-th leaf on all used cores and compare the
    final ranks of those leaves with the GPU's."""
    import oracle as O
    cpu = host_cpu_info()
    nthr = cpu["used"]
    x = np.random.default_rng(7).standard_normal(len(pts))
    out = {}
    R = None
    try:
        R = O.Reference()
        T = R.tree(pts, bmin, eta2, O.TREE_BBX)
        lv = T.leaves()
        if not (np.array_equal(lv["n1"], leaves["n1"]) and np.array_equal(lv["b2"], leaves["b2"])):
            raise RuntimeError("reference tree differs from the product tree")
        H = R.from_data(T, leaves["rank"], leaves["offset"], data)
        kind = "reference"
    except (FileNotFoundError, OSError, RuntimeError) as e:
        log(f"reference library unavailable ({e}); timing the oracle restatement")
        R = None
        o = O.Oracle()
        H = o.from_leaves(len(pts), leaves["b1"], leaves["n1"], leaves["b2"], leaves["n2"], leaves["rank"],
                          leaves["offset"], data)
        kind = "port"
    H.matvec(x)  # warm-up
    ts = []
    t_start = time.perf_counter()
    while len(ts) < 5 and (time.perf_counter() - t_start) < budget_s:
        t0 = time.perf_counter()
        y = H.matvec(x)
        ts.append(time.perf_counter() - t0)
    t = float(np.median(ts))
    err = float(np.max(np.abs(y_gpu_of(x) - y)) / np.max(np.abs(y)))
    log(f"GPU H-matvec vs the {kind}'s on the same x: rel.err {err:.2e}")
    out.update(value=round(B_alg / t / 1e9, 3), unit="GB/s", cores=1, kind=kind,
               sample=f"{len(ts)} serial mltaGeHVec calls (median {t:.3f} s each) on the identical "
                      f"H-matrix, N={len(pts)}",
               rel_err_vs_gpu=err, host_cpus=cpu)
    # all used cores: the reference's leaf products, leaf-parallel with a
    # private y per thread (its own mltaGeHVec_omp shares y between threads)
    try:
        if R is not None:
            par = lambda: R.matvec_omp(H, x, nthreads=nthr)   # noqa: E731
        else:
            par = lambda: o.matvec_omp(H, x, nthreads=nthr)   # noqa: E731
        par()
        t0 = time.perf_counter()
        reps = 0
        while reps < 3 and time.perf_counter() - t0 < budget_s:
            yo = par()
            reps += 1
        to = (time.perf_counter() - t0) / reps
        out["all_cores"] = {"value": round(B_alg / to / 1e9, 3), "unit": "GB/s", "cores": nthr, "kind": kind,
                            "ms": round(to * 1e3, 2),
                            "rel_diff_vs_serial": float(np.max(np.abs(yo - y)) / np.max(np.abs(y)))}
    except Exception as e:  # report, never hide
        out["all_cores"] = {"error": str(e)}
    if gmres is not None:
        b, tol, m = gmres
        t0 = time.perf_counter()
        xs, rc, res, steps = H.gmres(b, eps=tol, m=m, nsteps=2000)
        tg = time.perf_counter() - t0
        out["gmres"] = {"seconds": round(tg, 3), "rc": rc, "residual": res, "iterations": steps, "cores": 1,
                        "kind": kind, "sample": f"one serial GMRes({m}) solve to {tol:g} on the identical H-matrix"}
        log(f"CPU GMRes({m}): {tg:.2f} s, {steps} iterations, residual {res:.2e}")
        if R is not None:
            t0 = time.perf_counter()
            xs2, rc2, res2, steps2 = R.gmres_omp(H, b, eps=tol, m=m, nsteps=2000, nthreads=nthr)
            tg2 = time.perf_counter() - t0
            out["gmres_all_cores"] = {"seconds": round(tg2, 3), "rc": rc2, "iterations": steps2, "cores": nthr,
                                      "kind": kind, "sample": f"the reference's GMRes({m}) to {tol:g} with the "
                                                              "leaf-parallel operator"}
            log(f"CPU GMRes({m}) on {nthr} cores: {tg2:.2f} s, {steps2} iterations")
    del H
    if assembly is not None and R is not None:
        kern, kappa, weight, eps, stride, gpu_ranks = assembly
        HR, sec, sampled = R.assemble_omp(T, kern, kappa, eps, weight=weight, stride=stride, nthreads=nthr)
        kr = HR.layout()["rank"]
        del HR
        n1 = leaves["n1"].astype(np.float64)
        n2 = leaves["n2"].astype(np.float64)
        share = float((n1 * n2)[sampled].sum() / (n1 * n2).sum())
        adm = kr >= 0
        same = float(np.mean(kr[sampled] == gpu_ranks[sampled]))
        out["assembly"] = {"seconds_sample": round(sec, 3), "cores": nthr, "kind": kind, "stride": stride,
                           "leaves_sampled": int(sampled.sum()), "block_area_share": round(share, 4),
                           "extrapolated_s": round(sec / share, 2),
                           "final_ranks_equal_to_gpu": round(same, 4),
                           "sample": f"matgenGeH_omp's OpenMP leaf loop (apprx + ACA + cpyLrM_cmpr) on every "
                                     f"{stride}-th leaf of the leaf list, {nthr} threads; extrapolated by the "
                                     f"sampled share of the block area"}
        log(f"CPU assembly sample: {sec:.2f} s for {int(sampled.sum())} leaves ({share:.1%} of the block area, "
            f"{nthr} threads) -> {sec / share:.1f} s extrapolated; ranks equal to the GPU's on {same:.2%}")
        del adm
    return out


def symmetric_path(ctx, torch, pts, bmin, eta, weight, eps, tol, H, x, ms_geh, steps, X=None):
    """The symmetric storage of the same matrix (the reference's BEM example:
    subdivide_sym, apprx_sym with ACAr and packed Hermitian diagonal leaves,
    mltaHeHVec, CG on HeHMatrix): assembly time, stored values against the
    unsymmetric matrix, the product (forward part plus the transposed part of
    the off-diagonal leaves) and its difference from the unsymmetric product,
    CG to the solver tolerance."""
    from ahmed_amd import _lib
    N = len(pts)
    t0 = time.perf_counter()
    Ts = _lib.Tree(pts, bmin, eta * eta, _lib.TREE_BBX | _lib.TREE_SYM)
    t_tree = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    S, ast = _lib.HMatrix.assemble(ctx, Ts, _lib.KERNEL_LAPLACE, 0.0, weight, eps)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    ss = S.stats()
    y = torch.zeros(N, dtype=torch.float64, device="cuda")
    ms = timed(lambda: S.matvec(x, y), torch, steps, 2)
    y.zero_()
    S.matvec(x, y)
    yg = torch.zeros(N, dtype=torch.float64, device="cuda")
    H.matvec(x, yg)
    diff = float((y - yg).abs().max() / yg.abs().max())
    ms_p = None
    if X is not None:   # several right-hand sides (mltaHeHGeM)
        Yp = torch.zeros_like(X)
        S.matvec(X, Yp)
        ms_p = timed(lambda: S.matvec(X, Yp), torch, max(3, steps // 4), 1)
        del Yp
    b = torch.zeros(N, dtype=torch.float64, device="cuda")
    S.matvec(torch.ones(N, dtype=torch.float64, device="cuda"), b)
    xs = torch.zeros(N, dtype=torch.float64, device="cuda")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r = S.cg(b, xs, eps=tol, nsteps=5000)
    torch.cuda.synchronize()
    t_cg = time.perf_counter() - t0
    # stored values: the arena holds the diagonal leaves whole (n^2 values)
    lv = Ts.leaves()
    dg = lv["b1"] == lv["b2"]
    n = lv["n1"][dg].astype(np.float64)
    stored = ss["nvals"] - float((n * n - n * (n + 1) / 2).sum())
    out = {"workload": f"the same points, symmetric block tree (subdivide_sym), ACAr, eps {eps}",
           "tree_s": round(t_tree, 3), "gpu_assembly_s": round(t_asm, 3), "leaves": ss["n_leaves"],
           "stored_values": stored, "stored_vs_unsymmetric": round(stored / H.stats()["nvals"], 4),
           "matvec_ms": round(ms, 4), "matvec_ms_unsymmetric": round(ms_geh, 4),
           "rel_diff_vs_unsymmetric_product": diff,
           **({f"matvec_rhs{X.shape[0]}_ms": round(ms_p, 4)} if ms_p is not None else {}),
           "cg": {"seconds": round(t_cg, 4), "converged": r["converged"], "iterations": r["nsteps"],
                  "residual": r["eps"], "max_err_vs_ones": float((xs - 1).abs().max())}}
    log(f"symmetric storage: assembly {t_asm:.2f} s, {stored:.4g} values ({out['stored_vs_unsymmetric']:.3f} of "
        f"the unsymmetric), product {ms:.3f} ms (unsymmetric {ms_geh:.3f} ms), rel.diff {diff:.2e}; "
        f"CG {t_cg:.3f} s, {r['nsteps']} steps")
    del S
    return out


def small_system_gmres(ctx, torch):
    """GMRES(30) to 1e-8 on the reference's own example (unit sphere N=8192,
    leaf 32, eta 0.8, ACA eps 1e-6, Laplace with quadrature weight 4 pi/N,
    b = A*1), where a step's launches, not its product, set the step time.
    Modes: one-by-one launches (the default) and captured steps replayed as
    hipGraphs, each with the orthogonalisation and rotations in one
    workgroup per system (the automatic choice at this size) and in chunked
    launches.  Each mode: one untimed solve, then the median of 5 solves."""
    from ahmed_amd import _lib, geometry
    N = 8192
    tree = _lib.Tree(geometry.sphere(N, 1), 32, 0.8)
    H, _ = _lib.HMatrix.assemble(ctx, tree, _lib.KERNEL_LAPLACE, 0.0, 4 * np.pi / N, 1e-6)
    b = torch.zeros(N, dtype=torch.float64, device="cuda")
    H.matvec(torch.ones(N, dtype=torch.float64, device="cuda"), b)
    out = {"workload": "unit sphere N=8192, leaf 32, eta 0.8, ACA eps 1e-6, Laplace 4 pi/N, GMRES(30) to 1e-8"}
    xs_ref = None
    for name, graph, mode in (("eager", False, 0), ("graph", True, 0), ("eager_chunked", False, 1),
                              ("graph_chunked", True, 1)):
        H.set_graphs(graph)
        H.set_arnoldi(mode)
        dts = []
        for rep in range(6):
            xs = torch.zeros(N, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = H.gmres(b, xs, eps=1e-8, m=30, nsteps=300)
            torch.cuda.synchronize()
            dts.append(time.perf_counter() - t0)
        if xs_ref is None:
            xs_ref = xs
        t = float(np.median(dts[1:]))
        out[name] = {"seconds": round(t, 5), "iterations": r["nsteps"], "converged": r["converged"],
                     "us_per_step": round(t / max(r["nsteps"], 1) * 1e6, 1),
                     "x_equal_to_first_mode": bool(torch.equal(xs, xs_ref))}
        log(f"GMRES(30) N=8192 {name}: {t * 1e3:.2f} ms, {r['nsteps']} steps")
    H.set_arnoldi(0)
    H.set_graphs(False)
    out["graph_speedup_vs_eager"] = round(out["eager"]["seconds"] / out["graph"]["seconds"], 3)
    out["one_workgroup_speedup_vs_chunked"] = round(out["eager_chunked"]["seconds"] / out["eager"]["seconds"], 3)
    del H
    return out


def bench_batched(args):
    """64 independent H-matrices, N log-uniform in [4096, 65536] on seeded
    ellipsoids with axis ratios U(1,4), kernel exp(-k r)/(4 pi r) with
    k ~ U(0,5), ACA eps 1e-6, 8 RHS each: one batched assembly, every
    product in the form the autotuner picks (three launches or one), batched
    GMRES(50) with hipGraph-captured steps."""
    import torch
    from ahmed_amd import _lib, geometry
    probs = geometry.uneven_batch()
    n_mat = len(probs)
    trees = [_lib.Tree(pts, 64, 0.64) for pts, _, _ in probs]
    kernels = [(_lib.KERNEL_YUKAWA, kappa, w) for _, kappa, w in probs]
    sizes = [len(pts) for pts, _, _ in probs]
    ctx = _lib.Context(0)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    H, ast = _lib.HMatrix.assemble_batch(ctx, trees, kernels, eps=1e-6)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    st = H.stats()
    Ntot = H.n_rows
    log(f"batched assembly of {n_mat} matrices (N total {Ntot}): {t_asm:.2f} s, nvals {st['nvals']:.4g}")
    p = 8
    X = torch.randn(p, Ntot, dtype=torch.float64, device="cuda")
    Y = torch.zeros(p, Ntot, dtype=torch.float64, device="cuda")
    # the default form (three launches, 48K-value tiles), then the autotuned one
    ms_default = timed(lambda: H.matvec(X, Y), torch, args.steps, args.warmup)
    t0 = time.perf_counter()
    tile, ms_tuned, fused = H.autotune(p=p, reps=3)
    t_tune = time.perf_counter() - t0
    form_names = {0: "three_launches", 1: "row_panel_single_launch", 2: "persistent_single_launch"}
    log(f"autotuner ({t_tune:.1f} s): tile {tile} values, {form_names[fused]}, "
        f"{ms_tuned:.3f} ms (default {ms_default:.3f} ms)")
    # every form at the tuned tile size, then the tuned form for the line
    forms = {}
    for f in (0, 1, 2):
        H.set_fused(f)
        forms[form_names[f] + "_ms"] = round(timed(lambda: H.matvec(X, Y), torch, args.steps, args.warmup), 4)
    H.set_fused(fused)
    log("8 RHS: " + ", ".join(f"{k} {v:.3f}" for k, v in forms.items()))
    ms = timed(lambda: H.matvec(X, Y), torch, args.steps, args.warmup)
    B_alg = 8.0 * st["nvals"] + 8.0 * p * 3 * Ntot
    F_alg = 2.0 * p * st["nvals"]
    x1 = torch.randn(Ntot, dtype=torch.float64, device="cuda")
    y1 = torch.zeros(Ntot, dtype=torch.float64, device="cuda")
    ms1 = timed(lambda: H.matvec(x1, y1), torch, args.steps, args.warmup)
    log(f"batched H-matvec: 8 RHS {ms:.3f} ms ({B_alg / ms / 1e6:.1f} GB/s, {F_alg / ms / 1e9:.2f} TF/s); "
        f"1 RHS {ms1:.3f} ms ({(8.0 * st['nvals'] + 24.0 * Ntot) / ms1 / 1e6:.1f} GB/s)")
    b = torch.zeros(Ntot, dtype=torch.float64, device="cuda")
    H.matvec(torch.ones(Ntot, dtype=torch.float64, device="cuda"), b)
    # each mode twice: the first graph call captures the Arnoldi steps (kept
    # with the handle), the second replays them
    res = {}
    for name, graph in (("eager", False), ("graph", True)):
        dts = []
        for rep in range(2):
            xs = torch.zeros(Ntot, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            r = H.gmres_batch(b, xs, eps=1e-8, m=50, nsteps=2000, use_graph=graph)
            torch.cuda.synchronize()
            dts.append(time.perf_counter() - t0)
        res[name] = dict(seconds=round(dts[1], 4), first_call_s=round(dts[0], 4),
                         all_converged=bool(r["converged"].all()), max_steps=int(r["nsteps"].max()),
                         mean_steps=float(r["nsteps"].mean()), max_err_vs_ones=float((xs - 1).abs().max()))
        log(f"batched GMRES(50) {name}: {dts[1]:.4f} s "
            f"(first call {dts[0]:.4f} s), steps max {r['nsteps'].max()} mean {r['nsteps'].mean():.2f}")
    res["graph_speedup_vs_eager"] = round(res["eager"]["seconds"] / res["graph"]["seconds"], 3)
    # converged systems are frozen on the device: their product tasks stop,
    # so the products cost about mean/max of a lockstep solve's
    res["lockstep_products_avoided"] = round(1.0 - res["eager"]["mean_steps"] / res["eager"]["max_steps"], 3)
    line = {
        "metric": "batched H-matvec HBM GB/s (64 independent H-matrices, 8 RHS)",
        "value": round(B_alg / ms / 1e6, 2), "unit": "GB/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": round(ms, 4), "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f64", "data": "synthetic (seeded ellipsoids)",
        "config": {"workload": "64 H-matrices, N log-uniform in [4096, 65536], ellipsoids axis ratios U(1,4), "
                               "exp(-k r)/(4 pi r) with k ~ U(0,5), ACA eps 1e-6, 8 RHS, leaf 64, eta 0.8",
                   "config": "cfg4", "N_total": Ntot, "sizes_min_max": [min(sizes), max(sizes)],
                   "nvals": st["nvals"]},
        "TF/s": round(F_alg / ms / 1e9, 3), "matvec_1rhs_ms": round(ms1, 4),
        "autotune": {"tile_values": tile, "launch_form": form_names[fused], "seconds": round(t_tune, 2),
                     "default_form_ms": round(ms_default, 4), **forms},
        "fused_wait_aborted": H.fused_status(),
        "assembly_s": round(t_asm, 3), "aca_retries": ast["n_retries"], "gmres_batch": res,
    }
    print(json.dumps(line), flush=True)


def bench_learned_slabs(args):
    """Batched GPU ACA on N = 1M unit-cube points, kernel exp(-2 r)/r, eps 1e-6,
    with rank slabs sized three ways: the reference's worst case
    (n1*n2/(n1+n2)+4, capped at 256), a fixed 48 columns, and the MLP
    prediction.  ACA continues past under-sized slabs, so the three final
    matrices are identical; time, reruns and slab memory differ."""
    import torch
    from ahmed_amd import _lib, geometry
    from ahmed_amd.rank_mlp import leaf_features, load_params, predict_hints
    N, eps, kappa = args.n or 1048576, 1e-6, 2.0
    pts = geometry.cube_surface(N, 0)
    tree = _lib.Tree(pts, 64, 0.64)
    ctx = _lib.Context(0)
    t0 = time.perf_counter()
    feats, adm, maxk = leaf_features(pts, tree, eps, kappa)
    t_feat = time.perf_counter() - t0
    params = load_params()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    hint_adm, _ = predict_hints(ctx, feats, params, maxk=maxk)
    t_inf = time.perf_counter() - t0
    log(f"features for {len(adm)} admissible leaves {t_feat:.2f} s (host), MLP inference {t_inf * 1e3:.1f} ms")
    hints = {"worst_case": np.minimum(maxk, 256), "fixed48": None, "learned": hint_adm}
    out, ranks = {}, {}
    for name, h in hints.items():
        full = None
        if h is not None:
            full = np.full(tree.n_leaves, 48, np.uint32)
            full[adm] = h
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        H, st = _lib.HMatrix.assemble(ctx, tree, _lib.KERNEL_EXPR, kappa, 6.0 / N, eps, rank_hint=full,
                                      aca_ranks=True)
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        ranks[name] = (st["aca_rank"], H.stats()["nvals"], st["sum_rank"])
        out[name] = dict(seconds=round(dt, 3), aca_batches_s=round(st["aca_ms"] / 1e3, 3),
                         aca_reruns=st["n_retries"], slab_GB=round(st["slab_bytes"] / 1e9, 3),
                         peak_batch_slab_GB=round(st["peak_slab_bytes"] / 1e9, 3), sum_rank=st["sum_rank"],
                         nvals=H.stats()["nvals"])
        log(f"{name}: {dt:.3f} s (ACA batches {st['aca_ms'] / 1e3:.3f} s), reruns {st['n_retries']}, "
            f"slab {st['slab_bytes'] / 1e9:.2f} GB")
        H.close()
    same = all(np.array_equal(ranks[k][0], ranks["fixed48"][0]) and ranks[k][1] == ranks["fixed48"][1]
               for k in ranks)
    actual = ranks["fixed48"][0][adm]
    pred_q = dict(under_predicted=float(np.mean(hint_adm < actual)),
                  mean_slab_over_rank=float(np.mean(hint_adm / np.maximum(actual, 1))))
    line = {
        "metric": "GPU ACA assembly seconds with learned rank slabs", "value": out["learned"]["seconds"],
        "unit": "s", "n_gpus": 1, "steps": 1, "warmup": 0, "ms_per_step": out["learned"]["seconds"] * 1e3,
        "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic (seeded points on the unit-cube surface)",
        "config": {"workload": f"GPU ACA + recompression, N={N} unit-cube surface, exp(-2r)/r, eps 1e-6, leaf 64, "
                               "eta 0.8; MLP 8-64-64-1 BF16 rank slabs", "config": "cfg5", "N": N,
                   "admissible_leaves": int(len(adm))},
        "variants": out, "identical_matrices": bool(same), "prediction": pred_q,
        "feature_seconds_host": round(t_feat, 3), "inference_ms": round(t_inf * 1e3, 2),
    }
    print(json.dumps(line), flush=True)


def bench_trih(args):
    """Triangular H-solves (LtHGeM_solve / UtHGeM_solve / HLU_solve) on the
    FP64 sphere matrix of --config cfg2 (N = 262144, eps 1e-8): block factors
    of the assembled matrix (ahmed_amd.factors: LU of the dense diagonal
    leaves, the off-diagonal leaves unchanged), one persistent launch per
    solve.  Reported beside the reference's serial HLU_solve on the identical
    factors."""
    import torch
    from ahmed_amd import _lib, geometry
    from ahmed_amd.factors import block_factors
    N = args.n or 262144
    pts = geometry.sphere(N, 0)
    weight, eps, bmin, eta = 4 * np.pi / N, 1e-8, 64, 0.8
    ctx = _lib.Context(0)
    tree = _lib.Tree(pts, bmin, eta * eta)
    H, _ = _lib.HMatrix.assemble(ctx, tree, _lib.KERNEL_LAPLACE, 0.0, weight, eps, keep_leaves=True)
    leaves, data = H.download()
    t0 = time.perf_counter()
    layout = {f: leaves[f] for f in ("b1", "n1", "b2", "n2", "rank", "offset")}
    layout["data"] = data
    Lf, Uf = block_factors(layout)
    log(f"block factors: {time.perf_counter() - t0:.1f} s, L {len(Lf['data']) * 8 / 1e9:.2f} GB, "
        f"U {len(Uf['data']) * 8 / 1e9:.2f} GB")
    del layout
    res = {}
    fac = {}
    b = torch.randn(N, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(3))
    for w, F in (("L", Lf), ("U", Uf)):
        t0 = time.perf_counter()
        T = _lib.TriFactor.from_layout(ctx, N, w, F)
        t_up = time.perf_counter() - t0
        fac[w] = T
        info = T.info()
        x = b.clone()
        ms = timed(lambda: (x.copy_(b), T.solve(x)), torch, args.steps, args.warmup)
        B = 8.0 * info["nvals"] + 8.0 * 2 * N
        res[w] = {"ms": round(ms, 3), "GB/s_alg": round(B / (ms * 1e-3) / 1e9, 1),
                  "frac_hbm_peak": round(B / (ms * 1e-3) / 1e9 / HBM_PEAK_GBS, 4),
                  "diagonal_leaves": info["n_diag"], "queue_items": info["n_items"],
                  "nvals": info["nvals"], "upload_s": round(t_up, 2),
                  "us_per_diagonal_leaf": round(ms * 1e3 / info["n_diag"], 2)}
        log(f"{w} solve: {ms:.3f} ms, {res[w]['GB/s_alg']} GB/s algorithmic, "
            f"{info['n_diag']} diagonal leaves ({res[w]['us_per_diagonal_leaf']} us each), "
            f"{info['n_items']} items")
        # 8 right-hand sides in one launch (independent chains sharing the factor)
        Bm = torch.randn(8, N, dtype=torch.float64, device="cuda")
        Xm = Bm.clone()
        ms8 = timed(lambda: (Xm.copy_(Bm), T.solve(Xm)), torch, max(3, args.steps // 2), 1)
        B8 = 8.0 * info["nvals"] + 8.0 * 2 * N * 8
        res[w]["rhs8"] = {"ms": round(ms8, 3), "ms_per_rhs": round(ms8 / 8, 3),
                          "GB/s_alg": round(B8 / (ms8 * 1e-3) / 1e9, 1),
                          "frac_hbm_peak": round(B8 / (ms8 * 1e-3) / 1e9 / HBM_PEAK_GBS, 4)}
        log(f"{w} solve, 8 RHS in one launch: {ms8:.3f} ms ({ms8 / 8:.3f} ms per RHS)")
    x = b.clone()
    ms_lu = timed(lambda: (x.copy_(b), fac["L"].solve(x), fac["U"].solve(x)), torch, args.steps, args.warmup)
    res["HLU_solve_ms"] = round(ms_lu, 3)
    # GMRES with and without the preconditioner (b = H * 1, tol 1e-10)
    bb = torch.zeros(N, dtype=torch.float64, device="cuda")
    H.matvec(torch.ones(N, dtype=torch.float64, device="cuda"), bb)
    gm = {}
    for mode in ("plain", "hlu_solve"):
        xs = torch.zeros(N, dtype=torch.float64, device="cuda")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        r = H.gmres(bb, xs, eps=1e-10, m=50, nsteps=2000, prec=None if mode == "plain" else (fac["L"], fac["U"]))
        torch.cuda.synchronize()
        gm[mode] = {"seconds": round(time.perf_counter() - t0, 4), "iterations": r["nsteps"],
                    "converged": r["converged"], "residual": r["eps"]}
        log(f"GMRES(50) {mode}: {gm[mode]}")
    res["gmres_to_1e-10"] = gm
    cpu = None
    if not args.no_cpu:
        import oracle as O
        bh = b.cpu().numpy()
        try:
            R = O.Reference()
            T = R.tree(pts, bmin, eta * eta, O.TREE_BBX)
            F = R.factors_from_data(T, Lf, Uf)
            kind = "reference"
            run = lambda: F.solve("LU", bh)   # noqa: E731
        except (FileNotFoundError, OSError, RuntimeError) as e:
            log(f"reference library unavailable ({e}); timing the oracle restatement")
            o = O.Oracle()
            kind = "port"
            run = lambda: o.trih_solve(Uf, "U", o.trih_solve(Lf, "L", bh))   # noqa: E731
        ts = []
        t_start = time.perf_counter()
        while len(ts) < 3 and time.perf_counter() - t_start < 30:
            t0 = time.perf_counter()
            xc = run()
            ts.append(time.perf_counter() - t0)
        tc = float(np.median(ts))
        x = b.clone()
        fac["L"].solve(x)
        fac["U"].solve(x)
        diff = float(np.max(np.abs(x.cpu().numpy() - xc)) / np.max(np.abs(xc)))
        cpu = {"value": round(tc * 1e3, 1), "unit": "ms per HLU_solve", "cores": 1, "kind": kind,
               "sample": f"{len(ts)} serial HLU_solve calls on the identical factors, N={N}",
               "gpu_speedup": round(tc * 1e3 / ms_lu, 1), "rel_diff_gpu_vs_cpu": diff}
        log(f"CPU HLU_solve: {tc:.3f} s ({kind}); GPU {ms_lu:.3f} ms; rel.diff {diff:.2e}")
    B_lu = 8.0 * (fac["L"].info()["nvals"] + fac["U"].info()["nvals"]) + 8.0 * 4 * N
    line = {
        "metric": "triangular H-solve (HLU_solve) ms, N=262144 sphere, 1 GPU",
        "value": round(ms_lu, 3), "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": round(ms_lu, 3), "higher_is_better": False, "scaling": "weak", "vs_baseline": None,
        "dtype": "f64", "data": "synthetic (seeded sphere points; H-matrix assembled on the GPU; block factors)",
        "config": {"workload": f"HLU_solve (L then U triangular H-solve), N={N} unit sphere, leaf size {bmin}, "
                               f"eta {eta}, ACA eps {eps}, factors = LU of the diagonal leaves + off-diagonal leaves",
                   "N": N},
        "roofline": {"bound": "latency (diagonal chain) / hbm", "achieved": round(B_lu / (ms_lu * 1e-3) / 1e9, 1),
                     "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(B_lu / (ms_lu * 1e-3) / 1e9 / HBM_PEAK_GBS, 4), "traffic": None},
        "cpu_baseline": cpu,
        "solves": res,
    }
    print(json.dumps(line), flush=True)


def profiled_traffic(config: str, N: int):
    """HBM bytes per row-panel launch from the committed FETCH_SIZE / WRITE_SIZE
    summary of the same workload (None for other workloads)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hmatvec_traffic_cube1M.md")
    if config != "cfg3" or N != 1048576 or not os.path.exists(path):
        return None
    m = re.search(r"^traffic_bytes_per_launch\[cfg3\] = (\d+)", open(path).read(), re.M)
    return float(m.group(1)) if m else None


def stream_read_gbs(torch, gib: int = 4, reps: int = 5) -> float:
    """Achievable streaming-read rate of this GPU (torch sum over gib GiB of doubles)."""
    a = torch.ones((gib << 30) // 8, dtype=torch.float64, device="cuda")
    a.sum()
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        a.sum()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / reps
    del a
    return (gib << 30) / (ms * 1e-3) / 1e9


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--config", choices=["cfg3", "cfg2", "cfg4", "cfg5", "trih"], default="cfg3")
    ap.add_argument("--n", type=int, default=0, help="override N")
    ap.add_argument("--rhs", type=int, default=32)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--slow-alpha", type=float, default=100.0,
                    help="also solve with the off-diagonal weight times this (0: skip)")
    ap.add_argument("--copy-probe", action="store_true",
                    help="end with a plain 2 GiB read+write kernel (scale for FETCH_SIZE/WRITE_SIZE counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU only (asked for --gpus {args.gpus}); the multi-GPU "
              "sharding of the H-matvec (row-cluster intervals, all-gather of x) is designed, not built.", file=sys.stderr)
        sys.exit(2)

    import torch
    from ahmed_amd import _lib, geometry

    assert torch.cuda.is_available(), "bench.py needs a GPU"
    if args.config == "cfg4":
        return bench_batched(args)
    if args.config == "cfg5":
        return bench_learned_slabs(args)
    if args.config == "trih":
        return bench_trih(args)
    if args.config == "cfg2":
        N, eps, tol, geo = args.n or 262144, 1e-8, 1e-10, "unit sphere"
        pts = geometry.sphere(N, 0)
        weight = 4 * np.pi / N
    else:
        N, eps, tol, geo = args.n or 1048576, 1e-5, 1e-8, "unit-cube surface"
        pts = geometry.cube_surface(N, 0)
        weight = 6.0 / N
    bmin, eta = 64, 0.8
    ctx = _lib.Context(0)
    t0 = time.perf_counter()
    tree = _lib.Tree(pts, bmin, eta * eta)
    t_tree = time.perf_counter() - t0
    log(f"tree: {tree.n_leaves} leaves in {t_tree:.2f} s")
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    H, ast = _lib.HMatrix.assemble(ctx, tree, _lib.KERNEL_LAPLACE, 0.0, weight, eps,
                                   keep_leaves=not args.no_cpu, aca_ranks=True)
    torch.cuda.synchronize()
    t_asm = time.perf_counter() - t0
    st = H.stats()
    log(f"GPU assembly {t_asm:.2f} s: {st['n_lowrank']} low-rank / {st['n_dense']} dense, "
        f"nvals {st['nvals']:.4g}, ACA retries {ast['n_retries']}")
    # assembly work (VALU-bound): kernel entries evaluated by ACA (one row and
    # one column per step) and by the dense leaves, and the ACA updates
    # (2 k^2 (n1+n2) flops per admissible leaf of k steps)
    lv = tree.leaves()
    adm = lv["adm"].astype(bool)
    nsum = (lv["n1"].astype(np.float64) + lv["n2"].astype(np.float64))[adm]
    kaca = ast["aca_rank"].astype(np.float64)[adm]
    asm_work = {"kernel_evals_aca": float((kaca * nsum).sum()), "kernel_evals_dense": float(st["nvals_dense"]),
                "update_flops": float((2.0 * kaca * kaca * nsum).sum())}
    aca_gflops = asm_work["update_flops"] / (ast["aca_ms"] * 1e-3) / 1e9
    asm_work.update({"aca_phase_update_GFLOP/s": round(aca_gflops, 2),
                     "aca_phase_evals_per_s": round(asm_work["kernel_evals_aca"] / (ast["aca_ms"] * 1e-3), 1),
                     "frac_fp64_valu_peak": round(aca_gflops / 78600.0, 5)})
    log(f"assembly work: {asm_work['kernel_evals_aca']:.3g} ACA + {asm_work['kernel_evals_dense']:.3g} dense kernel "
        f"entries, {asm_work['update_flops']:.3g} update flops; ACA phase {aca_gflops:.1f} GFLOP/s "
        f"({100 * aca_gflops / 78600.0:.2f} % of the FP64 vector peak)")
    # the same assembly with ACAr (apprx_sym's row-pivoted ACA) on the admissible leaves
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Hr, astr = _lib.HMatrix.assemble(ctx, tree, _lib.KERNEL_LAPLACE, 0.0, weight, eps, rows_only=True)
    torch.cuda.synchronize()
    t_acar = time.perf_counter() - t0
    str_ = Hr.stats()
    acar = {"gpu_assembly_s": round(t_acar, 3), "aca_ms": round(astr["aca_ms"], 1),
            "mean_aca_rank": astr["sum_aca_rank"] / max(1, astr["n_candidates"]),
            "mean_rank": astr["sum_rank"] / max(1, astr["n_lowrank"]), "nvals": str_["nvals"],
            "n_lowrank": str_["n_lowrank"]}
    del Hr
    # the assemblies' scratch and the ACAr matrix sit in the allocation cache:
    # hand that memory back before the product and solver phases
    ctx.release_cached()
    log(f"GPU assembly with ACAr {t_acar:.2f} s (ACA batches {astr['aca_ms']:.0f} ms vs {ast['aca_ms']:.0f} ms), "
        f"nvals {str_['nvals']:.4g}")

    # ---- 1 RHS, FP64 ---------------------------------------------------------
    x = torch.randn(N, dtype=torch.float64, device="cuda", generator=torch.Generator("cuda").manual_seed(1))
    y = torch.zeros(N, dtype=torch.float64, device="cuda")
    B_alg = 8.0 * st["nvals"] + 8.0 * (N + 2 * N)
    ms = timed(lambda: H.matvec(x, y), torch, args.steps, args.warmup)
    gbs = B_alg / (ms * 1e-3) / 1e9
    log(f"FP64 H-matvec 1 RHS: {ms:.3f} ms, {gbs:.1f} GB/s")
    ms_a, ms_b = H.profile(x, y, reps=args.steps)
    bytes_b = 8.0 * (st["nvals_dense"] + st["nvals_u"]) + 8.0 * (N + st["t_len"] + 2 * N)
    bytes_a = 8.0 * st["nvals_v"] + 8.0 * (N + st["t_len"])
    ach_b = bytes_b / (ms_b * 1e-3) / 1e9
    log(f"phase A (V^T x) {ms_a:.3f} ms {bytes_a / (ms_a * 1e-3) / 1e9:.1f} GB/s; "
        f"phase B (row panels) {ms_b:.3f} ms {ach_b:.1f} GB/s")

    # ---- transposed product y += H^T x (mltaGeHhVec), FP64 ---------------------
    yt = torch.zeros(N, dtype=torch.float64, device="cuda")
    H.matvec_t(x, yt)          # builds the column index once
    mst = timed(lambda: H.matvec_t(x, yt), torch, args.steps, args.warmup)
    log(f"FP64 transposed H-matvec 1 RHS: {mst:.3f} ms, {B_alg / (mst * 1e-3) / 1e9:.1f} GB/s")

    # ---- p RHS, FP64 ---------------------------------------------------------
    p = args.rhs
    X = torch.randn(p, N, dtype=torch.float64, device="cuda")
    Y = torch.zeros(p, N, dtype=torch.float64, device="cuda")
    msp = timed(lambda: H.matvec(X, Y), torch, max(3, args.steps // 4), 1)
    F_alg = 2.0 * p * st["nvals"]
    Bp = 8.0 * st["nvals"] + 8.0 * p * (N + 2 * N)
    log(f"FP64 H-matvec {p} RHS: {msp:.3f} ms, {F_alg / (msp * 1e-3) / 1e12:.2f} TF/s")
    H.matvec_t(X, Y)           # sizes the 32-wide column-dot slots once
    mstp = timed(lambda: H.matvec_t(X, Y), torch, max(3, args.steps // 4), 1)
    log(f"FP64 transposed H-matvec {p} RHS: {mstp:.3f} ms, {F_alg / (mstp * 1e-3) / 1e12:.2f} TF/s")

    # ---- FP32 storage ----------------------------------------------------------
    H32 = H.convert()
    x32, y32 = x.float(), torch.zeros(N, dtype=torch.float32, device="cuda")
    ms32 = timed(lambda: H32.matvec(x32, y32), torch, args.steps, args.warmup)
    B32 = 4.0 * st["nvals"] + 4.0 * 3 * N
    yref = torch.zeros(N, dtype=torch.float64, device="cuda")
    H.matvec(x, yref)
    y32.zero_()
    H32.matvec(x32, y32)
    err32 = float((y32.double() - yref).abs().max() / yref.abs().max())
    log(f"FP32 H-matvec 1 RHS: {ms32:.3f} ms, {B32 / (ms32 * 1e-3) / 1e9:.1f} GB/s, rel.err vs FP64 {err32:.2e}")

    # ---- p RHS with FP32 storage, and with V^T X on BF16 MFMA ----------------
    Y64 = torch.zeros(p, N, dtype=torch.float64, device="cuda")
    H.matvec(X, Y64)
    X32p = X.float()
    Y32p = torch.zeros(p, N, dtype=torch.float32, device="cuda")
    ms32p = timed(lambda: H32.matvec(X32p, Y32p), torch, max(3, args.steps // 4), 1)
    Y32p.zero_()
    H32.matvec(X32p, Y32p)
    err32p = float((Y32p.double() - Y64).abs().max() / Y64.abs().max())
    Hb = H.convert(_lib.BF16)
    ms_bp = timed(lambda: Hb.matvec(X32p, Y32p), torch, max(3, args.steps // 4), 1)
    Y32p.zero_()
    Hb.matvec(X32p, Y32p)
    errbp = float((Y32p.double() - Y64).abs().max() / Y64.abs().max())
    ms_bf_a, ms_bf_b = Hb.profile(X32p, Y32p, reps=3)
    B_bf = 2.0 * st["nvals_v"] + 4.0 * (st["nvals_dense"] + st["nvals_u"]) + 4.0 * p * (N + 2 * N)
    log(f"{p} RHS: FP32 storage {ms32p:.3f} ms ({F_alg / (ms32p * 1e-3) / 1e12:.2f} TF/s, rel.err {err32p:.2e}); "
        f"BF16 V^T X {ms_bp:.3f} ms ({F_alg / (ms_bp * 1e-3) / 1e12:.2f} TF/s, rel.err {errbp:.2e}; "
        f"phase A {ms_bf_a:.3f} ms, phase B {ms_bf_b:.3f} ms)")
    del Hb, Y64, X32p, Y32p

    # ---- GMRES(50) time-to-tolerance: FP64 vs FP64 refinement around FP32 ----
    b = torch.zeros(N, dtype=torch.float64, device="cuda")
    H.matvec(torch.ones(N, dtype=torch.float64, device="cuda"), b)   # b = H * 1
    # each solve twice, the second timed; "fp64_graph": the first call
    # captures the Arnoldi steps into hipGraphs kept with the handle, the
    # timed second call replays them (the others launch the kernels one by one)
    gm = {}
    for mode in ("fp64", "fp64_graph", "mixed"):
        H.set_graphs(mode == "fp64_graph")
        dts = []
        for rep in range(2):
            xs = torch.zeros(N, dtype=torch.float64, device="cuda")
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            if mode != "mixed":
                r = H.gmres(b, xs, eps=tol, m=50, nsteps=2000)
                info = {"iterations": r["nsteps"], "matvecs": r["nmatvec"]}
            else:
                r = H.gmres_ir(H32, b, xs, eps=tol, m=50, inner_eps=1e-5, inner_steps=2000)
                info = {"refinements": r["outer"], "inner_iterations": r["inner"]}
            torch.cuda.synchronize()
            dts.append(time.perf_counter() - t0)
        dt = dts[1]
        info["first_call_s"] = round(dts[0], 4)
        res = torch.zeros(N, dtype=torch.float64, device="cuda")
        res.copy_(b)
        H.matvec(xs, res, -1.0)
        true = float(res.norm() / b.norm())
        err = float((xs - 1.0).abs().max())
        gm[mode] = dict(seconds=round(dt, 4), converged=r["converged"], reported_residual=r["eps"],
                        true_residual_fp64=true, max_err_vs_ones=err, **info)
        log(f"GMRES(50) {mode}: {dt:.3f} s, residual {true:.2e}, {info}")
    H.set_graphs(False)
    gm["tolerance"] = tol
    gm["speedup_mixed_vs_fp64"] = round(gm["fp64"]["seconds"] / gm["mixed"]["seconds"], 3)
    gm["graph_speedup_vs_eager"] = round(gm["fp64"]["seconds"] / gm["fp64_graph"]["seconds"], 3)
    # the same geometry with a stronger off-diagonal part (weight alpha * 6/N):
    # a system that needs tens of GMRES iterations, where the refinement's
    # FP32 inner products can outweigh its FP64 residual products
    if args.slow_alpha > 0:
        del H32
        ctx.release_cached()
        Hs, _ = _lib.HMatrix.assemble(ctx, tree, _lib.KERNEL_LAPLACE, 0.0, args.slow_alpha * weight, eps)
        Hs32 = Hs.convert()
        bs = torch.zeros(N, dtype=torch.float64, device="cuda")
        Hs.matvec(torch.ones(N, dtype=torch.float64, device="cuda"), bs)
        slow = {"weight": f"{args.slow_alpha:g} x the weight above", "tolerance": tol}
        for mode in ("fp64", "mixed"):
            dts = []
            for rep in range(2):
                xs = torch.zeros(N, dtype=torch.float64, device="cuda")
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                if mode == "fp64":
                    r = Hs.gmres(bs, xs, eps=tol, m=50, nsteps=5000)
                    info = {"iterations": r["nsteps"], "matvecs": r["nmatvec"]}
                else:
                    r = Hs.gmres_ir(Hs32, bs, xs, eps=tol, m=50, inner_eps=1e-5, inner_steps=5000)
                    info = {"refinements": r["outer"], "inner_iterations": r["inner"]}
                torch.cuda.synchronize()
                dts.append(time.perf_counter() - t0)
            dt = dts[1]
            info["first_call_s"] = round(dts[0], 4)
            res = bs.clone()
            Hs.matvec(xs, res, -1.0)
            slow[mode] = dict(seconds=round(dt, 4), converged=r["converged"],
                              true_residual_fp64=float(res.norm() / bs.norm()), **info)
            log(f"GMRES(50) {mode}, weight x{args.slow_alpha:g}: {dt:.3f} s, {info}")
        slow["speedup_mixed_vs_fp64"] = round(slow["fp64"]["seconds"] / slow["mixed"]["seconds"], 3)
        slow["speedup_bound_from_products"] = round(
            slow["fp64"]["matvecs"] * ms / (slow["mixed"]["refinements"] * ms + slow["mixed"]["inner_iterations"] * ms32), 3)
        gm["slower_system"] = slow
        del Hs, Hs32, bs
    gm["small_system"] = small_system_gmres(ctx, torch)
    sym = symmetric_path(ctx, torch, pts, bmin, eta, weight, eps, tol, H, x, ms, args.steps, X)
    ctx.release_cached()
    # what the products alone allow: the FP64 solve's products against the
    # refinement's FP64 residual products plus its FP32 inner products
    n64 = gm["fp64"]["matvecs"]
    nres = gm["mixed"]["refinements"]
    nin = gm["mixed"]["inner_iterations"]
    gm["speedup_bound_from_products"] = round(n64 * ms / (nres * ms + nin * ms32), 3)
    gm["target_speedup"] = 3.0

    sr = stream_read_gbs(torch)
    log(f"streaming read (torch sum over 4 GiB): {sr:.1f} GB/s")
    traffic = profiled_traffic(args.config, N)

    if args.copy_probe:
        a = torch.ones(1 << 28, dtype=torch.float64, device="cuda")
        c = torch.empty_like(a)
        torch.neg(a, out=c)
        torch.cuda.synchronize()
        log("copy probe: neg over 2 GiB (reads 2 GiB, writes 2 GiB)")
        del a, c

    cpu = None
    if not args.no_cpu:
        leaves, data = H.download()
        log(f"CPU baseline on {len(leaves)} leaves ({data.nbytes / 1e9:.1f} GB)")

        def y_gpu_of(xh):
            yd = torch.zeros(N, dtype=torch.float64, device="cuda")
            H.matvec(torch.as_tensor(xh, device="cuda"), yd)
            return yd.cpu().numpy()
        stride = 1 if N <= 262144 else 8
        cpu = cpu_baseline(pts, leaves, data, bmin, eta * eta, B_alg, y_gpu_of, gmres=(b.cpu().numpy(), tol, 50),
                           assembly=(_lib.KERNEL_LAPLACE, 0.0, weight, eps, stride, leaves["rank"]))
        if "gmres" in cpu:
            gm["speedup_mixed_vs_cpu_serial"] = round(cpu["gmres"]["seconds"] / gm["mixed"]["seconds"], 1)
            gm["speedup_fp64_vs_cpu_serial"] = round(cpu["gmres"]["seconds"] / gm["fp64"]["seconds"], 1)
        if "gmres_all_cores" in cpu:
            gm["speedup_mixed_vs_cpu_all_cores"] = round(cpu["gmres_all_cores"]["seconds"] / gm["mixed"]["seconds"], 1)
        log(f"CPU: {cpu}")
        del data

    line = {
        "metric": METRIC,
        "value": round(gbs, 2),
        "unit": "GB/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": f"synthetic (seeded points on the {geo}; H-matrix assembled on the GPU)",
        "config": {"workload": f"FP64 H-matvec, 1 RHS, N={N} {geo}, leaf size {bmin}, eta {eta}, "
                               f"ACA eps {eps}, Laplace kernel (quadrature weight {weight:.4g})",
                   "config": args.config,
                   "N": N, "leaf_size": bmin, "eta": eta, "aca_eps": eps, "rhs": 1,
                   "leaves": st["n_leaves"], "lowrank_leaves": st["n_lowrank"], "dense_leaves": st["n_dense"],
                   "nvals": st["nvals"], "bytes_per_step": B_alg},
        "roofline": {"bound": "hbm", "kernel": "panel_gemv_kernel<double,double,PHASE_B> (row panels)",
                     "achieved": round(ach_b, 2), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(ach_b / HBM_PEAK_GBS, 4), "traffic": traffic,
                     "traffic_over_algorithmic": round(traffic / bytes_b, 3) if traffic else None,
                     "stream_read_GBs": round(sr, 1), "frac_of_stream_read": round(ach_b / sr, 4),
                     "bytes_per_launch": bytes_b, "ms_per_launch": round(ms_b, 4),
                     "phase_a": {"ms": round(ms_a, 4), "GB/s": round(bytes_a / (ms_a * 1e-3) / 1e9, 2)}},
        "cpu_baseline": cpu,
        "parity": None if cpu is None else {
            "rel_err_gpu_vs_reference": cpu["rel_err_vs_gpu"], "bar": 1e-12,
            "ok": bool(cpu["rel_err_vs_gpu"] <= 1e-12), "against": cpu["kind"],
            "what": "GPU FP64 H-matvec vs the reference's mltaGeHVec on the identical H-matrix and x"},
        "matvec_multi_rhs": {"rhs": p, "ms": round(msp, 4),
                             "TF/s": round(F_alg / (msp * 1e-3) / 1e12, 3),
                             "mfma_frac_fp64": round(F_alg / (msp * 1e-3) / 1e12 / FP64_PEAK_TFS, 4),
                             "GB/s_alg": round(Bp / (msp * 1e-3) / 1e9, 2)},
        "matvec_transposed": {"ms": round(mst, 4), "GB/s": round(B_alg / (mst * 1e-3) / 1e9, 2),
                              f"rhs{p}": {"ms": round(mstp, 4), "TF/s": round(F_alg / (mstp * 1e-3) / 1e12, 3)}},
        "fp32_storage": {"ms": round(ms32, 4), "GB/s": round(B32 / (ms32 * 1e-3) / 1e9, 2),
                         "rel_err_vs_fp64": err32,
                         "rhs32": {"ms": round(ms32p, 4), "TF/s": round(F_alg / (ms32p * 1e-3) / 1e12, 3),
                                   "frac_fp32_peak": round(F_alg / (ms32p * 1e-3) / 1e12 / FP32_PEAK_TFS, 4),
                                   "rel_err_vs_fp64": err32p}},
        "bf16_vtx_rhs32": {"what": "U, dense in FP32; V^T X on BF16 MFMA (V, X rounded to BF16), FP32 accumulation",
                           "ms": round(ms_bp, 4), "TF/s": round(F_alg / (ms_bp * 1e-3) / 1e12, 3),
                           "GB/s_alg": round(B_bf / (ms_bp * 1e-3) / 1e9, 1),
                           "phase_a_ms": round(ms_bf_a, 4), "phase_b_ms": round(ms_bf_b, 4),
                           "rel_err_vs_fp64": errbp},
        "assembly": {"tree_s": round(t_tree, 3), "gpu_assembly_s": round(t_asm, 3), "aca_ms": round(ast["aca_ms"], 1),
                     "aca_retries": ast["n_retries"], "mean_aca_rank": ast["sum_aca_rank"] / max(1, ast["n_candidates"]),
                     "mean_rank": ast["sum_rank"] / max(1, ast["n_lowrank"]), "work": asm_work, "acar": acar},
        "gmres": gm,
        "symmetric": sym,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
