"""Benchmark: 2-D Helmholtz butterfly factor + apply at N = 2^20 on one MI355X.

Workload (the headline configuration): 2^20 sources equispaced in theta on
the curve r = 1 + 0.3 cos(5 theta), 2^20 targets equispaced on the circle
of radius 3, k for 10 points per wavelength along the curve (k = 73064.8),
quadtree leaf size 64, rank-estimate tolerance 1e-8, complex FP64, one
complex Gaussian right-hand side.  The factorization runs once (its wall
time, trees included, is `factor_s`), after an untimed factorization of
2^14 points that takes the process's one-time costs (`factor_warmup`); a
step is one apply of the factorization to the resident right-hand side,
so `value` is applies (RHS) per second.

Also reported: accuracy against exact FP64 dense rows at random targets
(the oracle, on the host), the 256-RHS apply in FP64 and FP32 (the mixed
precision comparison), the roofline of the dominant kernel (the apply,
HBM-bound) from HIP events, and the oracle CPU baseline on a bounded
sample.
"""
import argparse
import json
import os
import sys
import time
from concurrent.futures import ThreadPoolExecutor

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "factor s + apply RHS/s at N=2^20, mixed vs FP64 GPU at rel err ≤1e-6"
CURVE_L = 9.017203500515143
HBM_PEAK_GBS = 8000.0


def curve_problem(n, ppw=10):
    th = 2 * np.pi * np.arange(n) / n
    r = 1 + 0.3 * np.cos(5 * th)
    src = np.stack([r * np.cos(th), r * np.sin(th)], 1)
    tgt = np.stack([3 * np.cos(th), 3 * np.sin(th)], 1)
    return src, tgt, 2 * np.pi * n / (ppw * CURVE_L)


# ---- the CPU leg: the oracle (the C restatement of the reference, test
# infrastructure) computes the exact rows every accuracy figure is checked
# against, and is timed as the CPU baseline (cpu_baseline below).  Nothing
# timed on the GPU runs through it.

def cpu_baseline_exact_rows(k, tgt_rows, src, x, threads):
    """Exact FP64 rows of the single-layer kernel matrix times x
    (Neumaier-summed, oracle), on `threads` host threads."""
    import oracle as orc
    chunks = np.array_split(np.arange(len(tgt_rows)), threads)
    out = np.zeros(len(tgt_rows), complex)

    def work(ix):
        if len(ix):
            out[ix] = orc.dense_rows(k, tgt_rows[ix], src, x)

    with ThreadPoolExecutor(threads) as ex:
        list(ex.map(work, chunks))
    return out


def gmres_section(bf, n, tol, rank_const, rcond, restart, solve_tol):
    """Second-kind boundary equation I/2 + (S + C)W on the curve (n points,
    10 per wavelength, order-6 corrections), right-hand side a plane wave in
    direction 30 degrees: FP64 GMRES (no restart, `restart` Krylov vectors)
    against the mixed solver (FP64 residuals around GMRES(restart) with the
    FP32 apply of the orthogonalized factorization)."""
    th = 2 * np.pi * np.arange(n) / n
    r = 1 + 0.3 * np.cos(5 * th)
    dr = -1.5 * np.sin(5 * th)
    pts = np.stack([r * np.cos(th), r * np.sin(th)], 1)
    w = np.sqrt(r * r + dr * dr) * (2 * np.pi / n)
    k = 2 * np.pi * n / (10 * CURVE_L)
    t0 = time.perf_counter()
    tree = bf.Quadtree(pts, 64)
    helm = bf.Helm2(k)
    F = bf.FacHelm2.make_multilevel(helm, tree, tree, bf.options(tol=tol, rank_const=rank_const, pass_through=True,
                                                                 pinv_rcond=rcond))
    A = bf.BoundaryOperator(F, helm, tree, w, kr_order=6, alpha=0.5)
    factor_s = time.perf_counter() - t0
    d = np.array([np.cos(np.pi / 6), np.sin(np.pi / 6)])
    b = np.exp(1j * k * (pts @ d))[tree.perm]
    A.solve(b, solve_tol, restart)  # warm-up
    bf.lib().bfhipSynchronize()
    t0 = time.perf_counter()
    x64, it64 = A.solve(b, solve_tol, restart)
    t64 = time.perf_counter() - t0
    r64 = float(np.linalg.norm(A @ x64 - b) / np.linalg.norm(b))
    t0 = time.perf_counter()
    ortho_s = F.orthogonalize()
    F.prepare_mixed("fp32")
    prep_s = time.perf_counter() - t0
    A.solve_mixed(b, solve_tol, restart=restart)  # warm-up
    t0 = time.perf_counter()
    xm, info = A.solve_mixed(b, solve_tol, restart=restart)
    tm = time.perf_counter() - t0
    rm = float(np.linalg.norm(A @ xm - b) / np.linalg.norm(b))
    out = {"workload": f"I/2 + (S + C)W on r=1+0.3cos(5t), N={n}, k={k:.1f} (10 ppw), KR order 6, plane wave at 30 deg, "
                       f"factor tol {tol:g}, C={rank_const}, rcond {rcond:g}",
           "N": n, "factor_s": factor_s, "solve_tol": solve_tol, "restart": restart,
           "fp64_iterations": it64, "fp64_solve_s": t64, "fp64_rel_residual": r64,
           "mixed_prepare_s": prep_s, "mixed_orthogonalize_s": ortho_s, "mixed_iterations": info["iterations"],
           "mixed_refinements": info["refinements"], "mixed_solve_s": tm, "mixed_rel_residual": rm,
           "mixed_vs_fp64_solution": float(np.linalg.norm(xm - x64) / np.linalg.norm(x64)),
           "solve_speedup": t64 / tm}
    print(f"[bench] gmres N={n}: factor {factor_s:.2f}s; fp64 {it64} it {t64 * 1e3:.1f} ms res {r64:.1e}; mixed "
          f"{info['iterations']} it / {info['refinements']} refinements {tm * 1e3:.1f} ms res {rm:.1e} "
          f"(prepare {prep_s:.2f}s)", file=sys.stderr, flush=True)
    del A, F
    return out


def batch_section(bf, count, nrhs, steps, rows, tune_cache=None):
    """Independent problems: sources uniform in [0,1]^2, targets in
    [2,3]x[0,1] (seeded), N log-uniform in [2^12, 2^18], k for 8 points per
    wavelength, leaf size from {32, 64, 128}, tol 1e-7; factored one by one
    and as one block-diagonal operator, then `nrhs` right-hand sides per
    problem applied in FP64 and (orthogonalized) FP32."""
    rng = np.random.default_rng(3)
    probs = []
    for _ in range(count):
        n = int(round(2 ** rng.uniform(12, 18)))
        leaf = int(rng.choice([32, 64, 128]))
        src = rng.uniform(0, 1, (n, 2))
        tgt = rng.uniform(0, 1, (n, 2))
        tgt[:, 0] += 2
        probs.append((src, tgt, 2 * np.pi * np.sqrt(n) / 8, leaf))
    trees = [(bf.Quadtree(s_, leaf), bf.Quadtree(t_, leaf)) for s_, t_, _, leaf in probs]
    helms = [bf.Helm2(k) for _, _, k, _ in probs]
    opts = [bf.options(tol=1e-7, rank_const=1.25, pass_through=True, pinv_rcond=1e-11) for _ in probs]
    nsrc = sum(len(p_[0]) for p_ in probs)
    ntgt = sum(len(p_[1]) for p_ in probs)
    X = rng.standard_normal((nsrc, nrhs)) + 1j * rng.standard_normal((nsrc, nrhs))
    # one at a time
    bf.lib().bfhipSynchronize()
    t0 = time.perf_counter()
    singles = [bf.FacHelm2.make_multilevel(h, ts, tt, o) for h, (ts, tt), o in zip(helms, trees, opts)]
    seq_factor = time.perf_counter() - t0
    xs, c0 = [], 0
    for s_, _, _, _ in probs:
        xs.append(bf.DeviceArray.from_numpy(np.ascontiguousarray(X[c0:c0 + len(s_)])))
        c0 += len(s_)
    ys = [bf.DeviceArray(16 * len(t_) * nrhs) for _, t_, _, _ in probs]
    for F, dx, dy in zip(singles, xs, ys):
        bf.mul_device(F, dx, dy, nrhs, "fp64")
    bf.lib().bfhipSynchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        for F, dx, dy in zip(singles, xs, ys):
            bf.mul_device(F, dx, dy, nrhs, "fp64")
    bf.lib().bfhipSynchronize()
    seq_apply = (time.perf_counter() - t0) / steps
    Yseq = np.concatenate([dy.to_numpy(np.complex128, (len(t_), nrhs)) for dy, (_, t_, _, _) in zip(ys, probs)])
    err_rows = None
    if rows:
        errs = []
        for i in range(0, count, max(1, count // 4)):
            s_, t_, k, _ = probs[i]
            ts, tt = trees[i]
            r = np.sort(rng.choice(len(t_), size=min(rows, len(t_)), replace=False))
            x = np.empty(len(s_), complex)
            c0 = sum(len(p_[0]) for p_ in probs[:i])
            x[ts.perm] = X[c0:c0 + len(s_), 0]
            yd = cpu_baseline_exact_rows(k, t_[r], s_, x, 4)
            y = np.empty(len(t_), complex)
            r0 = sum(len(p_[1]) for p_ in probs[:i])
            y[tt.perm] = Yseq[r0:r0 + len(t_), 0]
            errs.append(float(np.linalg.norm(y[r] - yd) / np.linalg.norm(yd)))
        err_rows = max(errs)
    print(f"[bench] batch {count} problems ({nsrc} points): one by one factor {seq_factor:.2f}s, "
          f"{nrhs}-RHS fp64 apply {seq_apply * 1e3:.1f} ms", file=sys.stderr, flush=True)
    del singles, xs, ys
    # all at once
    t0 = time.perf_counter()
    B = bf.FacHelm2.make_batched(helms, [t[0] for t in trees], [t[1] for t in trees], opts)
    batch_factor = time.perf_counter() - t0
    bst = B.stats()
    print(f"[bench] batch factor breakdown: plan {bst['plan_s']:.2f}s, kernels+apply plan {bst['kernels_s']:.2f}s "
          f"(compress {bst['compress_ms']:.0f} ms, sample {bst['sample_ms']:.0f} ms), apply plan "
          f"{bst['apply_plan_s']:.2f}s", file=sys.stderr, flush=True)
    dX = bf.DeviceArray.from_numpy(X)
    dY = bf.DeviceArray(16 * ntgt * nrhs)
    if tune_cache:
        B.tune_apply(nrhs, "fp64", tune_cache)
    bf.mul_device(B, dX, dY, nrhs, "fp64")
    bf.lib().bfhipSynchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bf.mul_device(B, dX, dY, nrhs, "fp64")
    bf.lib().bfhipSynchronize()
    batch_apply = (time.perf_counter() - t0) / steps
    Y64 = dY.to_numpy(np.complex128, (ntgt, nrhs))
    t0 = time.perf_counter()
    B.orthogonalize()
    B.prepare_mixed("fp32")
    mixed_prep = time.perf_counter() - t0
    if tune_cache:
        B.tune_apply(nrhs, "fp32", tune_cache)
    dX32 = bf.DeviceArray.from_numpy(X.astype(np.complex64))
    dY32 = bf.DeviceArray(8 * ntgt * nrhs)
    bf.mul_device(B, dX32, dY32, nrhs, "fp32")
    bf.lib().bfhipSynchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        bf.mul_device(B, dX32, dY32, nrhs, "fp32")
    bf.lib().bfhipSynchronize()
    batch_apply32 = (time.perf_counter() - t0) / steps
    Y32 = dY32.to_numpy(np.complex64, (ntgt, nrhs)).astype(complex)
    out = {"workload": f"{count} independent problems: sources U[0,1]^2 -> targets U[2,3]x[0,1], N log-uniform in "
                       f"[2^12, 2^18] (total {nsrc}), k = 2 pi sqrt(N)/8 (8 ppw), leaf from {{32,64,128}}, tol 1e-7, "
                       f"C=1.25, rcond 1e-11, {nrhs} RHS each",
           "problems": count, "total_points": nsrc, "entries": B.stats()["entries"],
           "factor_one_by_one_s": seq_factor, "factor_batched_s": batch_factor,
           "factor_batched_breakdown": {"plan_s": bst["plan_s"], "kernels_and_apply_plan_s": bst["kernels_s"],
                                        "compress_ms": bst["compress_ms"], "sample_ms": bst["sample_ms"],
                                        "apply_plan_s": bst["apply_plan_s"]},
           "factor_speedup": seq_factor / batch_factor,
           "apply_fp64_one_by_one_ms": seq_apply * 1e3, "apply_fp64_batched_ms": batch_apply * 1e3,
           "apply_fp32_batched_ms": batch_apply32 * 1e3, "mixed_prepare_s": mixed_prep,
           "batched_vs_one_by_one": float(np.linalg.norm(Y64 - Yseq) / np.linalg.norm(Yseq)),
           "fp32_rel_err_vs_fp64": float(np.linalg.norm(Y32 - Y64) / np.linalg.norm(Y64)),
           "fp64_rel_err_vs_dense_rows_max": err_rows}
    print(f"[bench] batch {count} problems ({nsrc} points): factor one by one {seq_factor:.2f}s, batched "
          f"{batch_factor:.2f}s; {nrhs}-RHS apply fp64 one by one {seq_apply * 1e3:.1f} ms, batched "
          f"{batch_apply * 1e3:.1f} ms, fp32 {batch_apply32 * 1e3:.1f} ms (err {out['fp32_rel_err_vs_fp64']:.1e}); "
          f"fp64 vs dense rows {err_rows}", file=sys.stderr, flush=True)
    del B
    return out


def stream_section(bf, n, tol, oversample, epochs, rows, threads):
    """Streaming factorization (bfFacStreamer*) of the headline geometry at
    n points (n sources on r = 1 + 0.3 cos 5t -> n targets on the circle of
    radius 3, k for 10 points per wavelength), leaf 64, tol `tol`, row cuts
    from the second level down, min rows / columns 20, the column blocks sampled on the device beforehand: plain
    (every truncated SVD in full) against rank-predicted (randomized SVDs of
    width predicted rank + `oversample` where that is under half the block's
    smaller side, full SVD as fallback), the model trained on the
    (geometry, numerical rank) samples the plain run recorded.  Times are
    the feed loops only; errors against exact rows."""
    rng = np.random.default_rng(0)
    src, tgt, k = curve_problem(n)
    ts, tt = bf.Quadtree(src, 64), bf.Quadtree(tgt, 64)
    qs = bf.QuadtreeStruct.from_address(ts.ptr)
    nodes = [(int(a.offset[0]), int(a.offset[4]))
             for a in (bf.QuadtreeNodeStruct * qs.numNodes).from_address(qs.nodes)]
    sp, tp = src[ts.perm], tgt[tt.perm]
    helm = bf.Helm2(k)
    blocks = {}

    def stream(model):
        S = bf.FacStreamer(tt, ts, tol, init_depth=2, min_rows=20, min_cols=20)
        S.set_rank_model(model, k, oversample)
        t = 0.0
        while not S.done:
            c = S.current_column_node()
            if c not in blocks:
                a, e = nodes[c]
                blocks[c] = helm.get_kernel_matrix(sp[a:e], tp)
            t0 = time.perf_counter()
            S.feed(blocks[c])
            t += time.perf_counter() - t0
        return S, t

    stream(None)  # first pass: samples the blocks and warms the kernels
    S0, t_plain = stream(None)
    X, y = S0.rank_samples()
    model = bf.RankModel(1)
    t0 = time.perf_counter()
    mse = model.train(X, y, epochs=epochs, batch=512, lr=2e-3)
    t_train = time.perf_counter() - t0
    mae = float(np.abs(model.predict(X) - y).mean())
    S1, t_model = stream(model)
    st = S1.sketch_stats()
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    errs = [None, None]
    if rows:
        r = np.sort(rng.choice(n, size=min(rows, n), replace=False))
        yd = cpu_baseline_exact_rows(k, tp[r], sp, x, threads)
        errs = [float(np.linalg.norm((S.fac().mat() @ x)[r] - yd) / np.linalg.norm(yd)) for S in (S0, S1)]
    print(f"[bench] stream N={n} k={k:.1f} tol={tol:g}: plain {t_plain:.2f}s ({len(y)} SVDs), rank model "
          f"trained on its samples in {t_train:.2f}s (MAE {mae:.2f}), rank-predicted {t_model:.2f}s "
          f"({st['sketched']} sketched, {st['fell_back']} fell back); errors {errs[0]} / {errs[1]}",
          file=sys.stderr, flush=True)
    return {"workload": f"bfFacStreamer: {n} sources on r=1+0.3cos(5t) -> {n} targets on r=3, k={k:.1f} (10 ppw), "
                        f"leaf 64, tol {tol:g}, row tree from depth 2, min rows/cols 20",
            "N": n, "plain_s": t_plain, "rank_predicted_s": t_model, "speedup": t_plain / t_model,
            "svds": st["svds"], "sketched": st["sketched"], "fell_back": st["fell_back"],
            "train_samples": int(len(y)), "train_s": t_train, "train_mse": mse, "train_mae": mae,
            "plain_rel_err_vs_rows": errs[0], "rank_predicted_rel_err_vs_rows": errs[1]}


def cpu_baseline(n_sample, tol, rank_const, rcond, n_full_entries, threads, seconds=4.0):
    """Oracle (the C restatement) factorization of the same geometry with the
    same planner at n_sample points (one thread: the factorization is a
    serial C loop), then its apply timed as throughput on `threads` host
    threads, each applying its own right-hand side over and over for
    `seconds` (the C calls release the GIL, so the threads run on separate
    cores); the rate is scaled to the full size by the ratio of stored
    entries (the apply streams every entry once)."""
    import oracle as orc
    src, tgt, k = curve_problem(n_sample)
    t0 = time.perf_counter()
    O = orc.Fac(k, src, tgt, 64, tol, 128 * 128, True, rank_const=rank_const, pinv_rcond=rcond)
    t_fac = time.perf_counter() - t0
    lt, ft, bt = O.plan()
    entries = float((bt[:, 2] * bt[:, 3])[bt[:, 4] != 2].sum() + (lt[:, 3] * lt[:, 4])[lt[:, 0] == 0].sum())
    xs = [orc.complex_randn(n_sample) for _ in range(threads)]
    O.apply(xs[0])
    stop = time.perf_counter() + seconds

    def work(x):
        c = 0
        while time.perf_counter() < stop:
            O.apply(x)
            c += 1
        return c

    t0 = time.perf_counter()
    with ThreadPoolExecutor(threads) as ex:
        counts = list(ex.map(work, xs))
    wall = time.perf_counter() - t0
    rate = sum(counts) / wall  # applies per second at n_sample, all threads
    scaled = rate * entries / n_full_entries
    return {
        "value": scaled,
        "unit": "RHS/s",
        "cores": threads,
        "kind": "port",
        "sample": (f"oracle C restatement, same geometry/planner at N={n_sample} (factor {t_fac:.1f}s on 1 thread, "
                   f"{entries:.3g} stored entries); apply throughput {rate:.1f} RHS/s on {threads} threads over "
                   f"{wall:.1f}s ({sum(counts)} applies), scaled x{entries / n_full_entries:.2e} by stored entries "
                   f"to N=2^20"),
        "factor_s_sample": t_fac,
        "apply_rhs_per_s_sample": rate,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--tol", type=float, default=1e-8)
    ap.add_argument("--rank-const", type=float, default=1.25)
    ap.add_argument("--pinv-rcond", type=float, default=1e-11)
    ap.add_argument("--mixed-tol", type=float, default=1e-8)
    ap.add_argument("--mixed-rcond", type=float, default=1e-11)
    ap.add_argument("--mixed-rank-const", type=float, default=0.0, help="0 = same as --rank-const")
    ap.add_argument("--comp", type=int, default=2,
                    help="accumulation of the FP32 applies (bfhipSetApplyCompensation: 0 FP32, 1 FP64 sums, "
                         "2 + m FP32 matrix-core tiles folded into FP64 sums every m + 1 steps)")
    ap.add_argument("--rows", type=int, default=4096, help="exact dense rows for the error check (0 = skip)")
    ap.add_argument("--threads", type=int, default=16)
    ap.add_argument("--mixed-rhs", type=int, default=256)
    ap.add_argument("--mixed-steps", type=int, default=5)
    ap.add_argument("--mixed-ortho", type=int, default=1, help="orthogonalize the mixed factorization (1) or not (0)")
    ap.add_argument("--gmres-n", type=int, default=1 << 18, help="boundary-equation GMRES section size (0 = skip)")
    ap.add_argument("--gmres-restart", type=int, default=50)
    ap.add_argument("--gmres-tol", type=float, default=1e-10)
    ap.add_argument("--batch", type=int, default=64, help="independent-problems section size (0 = skip)")
    ap.add_argument("--batch-rhs", type=int, default=16)
    ap.add_argument("--stream-n", type=int, default=8192, help="streaming / rank-model section size (0 = skip)")
    ap.add_argument("--stream-tol", type=float, default=1e-7)
    ap.add_argument("--split-tol", type=float, default=1e-9,
                    help="tolerance of the hi+lo FP32 (fp32x2) section (0 = skip; needs --rows)")
    ap.add_argument("--tree", choices=["device", "host"], default="device",
                    help="build the two quadtrees on the GPU (bfhipQuadtreeInitWithLeafSize) or on the host")
    ap.add_argument("--factor-warmup", type=int, default=1 << 14,
                    help="points of an untimed factorization + apply run first, so the timed one does not carry "
                         "the process's one-time costs (kernel code loading, first touch); 0 = none")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-sample-n", type=int, default=8192)
    ap.add_argument("--tune", type=int, default=1,
                    help="run the apply autotuner (bfhipFacTuneApply) before each timed apply shape (untimed)")
    ap.add_argument("--graphs", type=int, default=1, help="replay the timed applies as hipGraphs")
    ap.add_argument("--tune-cache", default=os.path.join(os.environ.get("TMPDIR", "/tmp"), "bf_amd_apply_tune.txt"))
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also run a plain copy kernel of this many MiB (reference for HBM byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); "
              "this build runs the path on one GPU and has no multi-GPU sharding", file=sys.stderr)
        sys.exit(2)

    import butterfly_amd as bf

    if bf.lib().bfhipDeviceCount() < 1:
        raise SystemExit("no HIP device")
    bf.lib().bfhipSetApplyCompensation(args.comp)
    bf.set_apply_graphs(args.graphs)
    n = args.n
    src, tgt, k = curve_problem(n)
    print(f"[bench] N={n} k={k:.1f} leaf=64 tol={args.tol:g} C={args.rank_const} rcond={args.pinv_rcond:g}",
          file=sys.stderr, flush=True)

    opts = bf.options(tol=args.tol, rank_const=args.rank_const, pass_through=True, pinv_rcond=args.pinv_rcond)
    dev_tree = args.tree == "device"
    if args.factor_warmup:
        wsrc, wtgt, wk = curve_problem(args.factor_warmup)
        wts, wtt = bf.Quadtree(wsrc, 64, device=dev_tree), bf.Quadtree(wtgt, 64, device=dev_tree)
        Fw = bf.FacHelm2.make_multilevel(bf.Helm2(wk), wts, wtt, opts)
        Fw @ bf.complex_randn(args.factor_warmup)
        del Fw, wts, wtt
    t0 = time.perf_counter()
    ts, tt = bf.Quadtree(src, 64, device=dev_tree), bf.Quadtree(tgt, 64, device=dev_tree)
    tree_s = time.perf_counter() - t0
    F = bf.FacHelm2.make_multilevel(bf.Helm2(k), ts, tt, opts)
    factor_s = time.perf_counter() - t0
    st = F.stats()
    print(f"[bench] factor {factor_s:.2f}s (plan {st['plan_s']:.2f}s, sample {st['sample_ms']:.0f} ms, "
          f"compress {st['compress_ms']:.0f} ms) entries {st['entries']:.3e}", file=sys.stderr, flush=True)

    tune_s = {}
    if args.tune:
        t0 = time.perf_counter()
        F.tune_apply(1, "fp64", args.tune_cache)
        tune_s["fp64_1rhs"] = time.perf_counter() - t0
    apply_cfg = {"fp64_1rhs": F.apply_config(1)}
    bf.seed(0)
    x = bf.complex_randn(n)
    dx = bf.DeviceArray.from_numpy(x[ts.perm])
    dy = bf.DeviceArray(16 * n)
    for _ in range(args.warmup):
        bf.mul_device(F, dx, dy, 1, "fp64")
    bf.lib().bfhipSynchronize()
    bf.reset_kernel_times()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        bf.mul_device(F, dx, dy, 1, "fp64")
    bf.lib().bfhipSynchronize()
    t1 = time.perf_counter()
    ms_per_step = (t1 - t0) * 1e3 / args.steps
    sums, cnts = bf.kernel_time_sums()
    apply_ms = sums[2] / max(cnts[2], 1)
    bytes_per_apply = st["apply_factor_bytes"] + st["apply_vector_bytes"]
    achieved = bytes_per_apply / (apply_ms * 1e-3) / 1e9
    yt = dy.to_numpy(np.complex128, n)
    y = np.empty(n, complex)
    y[tt.perm] = yt
    print(f"[bench] apply {ms_per_step:.3f} ms/step (events {apply_ms:.3f} ms), {achieved:.0f} GB/s",
          file=sys.stderr, flush=True)

    if args.calibrate_copy:
        nb = args.calibrate_copy << 20
        a, b = bf.DeviceArray(nb), bf.DeviceArray(nb)
        bf.lib().bfhipDeviceCopy(a.ptr, b.ptr, nb)
        del a, b

    rel_err = None
    if args.rows:
        rng = np.random.default_rng(1)
        rows = np.sort(rng.choice(n, size=min(args.rows, n), replace=False))
        t0 = time.perf_counter()
        yd = cpu_baseline_exact_rows(k, tgt[rows], src, x, args.threads)
        rel_err = float(np.linalg.norm(y[rows] - yd) / np.linalg.norm(yd))
        print(f"[bench] rel err vs {len(rows)} exact rows: {rel_err:.3e} ({time.perf_counter() - t0:.1f}s)",
              file=sys.stderr, flush=True)

    # HBM bytes per apply measured by rocprofv3 FETCH_SIZE / WRITE_SIZE for
    # this kernel and workload, scaled by the counters' reading of a plain
    # copy (the summary file gives the commands)
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "k_apply1_traffic.json")
    if n == 1 << 20 and args.tol == 1e-8 and args.rank_const == 1.25 and os.path.exists(tpath):
        traffic = json.load(open(tpath))["traffic_bytes_per_apply"]

    mixed = None
    if args.mixed_rhs:
        nr = args.mixed_rhs
        bf.seed(2)
        X = bf.complex_randn(n * nr).reshape(n, nr)
        Xt = np.ascontiguousarray(X[ts.perm])
        dX = bf.DeviceArray.from_numpy(Xt)
        dY = bf.DeviceArray(16 * n * nr)
        if args.tune:
            t0 = time.perf_counter()
            F.tune_apply(nr, "fp64", args.tune_cache)
            tune_s[f"fp64_{nr}rhs"] = time.perf_counter() - t0
        apply_cfg[f"fp64_{nr}rhs"] = F.apply_config(nr)
        bf.mul_device(F, dX, dY, nr, "fp64")
        bf.lib().bfhipSynchronize()
        t0 = time.perf_counter()
        for _ in range(args.mixed_steps):
            bf.mul_device(F, dX, dY, nr, "fp64")
        bf.lib().bfhipSynchronize()
        t64 = (time.perf_counter() - t0) / args.mixed_steps
        Y64 = dY.to_numpy(np.complex128, (n, nr))
        del dX, dY
        del F  # free its HBM (factors and 256-RHS work buffers)
        # the mixed form: its own factorization at the looser tolerance,
        # stored and applied in FP32
        t0 = time.perf_counter()
        mopts = bf.options(tol=args.mixed_tol, rank_const=args.mixed_rank_const or args.rank_const, pass_through=True,
                           pinv_rcond=args.mixed_rcond)
        Fm = bf.FacHelm2.make_multilevel(bf.Helm2(k), ts, tt, mopts)
        ortho_s = Fm.orthogonalize() if args.mixed_ortho else 0.0
        Fm.prepare_mixed("fp32")
        factor_m = time.perf_counter() - t0
        # accuracy of the mixed factorization itself (its FP64 apply), then
        # drop the FP64 factors
        errm64_rows = None
        if args.rows:
            bf.mul_device(Fm, dx, dy, 1, "fp64")
            ym = np.empty(n, complex)
            ym[tt.perm] = dy.to_numpy(np.complex128, n)
            errm64_rows = float(np.linalg.norm(ym[rows] - yd) / np.linalg.norm(yd))
        bf.lib().bfhipFacReleaseFp64(Fm.ptr)
        if args.tune:
            t0 = time.perf_counter()
            Fm.tune_apply(nr, "fp32", args.tune_cache)
            Fm.tune_apply(1, "fp32", args.tune_cache)
            tune_s["fp32"] = time.perf_counter() - t0
        apply_cfg[f"fp32_{nr}rhs"] = Fm.apply_config(nr, "fp32")
        apply_cfg["fp32_1rhs"] = Fm.apply_config(1, "fp32")
        dX32 = bf.DeviceArray.from_numpy(Xt.astype(np.complex64))
        dY32 = bf.DeviceArray(8 * n * nr)
        bf.mul_device(Fm, dX32, dY32, nr, "fp32")
        bf.lib().bfhipSynchronize()
        t0 = time.perf_counter()
        for _ in range(args.mixed_steps):
            bf.mul_device(Fm, dX32, dY32, nr, "fp32")
        bf.lib().bfhipSynchronize()
        t32 = (time.perf_counter() - t0) / args.mixed_steps
        Y32 = dY32.to_numpy(np.complex64, (n, nr)).astype(complex)
        err32 = float(np.linalg.norm(Y32 - Y64) / np.linalg.norm(Y64))
        del dX32, dY32
        dx1 = bf.DeviceArray.from_numpy(x[ts.perm].astype(np.complex64))
        dy1 = bf.DeviceArray(8 * n)
        bf.mul_device(Fm, dx1, dy1, 1, "fp32")
        bf.lib().bfhipSynchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bf.mul_device(Fm, dx1, dy1, 1, "fp32")
        bf.lib().bfhipSynchronize()
        t32_1 = (time.perf_counter() - t0) / args.steps
        err32_rows = None
        if args.rows:
            y32 = np.empty(n, complex)
            y32[tt.perm] = dy1.to_numpy(np.complex64, n).astype(complex)
            err32_rows = float(np.linalg.norm(y32[rows] - yd) / np.linalg.norm(yd))
        mixed = {"nrhs": nr, "tol": args.mixed_tol, "pinv_rcond": args.mixed_rcond, "factor_s": factor_m,
                 "orthogonalize_s": ortho_s, "orthogonalized": bool(args.mixed_ortho),
                 "factor_speedup": factor_s / factor_m, "entries": Fm.stats()["entries"],
                 "fp64_apply_s": t64, "fp32_apply_s": t32, "fp64_rhs_per_s": nr / t64,
                 "fp32_rhs_per_s": nr / t32, "apply_speedup": t64 / t32, "fp32_rel_err_vs_fp64": err32,
                 "fp32_rel_err_vs_dense_rows": err32_rows, "fp64_rel_err_vs_dense_rows": errm64_rows,
                 "fp32_1rhs_ms": t32_1 * 1e3, "fp64_1rhs_ms": ms_per_step}
        print(f"[bench] mixed: factor {factor_m:.2f}s (orthogonalize {ortho_s:.2f}s); {nr} RHS fp64 "
              f"{t64 * 1e3:.1f} ms, fp32 {t32 * 1e3:.1f} ms, fp32 err vs fp64 {err32:.2e}, vs exact rows: "
              f"1 RHS fp32 {err32_rows}, its fp64 apply {errm64_rows}; 1 RHS fp32 {t32_1 * 1e3:.2f} ms",
              file=sys.stderr, flush=True)
        del Fm

    # the split form: a tighter factorization (tol 1e-9) stored as hi + lo
    # FP32 pairs, applied with FP32 vectors (pairs summed in FP64)
    split = None
    if args.split_tol and args.rows:
        F = None  # the FP64 factorization's HBM goes first
        t0 = time.perf_counter()
        sopts = bf.options(tol=args.split_tol, rank_const=args.rank_const, pass_through=True,
                           pinv_rcond=args.pinv_rcond)
        Fs = bf.FacHelm2.make_multilevel(bf.Helm2(k), ts, tt, sopts)
        fs_s = time.perf_counter() - t0
        bf.mul_device(Fs, dx, dy, 1, "fp64")
        bf.lib().bfhipSynchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bf.mul_device(Fs, dx, dy, 1, "fp64")
        bf.lib().bfhipSynchronize()
        ts64 = (time.perf_counter() - t0) / args.steps
        ys = np.empty(n, complex)
        ys[tt.perm] = dy.to_numpy(np.complex128, n)
        errs64 = float(np.linalg.norm(ys[rows] - yd) / np.linalg.norm(yd))
        sortho = Fs.orthogonalize() if args.mixed_ortho else 0.0
        Fs.prepare_mixed("fp32x2")
        bf.lib().bfhipFacReleaseFp64(Fs.ptr)
        dx1 = bf.DeviceArray.from_numpy(x[ts.perm].astype(np.complex64))
        dy1 = bf.DeviceArray(8 * n)
        bf.mul_device(Fs, dx1, dy1, 1, "fp32x2")
        bf.lib().bfhipSynchronize()
        t0 = time.perf_counter()
        for _ in range(args.steps):
            bf.mul_device(Fs, dx1, dy1, 1, "fp32x2")
        bf.lib().bfhipSynchronize()
        tsx = (time.perf_counter() - t0) / args.steps
        ysx = np.empty(n, complex)
        ysx[tt.perm] = dy1.to_numpy(np.complex64, n).astype(complex)
        errsx = float(np.linalg.norm(ysx[rows] - yd) / np.linalg.norm(yd))
        split = {"tol": args.split_tol, "factor_s": fs_s, "orthogonalize_s": sortho,
                 "entries": Fs.stats()["entries"], "fp64_1rhs_ms": ts64 * 1e3,
                 "fp64_rel_err_vs_dense_rows": errs64, "fp32x2_1rhs_ms": tsx * 1e3,
                 "fp32x2_rel_err_vs_dense_rows": errsx}
        print(f"[bench] split hi+lo at tol {args.split_tol:g}: factor {fs_s:.2f}s, {split['entries']:.3e} entries; "
              f"1 RHS fp64 {ts64 * 1e3:.2f} ms at {errs64:.2e}, fp32x2 {tsx * 1e3:.2f} ms at {errsx:.2e} "
              f"(vs exact rows)", file=sys.stderr, flush=True)
        del Fs, dx1, dy1

    gm = None
    if args.gmres_n:
        gm = gmres_section(bf, args.gmres_n, args.tol, args.rank_const, args.pinv_rcond, args.gmres_restart,
                           args.gmres_tol)

    bt = None
    if args.batch:
        bt = batch_section(bf, args.batch, args.batch_rhs, max(1, args.mixed_steps), 256 if args.rows else 0,
                           args.tune_cache if args.tune else None)

    sm = None
    if args.stream_n:
        try:
            sm = stream_section(bf, args.stream_n, args.stream_tol, 8, 200, 256 if args.rows else 0, args.threads)
        except bf.BfError as e:  # reported, not fatal: the headline line must still come out
            sm = {"error": str(e)}
            print(f"[bench] stream section failed: {e}", file=sys.stderr, flush=True)

    cpu = None
    if not args.no_cpu_baseline:
        cpu = cpu_baseline(args.cpu_sample_n, args.tol, args.rank_const, args.pinv_rcond, st["entries"],
                           args.threads)
        print(f"[bench] cpu baseline {cpu['value']:.3f} RHS/s ({cpu['sample']})", file=sys.stderr, flush=True)

    line = {
        "metric": METRIC,
        "value": 1e3 / ms_per_step,
        "unit": "RHS/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c128 (FP64 complex)",
        "data": "synthetic",
        "config": {
            "workload": "2^20 sources on r=1+0.3cos(5t) -> 2^20 targets on r=3, k=73064.8 (10 ppw), "
                        "leaf 64, rank tol 1e-8, FP64 factor + apply to 1 complex Gaussian RHS",
            "N": n, "k": k, "leaf_size": 64, "tol": args.tol, "rank_const": args.rank_const,
            "pinv_rcond": args.pinv_rcond,
            "planner": "pass-through", "nrhs": 1, "graphs": bool(args.graphs), "tuned": bool(args.tune),
        },
        "tune_s": tune_s,
        "apply_config": apply_cfg,
        "factor_s": factor_s,
        "factor_warmup": (f"one untimed N={args.factor_warmup} factorization + apply before the timed one"
                          if args.factor_warmup else None),
        "factor_breakdown": {"trees_s": tree_s, "trees": args.tree, "tree_plan_s": st["plan_s"], "sample_ms": st["sample_ms"],
                             "compress_ms": st["compress_ms"], "apply_plan_s": st["apply_plan_s"],
                             "entries": st["entries"], "lstsq_blocks": st["lstsq_blocks"],
                             "kernel_blocks": st["kernel_blocks"], "h0_evals": st["h0_evals"]},
        "rel_err_vs_dense_rows": rel_err,
        "dense_rows": args.rows,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "kernel": "k_apply1 (all stages of one apply)", "bytes_per_apply": bytes_per_apply,
                     "event_ms_per_apply": apply_ms},
        "mixed": mixed,
        "split": split,
        "gmres": gm,
        "batch": bt,
        "stream": sm,
        "cpu_baseline": cpu,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
