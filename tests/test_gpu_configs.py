"""The batched and the learned-rank workloads against exact FP64 dense rows
(GPU).

Batched independent problems: a draw in the manner of the benchmark's
batch -- per problem N log-uniform (here in [2^12, 2^15]), sources uniform
in [0,1]^2 and targets in [2,3]x[0,1], k for 8 points per wavelength, leaf
size from {32, 64, 128}, tolerance 1e-7 -- factored as one block-diagonal
operator; 16 right-hand sides per problem applied in FP64 and, after
orthogonalization, in FP32.  Every problem's output rows are checked
against the oracle's exact rows (Neumaier-summed FP64 sums of
(i/4) H0(k r) x over all sources).

Learned ranks: the streaming factorization of the curve r = 1 + 0.3 cos 5t
to the circle of radius 3 (10 points per wavelength along the curve) with
the rank model trained on the ranks the plain stream harvested, then the
randomized SVDs of width predicted rank + 8 with the full SVD as fallback;
its applies to 1 and to 16 right-hand sides against exact rows.
"""
import numpy as np
import pytest

import oracle as orc
from conftest import curve_problem

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def test_batched_draw_matches_dense_rows(gpu_lib):
    bf = gpu_lib
    rng = np.random.default_rng(5)
    count, nrhs, rows = 16, 16, 48
    probs = []
    for _ in range(count):
        n = int(round(2 ** rng.uniform(12, 15)))
        leaf = int(rng.choice([32, 64, 128]))
        src = rng.uniform(0, 1, (n, 2))
        tgt = rng.uniform(0, 1, (n, 2))
        tgt[:, 0] += 2
        probs.append((src, tgt, 2 * np.pi * np.sqrt(n) / 8, leaf))
    trees = [(bf.Quadtree(s, leaf), bf.Quadtree(t, leaf)) for s, t, _, leaf in probs]
    opts = [bf.options(tol=1e-7, rank_const=1.25, pass_through=True, pinv_rcond=1e-11) for _ in probs]
    B = bf.FacHelm2.make_batched([bf.Helm2(k) for _, _, k, _ in probs], [t[0] for t in trees],
                                 [t[1] for t in trees], opts)
    ns = [len(p[0]) for p in probs]
    X = rng.standard_normal((sum(ns), nrhs)) + 1j * rng.standard_normal((sum(ns), nrhs))
    Y64 = B @ X
    B.orthogonalize()
    B.prepare_mixed("fp32")
    dX = bf.DeviceArray.from_numpy(X.astype(np.complex64))
    dY = bf.DeviceArray(8 * nrhs * B.shape[0])
    bf.mul_device(B, dX, dY, nrhs, "fp32")
    Y32 = dY.to_numpy(np.complex64, (B.shape[0], nrhs)).astype(complex)
    c0 = r0 = 0
    worst64 = worst32 = 0.0
    for (src, tgt, k, leaf), (ts, tt) in zip(probs, trees):
        n = len(src)
        r = np.sort(rng.choice(n, rows, replace=False))
        for col in (0, nrhs - 1):
            x = np.empty(n, complex)
            x[ts.perm] = X[c0:c0 + n, col]
            want = orc.dense_rows(k, tgt[r], src, x)
            for Y, tag in ((Y64, 64), (Y32, 32)):
                y = np.empty(n, complex)
                y[tt.perm] = Y[r0:r0 + n, col]
                e = rel(y[r], want)
                if tag == 64:
                    worst64 = max(worst64, e)
                else:
                    worst32 = max(worst32, e)
        c0 += n
        r0 += n
    print(f"{count} problems ({sum(ns)} points): FP64 worst {worst64:.2e}, FP32 16 RHS worst {worst32:.2e}")
    assert worst64 < 1e-6 and worst32 < 5e-6


def test_rank_model_stream_on_the_curve(gpu_lib):
    bf = gpu_lib
    n, tol = 4096, 1e-7
    src, tgt, k = curve_problem(n)
    ts, tt = bf.Quadtree(src, 64), bf.Quadtree(tgt, 64)
    sp, tp = src[ts.perm], tgt[tt.perm]
    qs = bf.QuadtreeStruct.from_address(ts.ptr)
    nodes = [(int(a.offset[0]), int(a.offset[4]))
             for a in (bf.QuadtreeNodeStruct * qs.numNodes).from_address(qs.nodes)]
    helm = bf.Helm2(k)
    blocks = {}

    def stream(model):
        S = bf.FacStreamer(tt, ts, tol, init_depth=1, min_rows=20, min_cols=20)
        S.set_rank_model(model, k, 8)
        while not S.done:
            c = S.current_column_node()
            if c not in blocks:
                blocks[c] = helm.get_kernel_matrix(sp[nodes[c][0]:nodes[c][1]], tp)
            S.feed(blocks[c])
        return S

    S0 = stream(None)
    X, y = S0.rank_samples()
    model = bf.RankModel(4)
    model.train(X, y, epochs=300, batch=256, lr=2e-3)
    mae = float(np.abs(model.predict(X) - y).mean())
    S1 = stream(model)
    st = S1.sketch_stats()
    rng = np.random.default_rng(6)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    r = np.sort(rng.choice(n, 256, replace=False))
    want = orc.dense_rows(k, tp[r], sp, x)
    M0, M1 = S0.fac().mat(), S1.fac().mat()
    e0, e1 = rel((M0 @ x)[r], want), rel((M1 @ x)[r], want)
    # 16 right-hand sides through the device product
    Xr = rng.standard_normal((n, 16)) + 1j * rng.standard_normal((n, 16))
    Y = M1 @ Xr
    wantY = np.stack([orc.dense_rows(k, tp[r], sp, Xr[:, j]) for j in (0, 15)], 1)
    e16 = rel(Y[r][:, [0, 15]], wantY)
    print(f"curve N={n}: {len(y)} samples, MAE {mae:.2f}, stats {st}, errors plain {e0:.2e} model {e1:.2e} "
          f"16 RHS {e16:.2e}")
    assert len(y) > 50 and st["sketched"] > 0
    assert e0 < 20 * tol and e1 < 20 * tol and e16 < 20 * tol
