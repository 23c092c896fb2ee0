"""The rank model (FP32 MLP, HIP training and inference kernels) and the
randomized truncated SVDs it sizes (GPU).

Checkers: the float64 numpy restatement of the model (oracle/rank_mlp.py,
itself pinned by central differences) for the forward pass, the gradient
and Adam's first step; numpy's SVD for the randomized SVDs; the plain
streaming factorization (full SVDs) for the streamer with a model.

Tolerances: FP32 forward and gradient vs float64, 1e-4 relative (a few
hundred FP32 products per output); randomized SVDs accepted only when
||A - Q Q^H A||_F <= tol sigma_0, so their products are within 2 tol
sigma_0 of A (spectral norm).
"""
import numpy as np
import pytest

from oracle import rank_mlp as R

pytestmark = pytest.mark.gpu

IDENTITY_NORM = np.r_[np.zeros(6), np.ones(6), 0.0, 1.0].astype(np.float32)


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _model(bf, seed=3):
    m = bf.RankModel(seed)
    p, _ = m.params
    m.set_params(p, IDENTITY_NORM)  # features and targets used as given
    return m, p.astype(np.float64)


def test_rank_model_forward_matches_float64(gpu_lib):
    bf = gpu_lib
    m, p = _model(bf)
    X = np.random.default_rng(1).standard_normal((1000, 6)).astype(np.float32)  # 15 full tiles + a ragged one
    got = m.predict(X)
    want, _ = R.forward(p, X)
    assert rel(got, want) < 1e-4
    assert np.abs(got - want).max() < 1e-4 * np.abs(want).max()


def test_rank_model_gradient_and_adam_step_match_float64(gpu_lib):
    bf = gpu_lib
    m, p = _model(bf)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((777, 6)).astype(np.float32)
    y = rng.standard_normal(777).astype(np.float32)
    g, loss = m.gradient(X, y)
    lw, gw = R.loss_grad(p, X, y)
    assert abs(loss - lw) < 1e-5 * lw
    assert rel(g, gw) < 1e-4
    offs = np.cumsum([0] + [int(np.prod(s)) for s in R.SHAPES])
    for lo, hi in zip(offs[:-1], offs[1:]):  # every layer, not just the large ones
        assert rel(g[lo:hi], gw[lo:hi]) < 1e-3
    # one full-batch Adam step: p - lr g / (|g| + eps), where the gradient is
    # clear of FP32 noise
    lr = 1e-3
    m.train(X, y, epochs=1, batch=777, lr=lr)
    p1, _ = m.params
    want = R.adam_first_step(p, gw, lr)
    big = np.abs(gw) > 1e-4 * np.abs(gw).max()
    assert np.abs(p1 - want)[big].max() < 2e-3 * lr + 1e-7


def test_rank_model_learns_a_rank_law(gpu_lib):
    """Ranks from a smooth law of the features (k d, separation, tol) are
    learned to well under one rank on held-out samples."""
    bf = gpu_lib
    rng = np.random.default_rng(4)
    n = 24000
    X = np.stack([rng.uniform(-6, 0, n), rng.uniform(-6, 0, n), rng.uniform(0, 60, n), rng.uniform(1, 4, n),
                  rng.integers(2, 9, n), rng.uniform(-12, -4, n)], 1).astype(np.float32)
    r = np.ceil(0.5 * X[:, 2] / X[:, 3] - 1.2 * X[:, 5]).astype(np.float32)
    m = bf.RankModel(5)
    loss = m.train(X[:20000], r[:20000], epochs=40, batch=1024, lr=2e-3)
    pred = m.predict(X[20000:])
    mae = np.abs(pred - r[20000:]).mean()
    print("rank law: last-epoch MSE", loss, "held-out MAE", mae)
    assert mae < 0.5 and loss < 0.5


def _graded(rng, m, n, k, scale=1.0):
    """A = U diag(s) V^H with k singular values in [1e-3, 1] and the rest
    below 1e-13 (a clear gap at any cut in between)."""
    r = min(m, n)
    U, _ = np.linalg.qr(rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r)))
    V, _ = np.linalg.qr(rng.standard_normal((n, r)) + 1j * rng.standard_normal((n, r)))
    s = np.r_[np.geomspace(1, 1e-3, k), 1e-14 * rng.random(r - k)]
    return scale * (U * s) @ V.conj().T


def test_sketch_svd_matches_full_svd(gpu_lib):
    bf = gpu_lib
    rng = np.random.default_rng(6)
    tol = 1e-8
    mats = [_graded(rng, 300, 120, 20), _graded(rng, 90, 200, 15, 1e-3), _graded(rng, 64, 64, 40),
            _graded(rng, 600, 400, 24)]  # the last has more columns than the full SVD kernel takes
    widths = [28, 23, 48, 32]
    out = bf.svd_batched_sketch(mats, tol, widths, seed=11)
    for A, (U, W, k, fell), w in zip(mats, out, widths):
        s = np.linalg.svd(A, compute_uv=False)
        assert not fell and k == int(np.sum(s >= tol * s[0]))
        assert np.abs(U.conj().T @ U - np.eye(k)).max() < 1e-12
        assert np.linalg.norm(U @ W - A, 2) <= 2 * tol * s[0]
        assert np.allclose(np.linalg.norm(W, axis=1), s[:k], rtol=1e-10)
    # a width below the rank is rejected and redone in full
    small = mats[:3]
    out = bf.svd_batched_sketch(small, tol, [8, 8, 0], seed=11)
    ref = bf.svd_batched(small, tol)
    for (U, W, k, fell), (U0, W0, k0), wd in zip(out, ref, [8, 8, 0]):
        assert fell == (wd > 0) and k == k0
        assert np.array_equal(U, U0) and np.array_equal(W, W0)


def test_streamer_with_rank_model(gpu_lib):
    """Stream a Helmholtz kernel matrix with plain SVDs (recording the
    samples), train the model on them, stream again with randomized SVDs of
    width predicted rank + 8: the same matrix to the streamer's accuracy
    against exact rows (oracle), with the large SVDs sketched."""
    bf = gpu_lib
    import oracle as orc
    rng = np.random.default_rng(8)
    n = 4096
    src = rng.uniform(0, 1, (n, 2))
    tgt = rng.uniform(0, 1, (n, 2))
    tgt[:, 0] += 2
    k = 2 * np.pi * np.sqrt(n) / 8
    tol = 1e-8
    ts, tt = bf.Quadtree(src, 64), bf.Quadtree(tgt, 64)
    sp, tp = src[ts.perm], tgt[tt.perm]
    nodes = QuadNodes(bf, ts)
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
    st0 = S0.sketch_stats()
    assert len(y) == st0["svds"] > 100 and st0["sketched"] == 0
    m = bf.RankModel(9)
    m.train(X, y, epochs=200, batch=256, lr=2e-3)
    mae = np.abs(m.predict(X) - y).mean()
    S1 = stream(m)
    st1 = S1.sketch_stats()
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    rows = np.sort(rng.choice(n, 128, replace=False))
    want = orc.dense_rows(k, tp[rows], sp, x)
    e0 = rel((S0.fac().mat() @ x)[rows], want)
    e1 = rel((S1.fac().mat() @ x)[rows], want)
    print("streamer: samples", len(y), "train MAE", mae, "stats", st1, "errors", e0, e1)
    assert st1["sketched"] > 10 and e0 < 20 * tol and e1 < 20 * tol


def QuadNodes(bf, tree):
    """(first, last) point of every node of a Quadtree, by node index."""
    qs = bf.QuadtreeStruct.from_address(tree.ptr)
    arr = (bf.QuadtreeNodeStruct * qs.numNodes).from_address(qs.nodes)
    return [(int(a.offset[0]), int(a.offset[4])) for a in arr]
