"""The streaming factorization (bfFacStreamer*, bfFacGetMat) and the batched
device SVD against the numpy restatement in oracle/streamer.py (GPU).

Tolerances: the device SVD (one-sided Jacobi) and LAPACK agree on singular
values to ~1e-13 relative; the truncation decisions (row cuts, ranks) are
compared exactly; products of two factorizations that each reproduce the
matrix to ~tol agree to a small multiple of tol.
"""
import ctypes as C

import numpy as np
import pytest

import oracle as orc
from oracle.streamer import Streamer

pytestmark = pytest.mark.gpu


def rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


# shapes for every path of the batched SVD: one workgroup per matrix (small),
# QR then Jacobi in LDS (tall: m >= 2n, n <= 70, >= 2^15 entries), one launch
# per Jacobi step (large)
@pytest.mark.parametrize("m,n", [(40, 8), (300, 32), (17, 64), (128, 128), (5, 1), (1000, 20), (600, 300),
                                 (2048, 64), (700, 50), (64, 1500), (600, 70), (700, 67)])
def test_svd_batched_matches_numpy(gpu_lib, m, n):
    bf = gpu_lib
    if m >= 2 * n and m * n >= 1 << 15:
        # the tall path (QR + compact WY, Jacobi on the triangle in LDS) takes
        # widths up to 70: the widest ones sit at its LDS limit
        assert bf.lib().bfhipSvdTallMaxN() == 70
    rng = np.random.default_rng(m * 1000 + n)
    mats = []
    for scale in (1.0, 1e-3):
        G = rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))
        U0, s0, V0 = np.linalg.svd(G, full_matrices=False)
        # a graded spectrum with random exponents, sigma_0 = scale
        s0 = scale * np.sort(np.concatenate([[1.0], 10.0 ** (-14 * rng.random(len(s0) - 1))]))[::-1]
        mats.append((U0 * s0) @ V0)
    for tol in (1e-8, 1e-15):
        out = bf.svd_batched(mats, tol)
        for A, (U, W, k) in zip(mats, out):
            s = np.linalg.svd(A, compute_uv=False)
            kk = int(np.sum(s >= tol * s[0]))
            assert k == kk
            assert np.abs(U.conj().T @ U - np.eye(k)).max() < 1e-12
            sk = np.linalg.norm(W, axis=1)
            assert np.allclose(sk, s[:k], rtol=1e-11, atol=1e-14 * s[0])
            err = np.linalg.norm(U @ W - A, 2)
            # rounding of the rotations grows with the number of columns
            assert err <= (s[k] if k < len(s) else 0) * 1.0001 + 1e-13 * max(1, n / 64) * s[0]


@pytest.mark.parametrize("m,n", [(900, 40), (40, 900), (600, 70), (70, 600)])
def test_svd_batched_rank_deficient_with_zero_columns(gpu_lib, m, n):
    """Tall path (as is, or on the conjugate transpose) with exact rank 10
    and five exactly zero columns (rows): the pivoted QR ends with steps that
    have no reflector, which the compact WY step of the tall SVD skips."""
    bf = gpu_lib
    rng = np.random.default_rng(m + n)
    r = 10
    A = (rng.standard_normal((m, r)) + 1j * rng.standard_normal((m, r))) @ \
        (rng.standard_normal((r, n)) + 1j * rng.standard_normal((r, n)))
    zero = rng.choice(min(m, n), 5, replace=False)
    if m >= n:
        A[:, zero] = 0
    else:
        A[zero, :] = 0
    (U, W, k), = bf.svd_batched([A], 1e-8)
    s = np.linalg.svd(A, compute_uv=False)
    assert k == int(np.sum(s >= 1e-8 * s[0])) == r
    assert np.abs(U.conj().T @ U - np.eye(k)).max() < 1e-12
    assert np.allclose(np.linalg.norm(W, axis=1), s[:k], rtol=1e-11)
    assert np.linalg.norm(U @ W - A, 2) <= 1e-13 * s[0]


def _nodes(tree):
    import butterfly_amd as bf
    qs = bf.QuadtreeStruct.from_address(tree.ptr)
    return (bf.QuadtreeNodeStruct * qs.numNodes).from_address(qs.nodes)


def _stream_both(bf, A, src, tgt, col_leaf, row_leaf, **kw):
    ts, tt = bf.Quadtree(src, col_leaf), bf.Quadtree(tgt, row_leaf)
    sp, sn, _ = orc.tree(src, col_leaf)
    tp, tn, _ = orc.tree(tgt, row_leaf)
    assert np.array_equal(sp, ts.perm) and np.array_equal(tp, tt.perm)
    At = A(tgt[tp], src[sp])
    S = bf.FacStreamer(tt, ts, kw["tol"], init_depth=kw["init_depth"], min_rows=kw["min_rows"],
                       min_cols=kw["min_cols"])
    O = Streamer(tn, sn, **kw)
    cn = _nodes(ts)
    feeds = 0
    while not O.done:
        c = O.current_column_node()
        assert S.current_column_node() == c
        a, e = cn[c].offset[0], cn[c].offset[4]
        assert S.feed(At[:, a:e])
        O.feed(At[:, a:e])
        feeds += 1
    assert S.done and feeds > 1
    return S, O.fac(), At


def test_streamer_matches_oracle_on_kernel_matrix(gpu_lib):
    bf = gpu_lib
    rng = np.random.default_rng(4)
    n, tol = 2048, 1e-8
    src, tgt = rng.random((n, 2)), rng.random((n, 2)) + [2, 0]
    S, OF, At = _stream_both(bf, lambda t, s: orc.kernel_matrix(25.0, t, s), src, tgt, 64, 16,
                             tol=tol, init_depth=1, min_rows=20, min_cols=20)
    F = S.fac()
    assert np.array_equal(F.row_nodes(), np.array(OF.rowNodes))
    assert F.num_w == len(OF.W)
    assert F.factor_shape(0)[:2] == OF.Psi.shape
    for i, W in enumerate(OF.W):
        assert F.factor_shape(i + 1)[:2] == W.shape
    M = F.mat()
    x = rng.standard_normal((n, 3)) + 1j * rng.standard_normal((n, 3))
    y = M @ x
    yo = OF.apply(x)
    ye = At @ x
    print("device vs exact", rel(y, ye), "oracle vs exact", rel(yo, ye), "device vs oracle", rel(y, yo))
    assert rel(y, ye) < 20 * tol and rel(yo, ye) < 20 * tol
    assert rel(y, yo) < 20 * tol
    assert F.num_bytes() < At.size * 16 / 5


def test_streamer_exact_for_low_rank_and_partial_span(gpu_lib):
    """A rank-6 matrix streams to an exact factorization; before the last
    feed the span of partial factorizations reproduces the columns fed."""
    bf = gpu_lib
    rng = np.random.default_rng(3)
    n = 1024
    src, tgt = rng.random((n, 2)), rng.random((n, 2)) + [2, 0]
    L = rng.standard_normal((n, 6)) + 1j * rng.standard_normal((n, 6))
    R = rng.standard_normal((6, n)) + 1j * rng.standard_normal((6, n))
    ts, tt = bf.Quadtree(src, 32), bf.Quadtree(tgt, 16)
    A = L[tt.perm] @ R[:, ts.perm]
    S = bf.FacStreamer(tt, ts, 1e-10, init_depth=0, min_rows=20, min_cols=20)
    cn = _nodes(ts)
    fed = 0
    half = None
    while not S.done:
        c = S.current_column_node()
        a, e = cn[c].offset[0], cn[c].offset[4]
        assert S.feed(A[:, a:e])
        fed = e
        if half is None and fed >= n // 2 and not S.done:
            half = fed
            Mp = S.span_mat()
            xp = rng.standard_normal(Mp.shape[1]) + 0j
            assert Mp.shape[1] == half
            assert rel(Mp @ xp, A[:, :half] @ xp) < 1e-12
    F = S.fac()
    assert F.factor_shape(0)[1] == 6
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(F.mat() @ x, A @ x) < 1e-12


def test_streamer_cuts_follow_planted_ranks_on_device(gpu_lib):
    """The planted-rank matrix of the oracle's pin (test_oracle.py): the
    device streamer makes the same cuts -- final row cut [1, 9..20] with
    Psi columns 3 + 16 x 12 -- and reproduces the matrix."""
    bf = gpu_lib
    g = (np.arange(16) + 0.5) / 16
    X, Y = np.meshgrid(g, g)
    pts = np.stack([X.ravel(), Y.ravel()], 1)
    tt, ts = bf.Quadtree(pts, 16), bf.Quadtree(pts + [2.0, 0.0], 64)
    rng = np.random.default_rng(12)

    def cg(m, n):
        return rng.standard_normal((m, n)) + 1j * rng.standard_normal((m, n))

    A = np.zeros((256, 256), complex)
    A[0:64] = cg(64, 3) @ cg(3, 256)
    for j in range(4):
        c = slice(64 * j, 64 * (j + 1))
        A[64:128, c] = cg(64, 3) @ cg(3, 64) if j == 3 else cg(64, 64)
        A[128:192, c] = cg(64, 2) @ cg(2, 64)
        A[192:256, c] = cg(64, 20) @ cg(20, 64)
    S = bf.FacStreamer(tt, ts, 1e-10, init_depth=1, min_rows=1, min_cols=1)
    cn = _nodes(ts)
    while not S.done:
        c = S.current_column_node()
        assert S.feed(A[:, cn[c].offset[0]:cn[c].offset[4]])
    F = S.fac()
    assert list(F.row_nodes()) == [1] + list(range(9, 21))
    assert F.factor_shape(0)[:2] == (256, 3 + 16 * 12)
    x = rng.standard_normal(256) + 1j * rng.standard_normal(256)
    assert rel(F.mat() @ x, A @ x) < 1e-12
