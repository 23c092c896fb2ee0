"""The oracle pinned against known answers, closed forms, mpmath and numpy
(CPU only)."""
import json
import os

import mpmath as mp
import numpy as np
import pytest

import oracle as orc
from conftest import box_problem, curve_problem

HERE = os.path.dirname(os.path.abspath(__file__))


def test_splitmix64_known_answers():
    # published first outputs of splitmix64 seeded with 0
    L = orc.lib()
    L.orc_splitmix_seed(0)
    got = [L.orc_splitmix_draw() for _ in range(3)]
    assert got == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]


def test_xoshiro256plus_against_textbook_recurrence():
    # an independent restatement of xoshiro256+ 1.0 (Blackman & Vigna 2018)
    L = orc.lib()
    L.orc_splitmix_seed(7)
    s = [L.orc_splitmix_draw() for _ in range(4)]
    M = (1 << 64) - 1
    L.orc_seed(7)
    for _ in range(100):
        want = (s[0] + s[3]) & M
        t = (s[1] << 17) & M
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & M
        assert L.orc_next_u64() == want


def test_uniform_and_randn_statistics():
    orc.seed(0)
    u = np.array([orc.lib().orc_uniform() for _ in range(20000)])
    assert u.min() >= 0 and u.max() < 1
    assert abs(u.mean() - 0.5) < 0.01
    z = orc.complex_randn(50000)
    assert abs(z.real.mean()) < 0.02 and abs(z.imag.std() - 1) < 0.02
    # Box-Muller pairs: |z|^2 = -2 log u0 is exponential with mean 2
    assert abs(np.mean(np.abs(z) ** 2) - 2) < 0.05


def _xoshiro_uniforms(seed, count):
    """(x >> 11) 2^-53 of xoshiro256+ seeded through splitmix64, restated in
    Python from the published generators (rand.c:19-46)."""
    M = (1 << 64) - 1
    z = seed & M
    s = []
    for _ in range(4):
        z = (z + 0x9E3779B97F4A7C15) & M
        w = z
        w = ((w ^ (w >> 30)) * 0xBF58476D1CE4E5B9) & M
        w = ((w ^ (w >> 27)) * 0x94D049BB133111EB) & M
        s.append(w ^ (w >> 31))
    out = []
    for _ in range(count):
        out.append(((s[0] + s[3]) & M) >> 11)
        t = (s[1] << 17) & M
        s[2] ^= s[0]; s[3] ^= s[1]; s[1] ^= s[2]; s[0] ^= s[3]; s[2] ^= t
        s[3] = ((s[3] << 45) | (s[3] >> 19)) & M
    return [u * 2.0 ** -53 for u in out]


@pytest.mark.parametrize("n", [1, 2, 7, 10])
def test_box_muller_pairing_is_the_textbook_transform(n):
    """bfRealRandn (rand.c:52-72): consecutive uniforms (u0, u1) become
    (sqrt(-2 log u0) cos 2 pi u1, sqrt(-2 log u0) sin 2 pi u1); an odd
    length draws two more uniforms for the cosine deviate alone.  Swapping
    cos and sin, or the roles of u0 and u1, fails this bit for bit."""
    import math
    u = _xoshiro_uniforms(3, n + 2)
    want = []
    for i in range(n // 2):
        rad = math.sqrt(-2 * math.log(u[2 * i]))
        want += [rad * math.cos(2 * math.pi * u[2 * i + 1]), rad * math.sin(2 * math.pi * u[2 * i + 1])]
    if n % 2:
        j = 2 * (n // 2)
        want.append(math.sqrt(-2 * math.log(u[j])) * math.cos(2 * math.pi * u[j + 1]))
    orc.seed(3)
    got = np.zeros(n)
    orc.lib().orc_randn_fill(n, orc._dp(got))
    assert got.tolist() == want
    # bfComplexRandn is the real stream of length 2n read as (re, im) pairs
    orc.seed(3)
    z = orc.complex_randn(n // 2 + 1)
    orc.seed(3)
    r = np.zeros(2 * (n // 2 + 1))
    orc.lib().orc_randn_fill(len(r), orc._dp(r))
    assert np.array_equal(z.view(np.float64), r)


def test_box_sampling_draws_x_then_y():
    """bfPoint2SampleUniformlyFromBoundingBox (points.c:25-28): point i takes
    uniform 2i for x and 2i + 1 for y, each scaled to its side."""
    u = _xoshiro_uniforms(11, 8)
    orc.seed(11)
    P = orc.sample_box(4, 2.0, 3.0, -1.0, 0.5)
    for i in range(4):
        assert P[i, 0] == (3.0 - 2.0) * u[2 * i] + 2.0
        assert P[i, 1] == (0.5 - -1.0) * u[2 * i + 1] + -1.0


def test_h0_matches_mpmath_table():
    tab = json.load(open(os.path.join(HERE, "golden", "h0_table.json")))["rows"]
    x = np.array([float(r[0]) for r in tab])
    want = np.array([complex(float(r[1]), float(r[2])) for r in tab])
    got = orc.h0(x)
    err = np.abs(got - want) / np.abs(want)
    # glibc j0/y0: a few ulps, plus the phase conditioning x*eps at large x
    assert np.all(err <= 1e-14 + 4e-16 * x), (x[np.argmax(err)], err.max())


def test_h1_matches_mpmath_table():
    tab = json.load(open(os.path.join(HERE, "golden", "h1_table.json")))["rows"]
    x = np.array([float(r[0]) for r in tab])
    want = np.array([complex(float(r[1]), float(r[2])) for r in tab])
    err = np.abs(orc.h1(x) - want) / np.abs(want)
    assert np.all(err <= 1e-14 + 4e-16 * x), (x[np.argmax(err)], err.max())


def test_double_layer_is_the_source_normal_derivative():
    """D(t, s) = n_s . grad_s G(t, s): a central difference of the oracle's
    single-layer entries along the source normal (a sign or direction slip
    in n_s.(t - s) flips it); the combined field is alpha S + beta D; r = 0
    gives 0."""
    rng = np.random.default_rng(5)
    k = 11.0
    src = rng.uniform(-1, 1, (40, 2))
    ang = rng.uniform(0, 2 * np.pi, 40)
    nrm = np.stack([np.cos(ang), np.sin(ang)], 1)
    tgt = rng.uniform(2, 3, (30, 2))
    h = 1e-5
    fd = (orc.kernel_matrix(k, tgt, src + h * nrm) - orc.kernel_matrix(k, tgt, src - h * nrm)) / (2 * h)
    D = orc.kernel_matrix_layer(k, tgt, src, nrm, "double")
    assert np.abs(D - fd).max() < 1e-7 * np.abs(D).max()
    S = orc.kernel_matrix(k, tgt, src)
    a, b = 0.3 - 2j, -1.5 + 0.25j
    CF = orc.kernel_matrix_layer(k, tgt, src, nrm, "combined", a, b)
    assert np.abs(CF - (a * S + b * D)).max() < 1e-14 * np.abs(CF).max()
    Z = orc.kernel_matrix_layer(k, src[:5], src[:5], nrm[:5], "double")
    assert np.all(np.diag(Z) == 0)
    # dense rows (Neumaier-summed) agree with the matrix product
    x = rng.standard_normal(40) + 1j * rng.standard_normal(40)
    y = orc.dense_rows_layer(k, tgt, src, nrm, x, "combined", a, b)
    assert np.abs(y - CF @ x).max() < 1e-13 * np.abs(y).max()


def test_sprime_is_minus_the_target_normal_derivative():
    """S'(t, s) = (i/4) k H1(k r) n_t.(t - s)/r (helm2.c:45-53, :127-175).
    Since d/dx H0 = -H1, this is -n_t . grad_t G(t, s): a central difference
    of the oracle's single-layer entries along the target normal, with the
    sign the reference's formula carries (a sign slip, or the source normal
    used in place of the target's, fails it); r = 0 gives 0."""
    rng = np.random.default_rng(6)
    k = 9.0
    src = rng.uniform(-1, 1, (35, 2))
    tgt = rng.uniform(2, 3, (25, 2))
    ang = rng.uniform(0, 2 * np.pi, 25)
    tn = np.stack([np.cos(ang), np.sin(ang)], 1)
    sn = np.stack([np.cos(ang[:1] + 1.0), np.sin(ang[:1] + 1.0)], 1).repeat(35, 0)
    h = 1e-5
    fd = (orc.kernel_matrix(k, tgt + h * tn, src) - orc.kernel_matrix(k, tgt - h * tn, src)) / (2 * h)
    Sp = orc.kernel_matrix_layer(k, tgt, src, None, "sprime", tgt_normals=tn)
    assert np.abs(Sp + fd).max() < 1e-7 * np.abs(Sp).max()
    # the source normals play no part in S'
    assert np.array_equal(Sp, orc.kernel_matrix_layer(k, tgt, src, sn, "sprime", tgt_normals=tn))
    # S'(t, s) with n_t is D(s, t) with n_s := n_t and the roles swapped, up to the sign of t - s
    D = orc.kernel_matrix_layer(k, src, tgt, tn, "double")
    assert np.abs(Sp + D.T).max() < 1e-15 * np.abs(Sp).max()
    Z = orc.kernel_matrix_layer(k, tgt[:4], tgt[:4], None, "sprime", tgt_normals=tn[:4])
    assert np.all(np.diag(Z) == 0)
    x = rng.standard_normal(35) + 1j * rng.standard_normal(35)
    y = orc.dense_rows_layer(k, tgt, src, None, x, "sprime", tgt_normals=tn)
    assert np.abs(y - Sp @ x).max() < 1e-13 * np.abs(y).max()


def test_kernel_matrix_closed_form_and_zero_distance():
    k = 7.5
    tgt = np.array([[0.0, 0.0], [1.0, 2.0], [0.5, 0.5]])
    src = np.array([[0.0, 0.0], [3.0, -1.0]])
    K = orc.kernel_matrix(k, tgt, src)
    assert K[0, 0] == 0  # r == 0 gives 0 (helm2.c:114)
    for i in range(3):
        for j in range(2):
            r = float(np.hypot(*(src[j] - tgt[i])))
            if r == 0:
                continue
            want = complex(1j * mp.hankel1(0, k * r) / 4)
            assert abs(K[i, j] - want) < 1e-14 * abs(want)


def test_rank_estimate_closed_form():
    L = orc.lib()
    # p = ceil(k r1 r2 / (R - r1 - r2) - log10(eps))
    for k, r1, r2, R, eps in [(100.0, 0.3, 0.2, 1.5, 1e-15), (73064.8, 0.01, 0.02, 0.5, 1e-8)]:
        want = int(np.ceil(k * r1 * r2 / (R - r1 - r2) - np.log10(eps)))
        assert L.orc_rank_est(k, r1, 0.0, 0.0, r2, R, 0.0, eps) == want


def test_circle_points_closed_form():
    P = orc.circle_points(0.25, -1.0, 0.5, 7)
    th = 2 * np.pi / 7 * np.arange(7)
    np.testing.assert_allclose(P, np.stack([0.5 * np.cos(th) + 0.25, 0.5 * np.sin(th) - 1.0], 1), atol=1e-15)


@pytest.mark.parametrize("p,n", [(8, 5), (16, 40), (33, 20)])
def test_lstsq_well_conditioned_equals_numpy_solve(p, n):
    rng = np.random.default_rng(p)
    A = rng.standard_normal((p, p)) + 1j * rng.standard_normal((p, p))
    B = rng.standard_normal((p, n)) + 1j * rng.standard_normal((p, n))
    X, kept = orc.lstsq(A, B)
    assert kept == p
    np.testing.assert_allclose(X, np.linalg.solve(A, B), rtol=0, atol=1e-10 * np.abs(np.linalg.solve(A, B)).max())


def test_lstsq_truncation_matches_numpy_pinv_rule():
    # an exactly rank-deficient A: pinv keeps the 10 nonzero singular values
    rng = np.random.default_rng(3)
    U, _ = np.linalg.qr(rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20)))
    V, _ = np.linalg.qr(rng.standard_normal((20, 20)) + 1j * rng.standard_normal((20, 20)))
    s = np.r_[np.geomspace(1, 1e-6, 10), np.zeros(10)]
    A = (U * s) @ V.conj().T
    B = rng.standard_normal((20, 4)) + 1j * rng.standard_normal((20, 4))
    X, kept = orc.lstsq(A, B)
    assert kept == 10
    # numpy's pinv with rcond = max(m,n) eps (sigma_0 = 1 here, atol eps)
    Xn = np.linalg.pinv(A, rcond=20 * np.finfo(float).eps) @ B
    np.testing.assert_allclose(X, Xn, atol=1e-6)


def test_svd_values_match_numpy():
    rng = np.random.default_rng(5)
    A = rng.standard_normal((12, 12)) + 1j * rng.standard_normal((12, 12))
    np.testing.assert_allclose(orc.svd_values(A), np.linalg.svd(A, compute_uv=False), rtol=1e-13)


def test_quadtree_invariants():
    src, _, _ = box_problem(3000)
    for leaf in (1, 16):
        perm, nodes, bbox = orc.tree(src, leaf)
        assert np.array_equal(np.sort(perm), np.arange(len(src)))
        P = src[perm]
        for i, (depth, i0, i1, parent, *kids) in enumerate(nodes):
            pts = P[i0:i1]
            assert np.all(pts >= bbox[i, :2]) and np.all(pts <= bbox[i, 2:])
            live = [c for c in kids if c >= 0]
            if live:
                assert nodes[live[0]][1] == i0 and nodes[live[-1]][2] == i1
                assert sum(nodes[c][2] - nodes[c][1] for c in live) == i1 - i0
            else:
                assert i1 - i0 <= leaf or depth == 0
            # Z-order: quadrant q = 2 [x > sx] + [y > sy], ties go low
            for q, c in enumerate(kids):
                if c < 0:
                    continue
                sx, sy = (bbox[i, 0] + bbox[i, 2]) / 2, (bbox[i, 1] + bbox[i, 3]) / 2
                cp = P[nodes[c][1]:nodes[c][2]]
                assert np.all((cp[:, 0] > sx) == bool(q >> 1)) and np.all((cp[:, 1] > sy) == bool(q & 1))
        # root box is the bounding box stretched to a square
        assert abs((bbox[0, 2] - bbox[0, 0]) - (bbox[0, 3] - bbox[0, 1])) < 1e-15


def test_oracle_apply_equals_dense_rows_reference_rules():
    src, tgt, k = box_problem(2048)
    F = orc.Fac(k, src, tgt)
    sp, tp = F.perms()
    orc.seed(1)
    x = orc.complex_randn(len(src))
    y = np.empty(len(tgt), complex)
    y[tp] = F.apply(x[sp])[:, 0]
    rows = np.arange(0, len(tgt), 37)
    yd = orc.dense_rows(k, tgt[rows], src, x)
    assert np.linalg.norm(y[rows] - yd) / np.linalg.norm(yd) < 1e-11


def test_oracle_butterfly_leaf_reproduces_its_kernel_block():
    # densify one butterfly by applying it to unit vectors and compare with
    # the exact kernel block between its target and source nodes
    src, tgt, k = box_problem(4096)
    F = orc.Fac(k, src, tgt)
    lt, ft, bt = F.plan()
    assert lt[:, 0].all() and lt[0, 5] >= 3
    sp, tp = F.perms()
    row0, col0, m, n = lt[0, 1:5]
    cols = np.arange(col0, col0 + n, max(1, n // 12))
    X = np.zeros((len(src), len(cols)), complex)
    X[cols, np.arange(len(cols))] = 1
    Y = F.apply(X)[row0:row0 + m]
    K = orc.kernel_matrix(k, tgt[tp[row0:row0 + m]], src[sp[cols]])
    assert np.linalg.norm(Y - K) / np.linalg.norm(K) < 1e-11


def test_oracle_pass_through_curve_accuracy():
    src, tgt, k = curve_problem(4096)
    F = orc.Fac(k, src, tgt, leaf_size=64, tol=1e-10, pass_through=True)
    sp, tp = F.perms()
    orc.seed(2)
    x = orc.complex_randn(len(src))
    y = np.empty(len(tgt), complex)
    y[tp] = F.apply(x[sp])[:, 0]
    rows = np.arange(0, len(tgt), 61)
    yd = orc.dense_rows(k, tgt[rows], src, x)
    assert np.linalg.norm(y[rows] - yd) / np.linalg.norm(yd) < 1e-6


# ------------------------------------------------ streaming factorization

def _stream(A, rows_nodes, cols_nodes, **kw):
    from oracle.streamer import Streamer
    S = Streamer(rows_nodes, cols_nodes, **kw)
    while not S.done:
        c = S.current_column_node()
        assert S.ct.leaf(c)
        S.feed(A[:, S.ct.start(c):S.ct.end(c)])
    return S.fac()


def test_streamer_truncation_rule_closed_form():
    """linalg.c:26-35: keep sigma_k >= tol sigma_0; a diagonal matrix has
    its diagonal as singular values, so k is a count."""
    from oracle.streamer import tsvd
    s = np.array([4.0, 1.0, 3e-3, 4e-3 * 1e-3, 1e-9])
    A = np.diag(s).astype(complex)
    for tol, k in [(1e-2, 2), (7.5e-4, 3), (1e-6, 4), (2.4e-10, 5), (1e-12, 5)]:
        U, W, kk, truncated = tsvd(A, tol)
        assert kk == k and truncated == (k < 5)
        assert np.allclose(U @ W, np.diag(np.where(np.arange(5) < k, s, 0)), atol=1e-15)


def test_streamer_exact_for_a_low_rank_matrix():
    """A rank-6 matrix: every compressed row block keeps at most 6 terms and
    the product reproduces the matrix to rounding."""
    rng = np.random.default_rng(3)
    n = 1024
    src, tgt = rng.random((n, 2)), rng.random((n, 2)) + [2, 0]
    sp, sn, _ = orc.tree(src, 32)
    tp, tn, _ = orc.tree(tgt, 16)
    A = (rng.standard_normal((n, 6)) + 1j * rng.standard_normal((n, 6))) @ \
        (rng.standard_normal((6, n)) + 1j * rng.standard_normal((6, n)))
    F = _stream(A, tn, sn, tol=1e-10, init_depth=0, min_rows=20, min_cols=20)
    assert F.Psi.shape[1] == 6
    assert np.linalg.norm(F.dense() - A) / np.linalg.norm(A) < 1e-13


def test_streamer_cuts_follow_planted_ranks():
    """The row cuts of the streamer (fac_streamer.c:316-447, fac.c:509-573,
    :867-1050) on a matrix whose block ranks are planted, so the cuts follow
    from the ranks alone.  Rows: a 16 x 16 grid (quadtree leaf 16: root, four
    64-row nodes 1-4, sixteen 16-row leaves 5-20); columns: another grid,
    four 64-column leaves.  Row node 1 has rank 3 over all columns; node 2 is
    random on column leaves 1-3 and of rank 3 on leaf 4; node 3 has rank 2
    on each column leaf (independent factors); node 4 rank 20 on each.

    Feeding a leaf (from depth 1): a node whose block has rank below its
    sides keeps that many terms; a full-rank node splits; a full-rank leaf
    passes its rows through.  Leaves 1-3 cut [1, 9, 10, 11, 12, 3, 4] with
    3, 16 x 4, 2 and 20 columns of Psi; leaf 4 cuts [1, 2, 3, 4].  Merging
    the four: the coarsest common cut takes node 2 from leaf 4 over 9-12 from
    the others; a merged block compresses only if it drops terms and keeps
    fewer terms than rows, so node 1 keeps 3 terms, node 2 (rank 64 of 64
    rows) and node 4 (rank 64 of 80 columns, 64 rows) split into leaves that
    pass through, and node 3 (64 x 8 of full column rank: nothing dropped)
    splits too.  Final cut [1, 9..20], Psi columns 3 + 16 x 12."""
    from oracle.streamer import Streamer, eps_cut
    g = (np.arange(16) + 0.5) / 16
    X, Y = np.meshgrid(g, g)
    rows_pts = np.stack([X.ravel(), Y.ravel()], 1)
    _, tn, _ = orc.tree(rows_pts, 16)
    _, sn, _ = orc.tree(rows_pts + [2.0, 0.0], 64)
    assert [tuple(tn[i, 1:3]) for i in range(1, 5)] == [(0, 64), (64, 128), (128, 192), (192, 256)]
    assert all(tn[i, 4] < 0 for i in range(5, 21)) and len(sn) == 5
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
    S = Streamer(tn, sn, 1e-10, init_depth=1, min_rows=1, min_cols=1)
    order = []
    for step in range(3):
        c = S.current_column_node()
        order.append(c)
        S.feed(A[:, S.ct.start(c):S.ct.end(c)])
        f = S.facs[c]
        assert f.rowNodes == [1, 9, 10, 11, 12, 3, 4], f.rowNodes
        assert [e - a for a, e in f.colRanges] == [3, 16, 16, 16, 16, 2, 20]
    c = S.current_column_node()
    last = eps_cut(S.rt, S.init, 0, A[:, S.ct.start(c):S.ct.end(c)], S.spec, False)
    assert [nd for nd, _, _ in last] == [1, 2, 3, 4] and [P.shape[1] for _, P, _ in last] == [3, 3, 2, 20]
    S.feed(A[:, S.ct.start(c):S.ct.end(c)])
    assert S.done and sorted(order + [c]) == [1, 2, 3, 4]
    F = S.fac()
    assert F.rowNodes == [1] + list(range(9, 21)), F.rowNodes
    assert [e - a for a, e in F.colRanges] == [3] + [16] * 12
    assert np.linalg.norm(F.dense() - A) / np.linalg.norm(A) < 1e-12


def test_streamer_reproduces_a_kernel_matrix_to_its_tolerance():
    """Helmholtz kernel between separated boxes (k = 25): the streamed
    product is within a small multiple of tol of the matrix and stores far
    fewer entries; every SVD-made Psi block has orthonormal columns."""
    rng = np.random.default_rng(4)
    n, tol = 1024, 1e-8
    src, tgt = rng.random((n, 2)), rng.random((n, 2)) + [2, 0]
    sp, sn, _ = orc.tree(src, 64)
    tp, tn, _ = orc.tree(tgt, 16)
    A = orc.kernel_matrix(25.0, tgt[tp], src[sp])
    F = _stream(A, tn, sn, tol=tol, init_depth=1, min_rows=20, min_cols=20)
    assert np.linalg.norm(F.dense() - A) / np.linalg.norm(A) < 20 * tol
    assert F.entries() < A.size / 5
    Psi = F.Psi.toarray()
    for (a, e) in F.colRanges:
        blk = Psi[:, a:e]
        G = blk.conj().T @ blk
        if not np.allclose(np.abs(blk[np.abs(blk) > 0]), 1.0):  # identity blocks are trivially fine
            assert np.abs(G - np.eye(e - a)).max() < 1e-12


def test_field_form_is_the_same_operator():
    """Field-sample intermediates re-associate the charge-form product
    E Z_L^+ T_L ... T_1 Z_0^+ O_0 as (E Z_L^+)(T_L Z_{L-1}^+)...O_0: the two
    oracle factorizations must apply identically up to rounding."""
    from conftest import curve_problem
    src, tgt, k = curve_problem(2048)
    a = orc.Fac(k, src, tgt, 64, 1e-8, 128 * 128, True, rank_const=1.25, pinv_rcond=1e-11)
    b = orc.Fac(k, src, tgt, 64, 1e-8, 128 * 128, True, rank_const=1.25, pinv_rcond=1e-11, field_form=True)
    orc.seed(5)
    x = orc.complex_randn(len(src))
    ya, yb = a.apply(x)[:, 0], b.apply(x)[:, 0]
    assert np.linalg.norm(ya - yb) / np.linalg.norm(ya) < 1e-12
    kinds_b = b.plan()[2][:, 4]
    assert (kinds_b == 3).any() and not (kinds_b == 0).any()  # field blocks, no charge re-expansions


def test_oversampled_test_circles_are_at_least_as_accurate():
    """Least squares on 2p test points instead of p (same ranks): the
    operator stays within the factorization's accuracy of the exact rows and
    is no worse than the square (reference) form."""
    from conftest import curve_problem
    src, tgt, k = curve_problem(2048)
    orc.seed(6)
    x = orc.complex_randn(len(src))
    rows = np.arange(0, len(tgt), 17)
    errs = []
    for os_ in (1.0, 2.0):
        orc.lib().orc_set_test_oversample(os_)
        try:
            F = orc.Fac(k, src, tgt, 64, 1e-7, 128 * 128, True, rank_const=1.25, pinv_rcond=1e-9)
        finally:
            orc.lib().orc_set_test_oversample(1.0)
        sp, tp = F.perms()
        y = np.empty(len(tgt), complex)
        y[tp] = F.apply(x[sp])[:, 0]
        yd = orc.dense_rows(k, tgt[rows], src, x)
        errs.append(np.linalg.norm(y[rows] - yd) / np.linalg.norm(yd))
    assert errs[0] < 1e-5 and errs[1] <= errs[0] * 1.5


def test_kapur_rokhlin_single_layer_on_a_circle():
    """Fourier modes are eigenfunctions of the single layer on a circle:
    S e^{in t} = (i pi R/2) J_n(kR) H_n(kR) e^{in t}.  The trapezoid rule
    with the zero diagonal converges at first order; Kapur-Rokhlin order 2
    and order 6 must show their orders when N doubles (a wrong or missing
    gamma drops the rate to first order)."""
    from scipy.special import hankel1, jv
    from oracle.quadrature import kr_correct
    R, k, n_mode = 1.3, 7.0, 3
    lam = 1j * np.pi * R / 2 * jv(n_mode, k * R) * hankel1(n_mode, k * R)
    errs = {}
    for N in (256, 512):
        th = 2 * np.pi * np.arange(N) / N
        pts = np.stack([R * np.cos(th), R * np.sin(th)], 1)
        K = orc.kernel_matrix(k, pts, pts)
        sigma = np.exp(1j * n_mode * th)
        w = np.full(N, 2 * np.pi * R / N)
        for order in (0, 2, 6):
            y = kr_correct(K, order) @ (w * sigma)
            errs[order, N] = np.linalg.norm(y - lam * sigma) / np.linalg.norm(lam * sigma)
    rate = {o: np.log2(errs[o, 256] / errs[o, 512]) for o in (0, 2, 6)}
    assert 0.5 < rate[0] < 1.5, rate
    assert 2.0 < rate[2] < 3.5, rate
    assert rate[6] > 5.5, rate
    assert errs[6, 512] < 1e-6, errs


def test_rank_mlp_restatement_gradient_by_central_differences():
    """The float64 rank-model restatement: its analytic gradient against
    central differences of its own loss on coordinates of every layer, and
    the closed form of the all-zero network (output = b4, GELU(0) = 0)."""
    from oracle import rank_mlp as R
    rng = np.random.default_rng(7)
    p = 0.3 * rng.standard_normal(R.NPARAM)
    X = rng.standard_normal((40, 6))
    y = rng.standard_normal(40)
    _, g = R.loss_grad(p, X, y)
    offs = np.cumsum([0] + [int(np.prod(s)) for s in R.SHAPES])
    h = 1e-6
    for lo, hi in zip(offs[:-1], offs[1:]):
        for i in rng.choice(np.arange(lo, hi), size=min(6, hi - lo), replace=False):
            e = np.zeros(R.NPARAM)
            e[i] = h
            fd = (R.loss(p + e, X, y) - R.loss(p - e, X, y)) / (2 * h)
            assert abs(fd - g[i]) <= 1e-7 + 1e-5 * abs(fd), (i, fd, g[i])
    z = np.zeros(R.NPARAM)
    z[-1] = 0.7
    l, g0 = R.loss_grad(z, X, y)
    assert np.isclose(l, np.mean((0.7 - y) ** 2)) and np.isclose(g0[-1], 2 * np.mean(0.7 - y))
    assert np.all(g0[:-1] == 0)
