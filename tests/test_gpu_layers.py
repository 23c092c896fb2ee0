"""The double layer, the combined field and S' (GPU).

Kernels (helm2.c:38-90, :127-175): D = (i/4) k H1(k r) n_s.(t - s)/r with the
source unit normal n_s (0 for r < 1e-15), alpha S + beta D, and
S' = (i/4) k H1(k r) n_t.(t - s)/r with the target unit normal n_t.  The
checker is the oracle's C restatement with glibc's j1/y1 (the reference's
bf_H1): kernel entries to the phase-conditioned bound used for the single
layer, and the butterfly applies against exact FP64 dense rows of the same
layer to the factorization's accuracy.  The proxies of D and of the combined
field are the same layers with the circles' radial normals; those of S' are
single-layer charges, the target normals entering only the kernel blocks
(layer_pot.h:63-69, fac_helm2.c:58-59, :455-465).
"""
import numpy as np
import pytest

import oracle as orc
from conftest import CURVE_L
from oracle.quadrature import kr_correct

pytestmark = pytest.mark.gpu


def OPTS(bf):
    return bf.options(tol=1e-10, rank_const=1.25, pass_through=True, pinv_rcond=1e-11)


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


def starfish(n, ppw=10):
    """r = 1 + 0.3 cos 5t with outward unit normals; targets on r = 3."""
    th = 2 * np.pi * np.arange(n) / n
    r = 1 + 0.3 * np.cos(5 * th)
    dr = -1.5 * np.sin(5 * th)
    pts = np.stack([r * np.cos(th), r * np.sin(th)], 1)
    tx, ty = dr * np.cos(th) - r * np.sin(th), dr * np.sin(th) + r * np.cos(th)
    nrm = np.stack([ty, -tx], 1) / np.hypot(tx, ty)[:, None]
    tgt = np.stack([3 * np.cos(th), 3 * np.sin(th)], 1)
    w = np.hypot(tx, ty) * (2 * np.pi / n)
    return pts, nrm, tgt, w, 2 * np.pi * n / (ppw * CURVE_L)


LAYERS = {"double": "BF_LAYER_POTENTIAL_PV_DOUBLE", "combined": "BF_LAYER_POTENTIAL_COMBINED_FIELD",
          "sprime": "BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE"}


@pytest.mark.parametrize("layer", ["double", "combined", "sprime"])
def test_layer_kernel_matrix_matches_oracle(gpu_lib, layer):
    bf = gpu_lib
    rng = np.random.default_rng(41)
    src = rng.uniform(-1, 1, (300, 2))
    ang = rng.uniform(0, 2 * np.pi, 300)
    nrm = np.stack([np.cos(ang), np.sin(ang)], 1)
    tgt = rng.uniform(-1, 1, (200, 2))
    tgt[:3] = src[:3]  # coincident points: 0 on the diagonal
    tn = nrm[50:250]   # target normals (S')
    k = 37.0
    lp = getattr(bf, LAYERS[layer])
    alpha, beta = 1.0, -0.5j
    if layer == "sprime":
        K = bf.Helm2(k, lp).get_kernel_matrix(src, tgt, None, tn)
        Ko = orc.kernel_matrix_layer(k, tgt, src, None, layer, tgt_normals=tn)
    else:
        K = bf.Helm2(k, lp, alpha, beta).get_kernel_matrix(src, tgt, nrm)
        Ko = orc.kernel_matrix_layer(k, tgt, src, nrm, layer, alpha, beta)
    r = np.hypot(tgt[:, None, 0] - src[None, :, 0], tgt[:, None, 1] - src[None, :, 1])
    bound = 1e-13 * (1 + k * r) * np.abs(Ko).max()
    assert (np.abs(K - Ko) <= bound).all(), np.max(np.abs(K - Ko) / bound)
    assert K[0, 0] == 0 or layer == "combined"


@pytest.mark.parametrize("layer", ["double", "combined", "sprime"])
def test_layer_butterfly_matches_dense_rows(gpu_lib, layer):
    bf = gpu_lib
    n = 4096
    pts, nrm, tgt, _, k = starfish(n)
    tn = tgt / 3.0  # outward normals of the target circle
    lp = getattr(bf, LAYERS[layer])
    alpha, beta = 1.0, -0.5j
    if layer == "sprime":
        ts, tt = bf.Quadtree(pts, 64), bf.Quadtree(tgt, 64, normals=tn)
    else:
        ts, tt = bf.Quadtree(pts, 64, normals=nrm), bf.Quadtree(tgt, 64)
    F = bf.FacHelm2.make_multilevel(bf.Helm2(k, lp, alpha, beta), ts, tt, OPTS(bf))
    rng = np.random.default_rng(42)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    y = np.empty(n, complex)
    y[tt.perm] = F @ x[ts.perm]
    rows = np.arange(0, n, 37)
    if layer == "sprime":
        yd = orc.dense_rows_layer(k, tgt[rows], pts, None, x, layer, tgt_normals=tn[rows])
    else:
        yd = orc.dense_rows_layer(k, tgt[rows], pts, nrm, x, layer, alpha, beta)
    err = rel(y[rows], yd)
    print(layer, "butterfly vs dense rows:", err)
    assert err < 1e-8


def test_double_layer_boundary_equation(gpu_lib):
    """I/2 + D W on the curve against itself (no corrections: D is smooth),
    against the dense assembly; GMRES solves it (79 iterations for scipy's
    GMRES on the dense matrix at this size)."""
    bf = gpu_lib
    n = 1024
    pts, nrm, _, w, k = starfish(n, ppw=20)
    tree = bf.Quadtree(pts, 64, normals=nrm)
    helm = bf.Helm2(k, bf.BF_LAYER_POTENTIAL_PV_DOUBLE)
    F = bf.FacHelm2.make_multilevel(helm, tree, tree, OPTS(bf))
    A = bf.BoundaryOperator(F, helm, tree, w, kr_order=0, alpha=0.5)
    p = tree.perm
    D = orc.kernel_matrix_layer(k, pts, pts, nrm, "double")
    Ad = (0.5 * np.eye(n) + D * w[None, :])[np.ix_(p, p)]
    rng = np.random.default_rng(43)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(A @ x, Ad @ x) < 1e-8
    d = np.array([np.cos(np.pi / 6), np.sin(np.pi / 6)])
    b = np.exp(1j * k * (pts @ d))[p]
    xs, it = A.solve(b, 1e-10, 300)
    print("double-layer GMRES iterations:", it)
    assert rel(xs, np.linalg.solve(Ad, b)) < 1e-7 and abs(it - 79) <= 3
    with pytest.raises(bf.BfError):  # D needs no logarithmic corrections
        bf.BoundaryOperator(F, helm, tree, w, kr_order=6, alpha=0.5)


def test_combined_field_boundary_operator(gpu_lib):
    """I/2 + (alpha S + beta D) W + alpha C W: the Kapur-Rokhlin corrections
    C of the single-layer part are scaled by alpha."""
    bf = gpu_lib
    n = 2048
    pts, nrm, _, w, k = starfish(n)
    alpha, beta = -1j * k, 1.0
    tree = bf.Quadtree(pts, 64, normals=nrm)
    helm = bf.Helm2(k, bf.BF_LAYER_POTENTIAL_COMBINED_FIELD, alpha, beta)
    F = bf.FacHelm2.make_multilevel(helm, tree, tree, OPTS(bf))
    A = bf.BoundaryOperator(F, helm, tree, w, kr_order=6, alpha=0.5)
    p = tree.perm
    S = orc.kernel_matrix(k, pts, pts)
    CF = orc.kernel_matrix_layer(k, pts, pts, nrm, "combined", alpha, beta)
    Ad = (0.5 * np.eye(n) + (CF + alpha * (kr_correct(S, 6) - S)) * w[None, :])[np.ix_(p, p)]
    rng = np.random.default_rng(44)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(A @ x, Ad @ x) < 1e-8


def test_sprime_reference_rules_match_dense_rows(gpu_lib):
    """S' under the reference rules (leaf 1, ranks at 1e-15, dense below
    128^2) on the config-0 geometry with random target normals: the whole
    apply against exact FP64 dense rows of S'."""
    bf = gpu_lib
    from conftest import box_problem
    src, tgt, k = box_problem(2048)
    rng = np.random.default_rng(45)
    ang = rng.uniform(0, 2 * np.pi, len(tgt))
    tn = np.stack([np.cos(ang), np.sin(ang)], 1)
    ts, tt = bf.Quadtree(src, 1), bf.Quadtree(tgt, 1, normals=tn)
    F = bf.FacHelm2.make_multilevel(bf.Helm2(k, bf.BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE), ts, tt)
    x = rng.standard_normal(len(src)) + 1j * rng.standard_normal(len(src))
    y = np.empty(len(tgt), complex)
    y[tt.perm] = F @ x[ts.perm]
    yd = orc.dense_rows_layer(k, tgt, src, None, x, "sprime", tgt_normals=tn)
    err = rel(y, yd)
    print("S' reference rules vs dense rows:", err)
    assert err < 1e-10


def test_sprime_boundary_equation(gpu_lib):
    """I/2 + (S' + C) W, the reference's second-kind system
    (helm2_bie.c:25, :95-121): S' on the curve against itself with its
    outward normals, order-6 Kapur-Rokhlin corrections made of S' values
    (krComplexKernel, helm2.c:372-378), against the dense assembly from the
    oracle; GMRES against numpy's solve of that assembly."""
    bf = gpu_lib
    n = 2048
    pts, nrm, _, w, k = starfish(n)
    tree = bf.Quadtree(pts, 64, normals=nrm)
    helm = bf.Helm2(k, bf.BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE)
    F = bf.FacHelm2.make_multilevel(helm, tree, tree, OPTS(bf))
    A = bf.BoundaryOperator(F, helm, tree, w, kr_order=6, alpha=0.5)
    p = tree.perm
    Sp = orc.kernel_matrix_layer(k, pts, pts, None, "sprime", tgt_normals=nrm)
    Ad = (0.5 * np.eye(n) + kr_correct(Sp, 6) * w[None, :])[np.ix_(p, p)]
    rng = np.random.default_rng(46)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    err = rel(A @ x, Ad @ x)
    print("S' boundary operator vs dense assembly:", err)
    assert err < 1e-8
    d = np.array([np.cos(np.pi / 6), np.sin(np.pi / 6)])
    b = (1j * k * (nrm @ d) * np.exp(1j * k * (pts @ d)))[p]
    xs, it = A.solve(b, 1e-10, 300)
    print("S' GMRES iterations:", it)
    assert rel(xs, np.linalg.solve(Ad, b)) < 1e-7


def test_missing_normals_and_the_hypersingular_layer_are_refused(gpu_lib):
    bf = gpu_lib
    pts, nrm, tgt, _, k = starfish(1024)
    ts, tt = bf.Quadtree(pts, 64), bf.Quadtree(tgt, 64)
    with pytest.raises(bf.BfError):  # no source normals
        bf.FacHelm2.make_multilevel(bf.Helm2(k, bf.BF_LAYER_POTENTIAL_PV_DOUBLE), ts, tt, bf.options(pass_through=True))
    tsn = bf.Quadtree(pts, 64, normals=nrm)
    with pytest.raises(bf.BfError):  # S' without target normals
        bf.FacHelm2.make_multilevel(bf.Helm2(k, bf.BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE), tsn, tt,
                                    bf.options(pass_through=True))
    ttn = bf.Quadtree(tgt, 64, normals=tgt / 3.0)
    with pytest.raises(bf.BfError):  # the normal derivative of D (not sampled by the reference either)
        bf.FacHelm2.make_multilevel(bf.Helm2(k, 4), tsn, ttn, bf.options(pass_through=True))
