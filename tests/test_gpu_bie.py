"""Second-kind boundary integral operator and GMRES on the device (GPU).

The operator is A = I/2 + (S + C) W on the curve r = 1 + 0.3 cos(5 theta)
(helm2_bie.c:109-121 with the single layer S): S the device butterfly
factorization of the curve against itself, C the Kapur-Rokhlin corrections,
W the trapezoid weights.  The checker is the dense matrix assembled from
the oracle's kernel entries and its numpy quadrature restatement; the
solves are checked against numpy's dense LU solve, and the unrestarted GMRES
iteration count against scipy's GMRES on the dense matrix (same Krylov
spaces, so the counts agree up to the factorization's perturbation).

Tolerances: the factorization is built at 1e-10, so operator and solution
agree with the dense ones to ~1e-9; the FP32 inner solves of the mixed
solver are refined with FP64 residuals down to the FP64 tolerance.
"""
import numpy as np
import pytest

import oracle as orc
from conftest import curve_problem
from oracle.quadrature import bie_dense, curve_weights

pytestmark = pytest.mark.gpu

N = 2048


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


@pytest.fixture(scope="module")
def system(gpu_lib):
    bf = gpu_lib
    pts, _, k = curve_problem(N)
    tree = bf.Quadtree(pts, 64)
    helm = bf.Helm2(k)
    F = bf.FacHelm2.make_multilevel(helm, tree, tree, bf.options(tol=1e-10, pass_through=True))
    w = curve_weights(N)
    A = bf.BoundaryOperator(F, helm, tree, w, kr_order=6, alpha=0.5)
    K = orc.kernel_matrix(k, pts, pts)
    p = tree.perm
    dense = {order: bie_dense(K, w, order, 0.5)[np.ix_(p, p)] for order in (0, 6)}
    # plane wave in direction 30 degrees, tree order
    d = np.array([np.cos(np.pi / 6), np.sin(np.pi / 6)])
    b = np.exp(1j * k * (pts @ d))[p]
    return dict(bf=bf, F=F, A=A, dense=dense, b=b, tree=tree, helm=helm, w=w, k=k)


def test_boundary_apply_matches_dense(system):
    bf, A, dense = system["bf"], system["A"], system["dense"]
    rng = np.random.default_rng(21)
    x = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    assert rel(A @ x, dense[6] @ x) < 1e-8
    A0 = bf.BoundaryOperator(system["F"], system["helm"], system["tree"], system["w"], kr_order=0, alpha=0.5)
    assert rel(A0 @ x, dense[0] @ x) < 1e-8
    # the corrections are not a no-op at this resolution
    assert rel(dense[6] @ x, dense[0] @ x) > 1e-6
    X = rng.standard_normal((N, 5)) + 1j * rng.standard_normal((N, 5))
    dX = bf.DeviceArray.from_numpy(X)
    dY = bf.DeviceArray(16 * N * 5)
    A.mul_device(dX, dY, 5, "fp64")
    Y = dY.to_numpy(np.complex128, (N, 5))
    for c in range(5):
        assert rel(Y[:, c], A @ X[:, c]) < 1e-12


def test_boundary_operator_rejects_bad_arguments(system):
    bf = system["bf"]
    with pytest.raises(bf.BfError):
        bf.BoundaryOperator(system["F"], system["helm"], system["tree"], system["w"], kr_order=4)
    with pytest.raises(bf.BfError):
        system["A"] @ np.zeros(N + 1, complex)


def test_gmres_solves_the_dense_system(system):
    from scipy.sparse.linalg import gmres
    A, Ad, b = system["A"], system["dense"][6], system["b"]
    tol = 1e-10
    x, it = A.solve(b, tol, 300)
    xd = np.linalg.solve(Ad, b)
    assert rel(x, xd) < 1e-8
    assert np.linalg.norm(Ad @ x - b) / np.linalg.norm(b) < 1e-8
    count = [0]

    def cb(_):
        count[0] += 1

    gmres(Ad, b, rtol=tol, atol=0, restart=300, maxiter=1, callback=cb, callback_type="pr_norm")
    print("GMRES iterations: device", it, "scipy on the dense matrix", count[0])
    assert 0 < it <= 300 and abs(it - count[0]) <= 2


def test_gmres_columns_and_initial_guess(system):
    A, b = system["A"], system["b"]
    rng = np.random.default_rng(22)
    B = np.stack([b, rng.standard_normal(N) + 1j * rng.standard_normal(N)], 1)
    X, it = A.solve(B, 1e-10, 300)
    x0, _ = A.solve(b, 1e-10, 300)
    assert rel(X[:, 0], x0) < 1e-12
    # starting from the solution leaves nothing to do
    X2, it2 = A.solve(B, 1e-6, 300, x0=X)
    assert it2 <= 1 and rel(X2, X) < 1e-6


def test_mixed_gmres_reaches_the_fp64_tolerance(system):
    """FP64 residuals around FP32 inner GMRES(50): the refined solution meets
    the FP64 tolerance although one FP32 pass cannot."""
    F, A, Ad, b = system["F"], system["A"], system["dense"][6], system["b"]
    x64, _ = A.solve(b, 1e-10, 300)
    F.orthogonalize()
    F.prepare_mixed("fp32")
    x, info = A.solve_mixed(b, 1e-10, restart=50, max_iter=1000)
    print("mixed GMRES:", info)
    assert info["converged"] and info["rel_res"] <= 1e-10
    assert info["refinements"] >= 2
    assert rel(x, x64) < 1e-8
    assert np.linalg.norm(Ad @ x - b) / np.linalg.norm(b) < 1e-8
