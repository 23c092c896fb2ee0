"""Boundary quadrature for the integral-equation checks, restated with numpy.

TEST INFRASTRUCTURE ONLY: the checker of butterfly_amd's boundary operator
and GMRES (bfhipHelm2BieNew, bfSolveGMRES).  Only tests/ load it.

  - Kapur-Rokhlin corrections of a logarithmic singularity for the
    trapezoid rule on a closed curve, orders 2 and 6: entry (i, i +- (p+1))
    of the kernel matrix gains gamma_p times itself (quadrature.c:12-27,
    :126-166; Kapur & Rokhlin, SIAM J. Numer. Anal. 34 (1997));
  - the trapezoid weights |gamma'(theta_j)| 2 pi / N of the curve
    r(theta) = 1 + 0.3 cos(5 theta);
  - the dense second-kind matrix alpha I + (K + C) W (helm2_bie.c:95-121).

Pinned by the single layer on a circle, whose Fourier modes are exact
eigenfunctions: S e^{i n theta} = (i pi R / 2) J_n(k R) H_n(k R) e^{i n theta}
(tests/test_oracle.py).
"""
import numpy as np

KR = {
    2: np.array([1.825748064736159, -1.325748064736159]),
    6: np.array([4.967362978287758, -16.20501504859126, 25.85153761832639,
                 -22.22599466791883, 9.930104998037539, -1.817995878141594]),
}


def kr_correct(K, order):
    """K + C for a kernel matrix K in boundary order (zero diagonal)."""
    if order == 0:
        return K.copy()
    n = K.shape[0]
    A = K.copy()
    i = np.arange(n)
    for p, g in enumerate(KR[order]):
        for j in ((i + p + 1) % n, (i - p - 1) % n):
            A[i, j] += g * K[i, j]
    return A


def curve_weights(n):
    th = 2 * np.pi * np.arange(n) / n
    r = 1 + 0.3 * np.cos(5 * th)
    dr = -1.5 * np.sin(5 * th)
    return np.sqrt(r * r + dr * dr) * (2 * np.pi / n)


def bie_dense(K, w, order, alpha=0.5):
    return alpha * np.eye(K.shape[0]) + kr_correct(K, order) * w[None, :]
