/*
 * Hankel functions H0^(1)(x) = J0(x) + i Y0(x) and H1^(1)(x) = J1 + i Y1
 * in FP64 for x > 0, the 2-D Helmholtz single-layer kernel
 * G = (i/4) H0^(1)(k r) and the double-layer kernel (i k / 4) H1(k r) n.(t-s)/r.
 *
 * Written for the device (it is also compiled for the host so the CPU
 * tests can check it against mpmath and glibc).  The reference takes H0
 * from glibc's j0/y0 (bessel.c:284-296) and forms G in helm2.c:93-125.
 *
 * Three branches, chosen per lane; a wave whose arguments fall in one
 * branch runs only that branch:
 *   x <= 4        power series of J0 and Y0 (cancellation stays below 4x)
 *   4 < x < 24    P, Q of the modulus/phase form from Chebyshev expansions
 *                 in 1/x
 *   x >= 24       P, Q from the Hankel asymptotic series (20 terms each)
 * with H0 = sqrt(2/(pi x)) (P + iQ) exp(i(x - pi/4)) above 4.  The phase
 * is reduced by a three-part Cody-Waite split of pi/2, so it stays exact
 * to a few ulps for the x ~ 3e5 arguments of the large problems.
 * Coefficients come from gen_h0_coeffs.py (mpmath, 50 digits).
 */
#pragma once

#include <math.h>

#include "h0_coeffs.h"

#if defined(__HIPCC__)
#define BF_HD __host__ __device__
#else
#define BF_HD
#endif

BF_HD static inline void bf_h0_series(double x, double *jr, double *yr) {
  double z = 0.25 * x * x;
  double j = bf_h0_ser_j[BF_H0_NSER - 1];
  double s = bf_h0_ser_y[BF_H0_NSER - 1];
  for (int k = BF_H0_NSER - 2; k >= 0; --k) {
    j = fma(j, z, bf_h0_ser_j[k]);
    s = fma(s, z, bf_h0_ser_y[k]);
  }
  *jr = j;
  *yr = BF_TWO_OVER_PI * ((log(0.5 * x) + BF_EULER_GAMMA) * j + s);
}

BF_HD static inline void bf_h0_pq(double x, double *P, double *Q) {
  if (x < BF_H0_X_ASYM) {
    /* Clenshaw on t in [-1, 1] */
    double y = 1.0 / x;
    double t = (2.0 * y - (BF_H0_CHEB_YA + BF_H0_CHEB_YB)) / (BF_H0_CHEB_YB - BF_H0_CHEB_YA);
    double t2 = 2.0 * t;
    double bp1 = 0, bp2 = 0, bq1 = 0, bq2 = 0;
    for (int k = BF_H0_NCHEB - 1; k >= 1; --k) {
      double bp = fma(t2, bp1, bf_h0_cheb_p[k] - bp2);
      double bq = fma(t2, bq1, bf_h0_cheb_q[k] - bq2);
      bp2 = bp1; bp1 = bp;
      bq2 = bq1; bq1 = bq;
    }
    *P = fma(t, bp1, bf_h0_cheb_p[0] - bp2);
    *Q = fma(t, bq1, bf_h0_cheb_q[0] - bq2);
  } else {
    double w = 1.0 / x;
    double w2 = w * w;
    double p = bf_h0_asym_p[BF_H0_NASYM - 1];
    double q = bf_h0_asym_q[BF_H0_NASYM - 1];
    for (int k = BF_H0_NASYM - 2; k >= 0; --k) {
      p = fma(p, w2, bf_h0_asym_p[k]);
      q = fma(q, w2, bf_h0_asym_q[k]);
    }
    *P = p;
    *Q = q * w;
  }
}

/* cos and sin of x - pi/4 */
BF_HD static inline void bf_phase(double x, double *c, double *s) {
  double n = rint(x * BF_TWO_OVER_PI);
  double r = fma(-n, BF_PIO2_1, x);
  r = fma(-n, BF_PIO2_2, r);
  r = fma(-n, BF_PIO2_3, r);
  double u = (r - BF_PIO4_HI) - BF_PIO4_LO;
  double su, cu;
  sincos(u, &su, &cu);
  long q = ((long)n) & 3;
  double cc = cu, ss = su;
  if (q == 1) { cc = -su; ss = cu; }
  else if (q == 2) { cc = -cu; ss = -su; }
  else if (q == 3) { cc = su; ss = -cu; }
  *c = cc;
  *s = ss;
}

/* H0^(1)(x) = J0 + i Y0, x > 0 */
BF_HD static inline void bf_h0(double x, double *jr, double *yr) {
  if (x <= BF_H0_X_SMALL) {
    bf_h0_series(x, jr, yr);
    return;
  }
  double P, Q, c, s;
  bf_h0_pq(x, &P, &Q);
  bf_phase(x, &c, &s);
  double amp = sqrt(BF_TWO_OVER_PI / x);
  *jr = amp * (P * c - Q * s);
  *yr = amp * (P * s + Q * c);
}

/* ---- H1^(1)(x) = J1 + i Y1 for the double layer, same three branches */
BF_HD static inline void bf_h1_series(double x, double *jr, double *yr) {
  double z = 0.25 * x * x;
  double j = bf_h1_ser_j[BF_H0_NSER - 1];
  double s = bf_h1_ser_y[BF_H0_NSER - 1];
  for (int k = BF_H0_NSER - 2; k >= 0; --k) {
    j = fma(j, z, bf_h1_ser_j[k]);
    s = fma(s, z, bf_h1_ser_y[k]);
  }
  const double hx = 0.5 * x;
  const double j1 = hx * j;
  *jr = j1;
  *yr = BF_TWO_OVER_PI * (log(hx) * j1 - 1.0 / x) + hx * s;
}

BF_HD static inline void bf_h1_pq(double x, double *P, double *Q) {
  if (x < BF_H0_X_ASYM) {
    double y = 1.0 / x;
    double t = (2.0 * y - (BF_H0_CHEB_YA + BF_H0_CHEB_YB)) / (BF_H0_CHEB_YB - BF_H0_CHEB_YA);
    double t2 = 2.0 * t;
    double bp1 = 0, bp2 = 0, bq1 = 0, bq2 = 0;
    for (int k = BF_H1_NCHEB - 1; k >= 1; --k) {
      double bp = fma(t2, bp1, bf_h1_cheb_p[k] - bp2);
      double bq = fma(t2, bq1, bf_h1_cheb_q[k] - bq2);
      bp2 = bp1; bp1 = bp;
      bq2 = bq1; bq1 = bq;
    }
    *P = fma(t, bp1, bf_h1_cheb_p[0] - bp2);
    *Q = fma(t, bq1, bf_h1_cheb_q[0] - bq2);
  } else {
    double w = 1.0 / x;
    double w2 = w * w;
    double p = bf_h1_asym_p[BF_H0_NASYM - 1];
    double q = bf_h1_asym_q[BF_H0_NASYM - 1];
    for (int k = BF_H0_NASYM - 2; k >= 0; --k) {
      p = fma(p, w2, bf_h1_asym_p[k]);
      q = fma(q, w2, bf_h1_asym_q[k]);
    }
    *P = p;
    *Q = q * w;
  }
}

/* H1^(1)(x) = J1 + i Y1, x > 0; above X_SMALL with the phase
 * x - 3 pi/4 = (x - pi/4) - pi/2: cos = sin(x - pi/4), sin = -cos(x - pi/4) */
BF_HD static inline void bf_h1(double x, double *jr, double *yr) {
  if (x <= BF_H0_X_SMALL) {
    bf_h1_series(x, jr, yr);
    return;
  }
  double P, Q, c, s;
  bf_h1_pq(x, &P, &Q);
  bf_phase(x, &c, &s);
  double amp = sqrt(BF_TWO_OVER_PI / x);
  *jr = amp * (P * s + Q * c);
  *yr = amp * (Q * s - P * c);
}

/* double layer with the source normal n: (i/4) k H1(k r) (n . (t - s)) / r,
 * 0 for r < 1e-15 (helm2.c:51-58); `dot` = n . (t - s) */
BF_HD static inline void bf_helm2_double(double k, double r, double dot, double *gr, double *gi) {
  if (r < 1e-15) {
    *gr = 0;
    *gi = 0;
    return;
  }
  double j, y;
  bf_h1(k * r, &j, &y);
  const double sc = 0.25 * k * dot / r;
  *gr = -sc * y;
  *gi = sc * j;
}

/* G = (i/4) H0(k r) with G = 0 at r == 0 (helm2.c:114) */
BF_HD static inline void bf_helm2_single(double k, double r, double *gr, double *gi) {
  if (r == 0) {
    *gr = 0;
    *gi = 0;
    return;
  }
  double j, y;
  bf_h0(k * r, &j, &y);
  *gr = -0.25 * y;
  *gi = 0.25 * j;
}
