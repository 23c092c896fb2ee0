/*
 * A second-kind boundary integral operator on the device and GMRES for it
 * (helm2_bie.c:95-172, quadrature.c:12-200, linalg.c:47-317).
 *
 *   A x = alpha x + (S + C) W x
 *
 * S is a device butterfly factorization of the kernel on the boundary
 * points (sources = targets, zero on the diagonal as helm2.c:114), W the
 * quadrature weights, C the Kapur-Rokhlin corrections of the logarithmic
 * singularity: C[i][j] = g_p G(x_i <- x_j) for the 2 x order neighbours
 * j = i +- (p + 1) along the boundary (quadrature.c:126-166).  Everything
 * is in tree order on the device; one application is a row scaling, the
 * butterfly apply and one fused kernel for alpha x + C (W x).
 *
 * GMRES keeps its Krylov basis in HBM in FP64 and orthogonalizes with
 * classical Gram-Schmidt applied twice (two matrix-vector products with
 * the basis per step, which keeps the basis orthogonal to working
 * precision); the small Hessenberg problem (Givens rotations) is on the
 * host.  The mixed solver refines in FP64: r = b - A x with the FP64
 * apply, an inner GMRES on A d = r with the FP32 apply, x += d.
 */
#include <complex.h>
#include <math.h>

#include "bf_host.h"

/* Kapur-Rokhlin end corrections for a logarithmic singularity
 * (Kapur & Rokhlin, SIAM J. Numer. Anal. 34 (1997), the gamma table; the
 * same values as quadrature.c:12-27) */
static const double KR2[2] = {1.825748064736159, -1.325748064736159};
static const double KR6[6] = {4.967362978287758, -16.20501504859126, 25.85153761832639,
                              -22.22599466791883, 9.930104998037539, -1.817995878141594};

typedef struct BfMatHipBie {
  BfMat super;
  BfMat *fac; /* not owned */
  int64_t n;
  double alpha;
  int ncorr;
  double *dW;
  int32_t *dCol;
  double *dC;
  void *dU, *dTmp;
  size_t workBytes;
} BfMatHipBie;

static BfType bie_type(BfMat const *m) { (void)m; return BF_TYPE_MAT_HIP_BIE; }
static BfSize bie_bytes(BfMat const *m) {
  const BfMatHipBie *A = (const BfMatHipBie *)m;
  return bfMatNumBytes(A->fac) + (BfSize)A->n * (8 + (BfSize)A->ncorr * 20);
}
static void bie_delete(BfMat **mp) {
  BfMatHipBie *A = (BfMatHipBie *)*mp;
  bfk_free(A->dW);
  bfk_free(A->dCol);
  bfk_free(A->dC);
  bfk_free(A->dU);
  bfk_free(A->dTmp);
  free(A);
  *mp = NULL;
}
static BfMat *bie_mul(BfMat const *A, BfMat const *X);
static BfMatVtable bieVtable = {bie_type, bie_bytes, bie_delete, NULL, bie_mul, NULL};

BfMat *bfhipHelm2BieNew(BfMat *fac, BfHelm2 const *helm, BfPoints2 const *points, BfReal const *weights,
                        BfPerm const *perm, BfSize krOrder, BfReal alpha) {
  if (!fac || bfMatGetType(fac) != BF_TYPE_MAT_HIP_BUTTERFLY || !helm || !points || !weights || !perm ||
      fac->numRows != fac->numCols || fac->numRows != points->size || perm->size != points->size ||
      (krOrder != 0 && krOrder != 2 && krOrder != 6) || (krOrder && points->size < 2 * krOrder + 1)) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  /* the corrections treat the logarithmic singularity of the single layer:
   * none for the (smooth) double layer, alpha times them for the combined
   * field alpha S + beta D; for S' they are S' values at the neighbours with
   * the target point's normal (krComplexKernel, helm2.c:372-378), read from
   * the factorization's target normals (the boundary's own) */
  BfMatHipButterfly const *Fb = (BfMatHipButterfly const *)fac;
  const BfLayerPotential lp = Fb->helm.layerPot;
  const int sprime = lp == BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE;
  if ((lp == BF_LAYER_POTENTIAL_PV_DOUBLE && krOrder != 0) || (sprime && krOrder != 0 && !Fb->dTgtN)) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  const BfComplex cscale = lp == BF_LAYER_POTENTIAL_COMBINED_FIELD ? ((BfMatHipButterfly const *)fac)->helm.alpha : 1;
  const int64_t n = (int64_t)points->size;
  const int ord = (int)krOrder, nc = 2 * ord;
  const double *g = ord == 2 ? KR2 : KR6;
  BfMatHipBie *A = calloc(1, sizeof(BfMatHipBie));
  bfMatInitBase(&A->super, &bieVtable, (BfSize)n, (BfSize)n);
  A->fac = fac;
  A->n = n;
  A->alpha = alpha;
  A->ncorr = nc;
  /* weights and corrections in tree order */
  BfSize *rev = malloc((size_t)n * sizeof(BfSize));
  for (int64_t i = 0; i < n; ++i) rev[perm->index[i]] = (BfSize)i;
  double *w = malloc((size_t)n * sizeof(double));
  for (int64_t i = 0; i < n; ++i) w[i] = weights[perm->index[i]];
  int64_t np = n * nc;
  int32_t *col = malloc((size_t)(np ? np : 1) * sizeof(int32_t));
  double *pa = malloc((size_t)(np ? np : 1) * 2 * sizeof(double)), *pb = malloc((size_t)(np ? np : 1) * 2 * sizeof(double));
  double *gam = malloc((size_t)(np ? np : 1) * sizeof(double));
  /* S': the normal at each row's point (tree order on the device) */
  double *tn = sprime && np ? malloc((size_t)n * 16) : NULL, *pn = sprime && np ? malloc((size_t)np * 16) : NULL;
  int rc = 0;
  if (tn) rc = bfk_d2h(tn, Fb->dTgtN, (size_t)n * 16);
  for (int64_t i = 0; i < n; ++i) {
    const int64_t oi = (int64_t)perm->index[i];
    for (int p = 0; p < ord; ++p)
      for (int side = 0; side < 2; ++side) {
        const int64_t oj = side == 0 ? (oi + p + 1) % n : (oi + n - p - 1) % n;
        const int64_t q = i * nc + 2 * p + side;
        col[q] = (int32_t)rev[oj];
        pa[2 * q] = points->data[oi][0];
        pa[2 * q + 1] = points->data[oi][1];
        pb[2 * q] = points->data[oj][0];
        pb[2 * q + 1] = points->data[oj][1];
        gam[q] = g[p];
        if (pn && tn) {
          pn[2 * q] = tn[2 * i];
          pn[2 * q + 1] = tn[2 * i + 1];
        }
      }
  }
  void *dA = NULL, *dB = NULL, *dNA = NULL;
  A->dW = bfk_malloc((size_t)n * 8);
  A->dCol = bfk_malloc((size_t)(np ? np : 1) * 4);
  A->dC = bfk_malloc((size_t)(np ? np : 1) * 16);
  dA = bfk_malloc((size_t)(np ? np : 1) * 16);
  dB = bfk_malloc((size_t)(np ? np : 1) * 16);
  if (pn) dNA = bfk_malloc((size_t)np * 16);
  if (!A->dW || !A->dCol || !A->dC || !dA || !dB || (sprime && np && (!tn || !pn || !dNA))) rc = rc ? rc : -2;
  if (!rc && dNA) rc = bfk_h2d(dNA, pn, (size_t)np * 16);
  if (!rc) rc = bfk_h2d(A->dW, w, (size_t)n * 8);
  if (!rc && np) rc = bfk_h2d(A->dCol, col, (size_t)np * 4);
  if (!rc && np) rc = bfk_h2d(dA, pa, (size_t)np * 16);
  if (!rc && np) rc = bfk_h2d(dB, pb, (size_t)np * 16);
  if (!rc && np) rc = bfk_pair_green(np, dA, dB, (const double *)dNA, helm->k, A->dC);
  if (!rc && np) { /* scale by the correction weights on the host */
    double *c = malloc((size_t)np * 16);
    rc = bfk_d2h(c, A->dC, (size_t)np * 16);
    for (int64_t q = 0; q < np && !rc; ++q) {
      const BfComplex v = (c[2 * q] + I * c[2 * q + 1]) * gam[q] * cscale;
      c[2 * q] = creal(v);
      c[2 * q + 1] = cimag(v);
    }
    if (!rc) rc = bfk_h2d(A->dC, c, (size_t)np * 16);
    free(c);
  }
  bfk_free(dA);
  bfk_free(dB);
  bfk_free(dNA);
  free(rev); free(w); free(col); free(pa); free(pb); free(gam); free(tn); free(pn);
  if (rc) {
    BfMat *m = &A->super;
    bie_delete(&m);
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  return &A->super;
}

static int ensure_work(BfMatHipBie *A, size_t bytes) {
  if (bytes <= A->workBytes) return 0;
  bfk_free(A->dU);
  bfk_free(A->dTmp);
  A->dU = bfk_malloc(bytes);
  A->dTmp = bfk_malloc(bytes);
  A->workBytes = bytes;
  return (!A->dU || !A->dTmp) ? -2 : 0;
}

/* y = A x on device vectors (n x nrhs, tree order), FP64 or FP32 */
int bfhipBieMulDevice(BfMat const *op, void const *dX, void *dY, BfSize nrhs, BfhipPrecision prec) {
  if (!op || bfMatGetType(op) != BF_TYPE_MAT_HIP_BIE || !dX || !dY || nrhs == 0 || prec == BFHIP_PRECISION_FP32X2) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  BfMatHipBie *A = (BfMatHipBie *)op;
  const int fp32 = prec == BFHIP_PRECISION_FP32;
  int rc = ensure_work(A, (size_t)A->n * nrhs * 16);
  if (!rc) rc = bfk_scale_rows(A->n, (int)nrhs, A->dW, dX, A->dU, fp32);
  if (!rc) rc = bfhipMatMulDevice(A->fac, A->dU, dY, nrhs, prec);
  if (!rc) rc = bfk_bie_post(A->n, (int)nrhs, A->ncorr, A->alpha, A->dCol, A->dC, dX, A->dU, dY, fp32);
  if (!rc) rc = bfk_sync();
  if (rc) bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
  return rc;
}

static BfMat *bie_mul(BfMat const *Am, BfMat const *Xm) {
  const BfMatHipBie *A = (const BfMatHipBie *)Am;
  if (!Xm || bfMatGetType(Xm) != BF_TYPE_MAT_DENSE_COMPLEX || (int64_t)Xm->numRows != A->n) {
    bfSetError(BF_ERROR_INCOMPATIBLE_SHAPES);
    return NULL;
  }
  const BfMatDenseComplex *X = (const BfMatDenseComplex *)Xm;
  const int64_t n = A->n, nrhs = (int64_t)Xm->numCols;
  BfComplex *hx = malloc((size_t)(n * nrhs) * sizeof(BfComplex));
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < nrhs; ++j) hx[i * nrhs + j] = X->data[i * X->rowStride + j * X->colStride];
  void *dx = bfk_malloc((size_t)(n * nrhs) * 16), *dy = bfk_malloc((size_t)(n * nrhs) * 16);
  int rc = (!dx || !dy) ? -2 : 0;
  if (!rc) rc = bfk_h2d(dx, hx, (size_t)(n * nrhs) * 16);
  if (!rc) rc = bfhipBieMulDevice(Am, dx, dy, (BfSize)nrhs, BFHIP_PRECISION_FP64);
  BfMatDenseComplex *Y = NULL;
  if (!rc) {
    Y = bfMatDenseComplexNew();
    bfMatDenseComplexInit(Y, (BfSize)n, (BfSize)nrhs);
    rc = bfk_d2h(Y->data, dy, (size_t)(n * nrhs) * 16);
  }
  bfk_free(dx);
  bfk_free(dy);
  free(hx);
  if (rc) {
    if (Y) bfMatDenseComplexDeinitAndDealloc(&Y);
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  return &Y->super;
}

/* ---------------------------------------------------------------- GMRES */

typedef struct {
  BfMatHipBie *A;
  int64_t n;
  int m;            /* Krylov dimension per cycle */
  double *dV;       /* (m + 1) x n, column j at dV + 2 j n */
  float *dv32, *dw32;
} Gmres;

static double *vcol(Gmres *G, int j) { return G->dV + (size_t)2 * G->n * j; }

/* w = A v (dv, dw FP64 device vectors) with the FP64 or FP32 apply */
static int op_apply(Gmres *G, const double *dv, double *dw, int fp32) {
  if (!fp32) return bfhipBieMulDevice(&G->A->super, dv, dw, 1, BFHIP_PRECISION_FP64);
  int rc = bfk_convert(G->n, dv, G->dv32, 1);
  if (!rc) rc = bfhipBieMulDevice(&G->A->super, G->dv32, G->dw32, 1, BFHIP_PRECISION_FP32);
  if (!rc) rc = bfk_convert(G->n, G->dw32, dw, 0);
  return rc;
}

static int dev_norm(Gmres *G, const double *dx, double *out) {
  double h[2];
  int rc = bfk_dotc_multi(G->n, 1, dx, G->n, dx, h);
  *out = sqrt(h[0] > 0 ? h[0] : 0);
  return rc;
}

/* GMRES(m) with restarts for A x = b (FP64 device vectors; x holds the
 * initial guess), stopping when the recursive residual estimate reaches
 * tol |b|; returns the inner iterations used in *iters */
static int gmres_solve(Gmres *G, const double *db, double *dx, double tol, int maxIter, int fp32, int *iters,
                       double *bnormOut) {
  const int64_t n = G->n;
  const int m = G->m;
  double bnorm;
  int rc = dev_norm(G, db, &bnorm);
  if (bnormOut) *bnormOut = bnorm;
  *iters = 0;
  if (rc || bnorm == 0) return rc;
  double *H = calloc((size_t)(m + 1) * m * 2, sizeof(double)); /* column-major (m+1) x m complex */
  double *cs = calloc((size_t)m, sizeof(double)), *sn = calloc((size_t)m * 2, sizeof(double));
  double *gv = calloc((size_t)(m + 1) * 2, sizeof(double));
  double *h = calloc((size_t)(m + 1) * 2, sizeof(double)), *h2 = calloc((size_t)(m + 1) * 2, sizeof(double));
  double *y = calloc((size_t)m * 2, sizeof(double));
  double one[2] = {1, 0};
  int done = 0;
  while (!rc && !done && *iters < maxIter) {
    /* r = b - A x into V_0 */
    double *v0 = vcol(G, 0);
    rc = op_apply(G, dx, v0, fp32);
    if (!rc) rc = bfk_copy(db, vcol(G, 1), (size_t)n * 16); /* V_1 as scratch */
    if (!rc) rc = bfk_axpy_multi(n, 1, v0, n, one, 1.0, vcol(G, 1));
    if (!rc) rc = bfk_copy(vcol(G, 1), v0, (size_t)n * 16);
    double beta = 0;
    if (!rc) rc = dev_norm(G, v0, &beta);
    if (rc) break;
    if (beta <= tol * bnorm) break;
    rc = bfk_axpy_multi(n, 0, v0, n, NULL, 1.0 / beta, v0);
    memset(gv, 0, (size_t)(m + 1) * 2 * sizeof(double));
    gv[0] = beta;
    int k = 0;
    for (int j = 0; j < m && !rc && *iters < maxIter; ++j) {
      double *w = vcol(G, j + 1);
      rc = op_apply(G, vcol(G, j), w, fp32);
      /* classical Gram-Schmidt, twice */
      if (!rc) rc = bfk_dotc_multi(n, j + 1, G->dV, n, w, h);
      if (!rc) rc = bfk_axpy_multi(n, j + 1, G->dV, n, h, 1.0, w);
      if (!rc) rc = bfk_dotc_multi(n, j + 1, G->dV, n, w, h2);
      if (!rc) rc = bfk_axpy_multi(n, j + 1, G->dV, n, h2, 1.0, w);
      double hn = 0;
      if (!rc) rc = dev_norm(G, w, &hn);
      if (rc) break;
      for (int i = 0; i <= j; ++i) {
        H[2 * (j * (m + 1) + i)] = h[2 * i] + h2[2 * i];
        H[2 * (j * (m + 1) + i) + 1] = h[2 * i + 1] + h2[2 * i + 1];
      }
      H[2 * (j * (m + 1) + j + 1)] = hn;
      H[2 * (j * (m + 1) + j + 1) + 1] = 0;
      if (hn > 0) rc = bfk_axpy_multi(n, 0, w, n, NULL, 1.0 / hn, w);
      /* previous rotations on column j: [c s; -conj(s) c] */
      for (int i = 0; i < j; ++i) {
        double *a = &H[2 * (j * (m + 1) + i)], *b = &H[2 * (j * (m + 1) + i + 1)];
        const double c = cs[i], sr = sn[2 * i], si = sn[2 * i + 1];
        const double ar = a[0], ai = a[1], br = b[0], bi = b[1];
        a[0] = c * ar + (sr * br - si * bi);
        a[1] = c * ai + (sr * bi + si * br);
        b[0] = -(sr * ar + si * ai) + c * br;
        b[1] = -(sr * ai - si * ar) + c * bi;
      }
      /* new rotation zeroing H[j+1][j] */
      {
        double *a = &H[2 * (j * (m + 1) + j)], *b = &H[2 * (j * (m + 1) + j + 1)];
        const double aa = hypot(a[0], a[1]), bb = hypot(b[0], b[1]), r = hypot(aa, bb);
        double c = 1, sr = 0, si = 0;
        if (r > 0) {
          if (aa == 0) {
            c = 0;
            sr = b[0] / bb; /* s = conj(b)/|b| with a = 0 */
            si = -b[1] / bb;
          } else {
            c = aa / r;
            /* s = (a/|a|) conj(b) / r */
            const double pr = a[0] / aa, pi = a[1] / aa;
            sr = (pr * b[0] + pi * b[1]) / r;
            si = (pi * b[0] - pr * b[1]) / r;
          }
        }
        cs[j] = c;
        sn[2 * j] = sr;
        sn[2 * j + 1] = si;
        const double ar = a[0], ai = a[1], br = b[0], bi = b[1];
        a[0] = c * ar + (sr * br - si * bi);
        a[1] = c * ai + (sr * bi + si * br);
        b[0] = 0;
        b[1] = 0;
        const double gr = gv[2 * j], gi = gv[2 * j + 1];
        gv[2 * j] = c * gr;
        gv[2 * j + 1] = c * gi;
        gv[2 * (j + 1)] = -(sr * gr + si * gi);
        gv[2 * (j + 1) + 1] = -(sr * gi - si * gr);
      }
      ++*iters;
      k = j + 1;
      if (hypot(gv[2 * k], gv[2 * k + 1]) <= tol * bnorm || hn == 0) {
        done = 1;
        break;
      }
    }
    if (rc || k == 0) break;
    /* y = H(0:k, 0:k)^-1 g(0:k) (upper triangular) */
    for (int i = k - 1; i >= 0; --i) {
      double sr = gv[2 * i], si = gv[2 * i + 1];
      for (int l = i + 1; l < k; ++l) {
        const double *hh = &H[2 * (l * (m + 1) + i)];
        sr -= hh[0] * y[2 * l] - hh[1] * y[2 * l + 1];
        si -= hh[0] * y[2 * l + 1] + hh[1] * y[2 * l];
      }
      const double *d = &H[2 * (i * (m + 1) + i)];
      const double dd = d[0] * d[0] + d[1] * d[1];
      y[2 * i] = (sr * d[0] + si * d[1]) / dd;
      y[2 * i + 1] = (si * d[0] - sr * d[1]) / dd;
    }
    for (int i = 0; i < k; ++i) { y[2 * i] = -y[2 * i]; y[2 * i + 1] = -y[2 * i + 1]; }
    rc = bfk_axpy_multi(n, k, G->dV, n, y, 1.0, dx); /* x += V y */
  }
  free(H); free(cs); free(sn); free(gv); free(h); free(h2); free(y);
  return rc;
}

static int gmres_init(Gmres *G, BfMatHipBie *A, int m, int fp32) {
  memset(G, 0, sizeof(*G));
  G->A = A;
  G->n = A->n;
  G->m = m;
  G->dV = bfk_malloc((size_t)(m + 1) * (size_t)A->n * 16);
  if (!G->dV) return -2;
  if (fp32) {
    G->dv32 = bfk_malloc((size_t)A->n * 8);
    G->dw32 = bfk_malloc((size_t)A->n * 8);
    if (!G->dv32 || !G->dw32) return -2;
  }
  return 0;
}

static void gmres_free(Gmres *G) {
  bfk_free(G->dV);
  bfk_free(G->dv32);
  bfk_free(G->dw32);
}

/* linalg.c:47-317: X = A^-1 B by GMRES without restarts (maxNumIter Krylov
 * vectors), one right-hand side at a time; A is a device boundary operator
 * (bfhipHelm2BieNew); B and X0 host dense complex.  Left preconditioning
 * (M) is not provided. */
BfMat *bfSolveGMRES(BfMat const *Am, BfMat const *Bm, BfMat *X0, BfReal tol, BfSize maxNumIter, BfSize *numIter,
                    BfMat const *M) {
  if (M) {
    bfSetError(BF_ERROR_NOT_IMPLEMENTED);
    return NULL;
  }
  if (!Am || bfMatGetType(Am) != BF_TYPE_MAT_HIP_BIE || !Bm || bfMatGetType(Bm) != BF_TYPE_MAT_DENSE_COMPLEX ||
      maxNumIter == 0 || Bm->numRows != Am->numRows ||
      (X0 && (bfMatGetType(X0) != BF_TYPE_MAT_DENSE_COMPLEX || X0->numRows != Bm->numRows ||
              X0->numCols != Bm->numCols))) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  BfMatHipBie *A = (BfMatHipBie *)Am;
  const BfMatDenseComplex *B = (const BfMatDenseComplex *)Bm, *X0d = (const BfMatDenseComplex *)X0;
  const int64_t n = A->n, nrhs = (int64_t)Bm->numCols;
  BfMatDenseComplex *X = bfMatDenseComplexNew();
  bfMatDenseComplexInit(X, (BfSize)n, (BfSize)nrhs);
  Gmres G;
  int rc = gmres_init(&G, A, (int)maxNumIter, 0);
  void *db = bfk_malloc((size_t)n * 16), *dx = bfk_malloc((size_t)n * 16);
  BfComplex *col = malloc((size_t)n * sizeof(BfComplex));
  if (!db || !dx) rc = -2;
  int maxIt = 0;
  for (int64_t r = 0; r < nrhs && !rc; ++r) {
    for (int64_t i = 0; i < n; ++i) col[i] = B->data[i * B->rowStride + r * B->colStride];
    rc = bfk_h2d(db, col, (size_t)n * 16);
    for (int64_t i = 0; i < n; ++i) col[i] = X0d ? X0d->data[i * X0d->rowStride + r * X0d->colStride] : 0;
    if (!rc) rc = bfk_h2d(dx, col, (size_t)n * 16);
    int it = 0;
    if (!rc) rc = gmres_solve(&G, db, dx, tol, (int)maxNumIter, 0, &it, NULL);
    if (!rc) rc = bfk_d2h(col, dx, (size_t)n * 16);
    for (int64_t i = 0; i < n && !rc; ++i) X->data[i * X->rowStride + r * X->colStride] = col[i];
    if (it > maxIt) maxIt = it;
  }
  if (numIter) *numIter = (BfSize)maxIt;
  gmres_free(&G);
  bfk_free(db);
  bfk_free(dx);
  free(col);
  if (rc) {
    bfMatDenseComplexDeinitAndDealloc(&X);
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  return &X->super;
}

/* Mixed precision: FP64 residuals, inner GMRES(restart) on A d = r with
 * the FP32 apply (bfhipFacPrepareMixed must have run on the factorization).
 * b and x are host vectors in tree order.  Returns 0 when |b - A x| <=
 * tol |b| with the FP64 apply. */
int bfhipSolveGmresMixed(BfMat const *Am, BfComplex const *b, BfComplex *x, BfReal tol, BfSize restart,
                         BfSize maxIter, BfSize *numIter, BfSize *numRefine, BfReal *relRes) {
  if (!Am || bfMatGetType(Am) != BF_TYPE_MAT_HIP_BIE || !b || !x || restart == 0 || maxIter == 0) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  BfMatHipBie *A = (BfMatHipBie *)Am;
  const int64_t n = A->n;
  Gmres G;
  int rc = gmres_init(&G, A, (int)restart, 1);
  double *db = bfk_malloc((size_t)n * 16), *dx = bfk_malloc((size_t)n * 16), *dr = bfk_malloc((size_t)n * 16);
  double *dd = bfk_malloc((size_t)n * 16), *dt = bfk_malloc((size_t)n * 16);
  if (!db || !dx || !dr || !dd || !dt) rc = -2;
  if (!rc) rc = bfk_h2d(db, b, (size_t)n * 16);
  if (!rc) rc = bfk_memset(dx, (size_t)n * 16);
  double bnorm = 0, rel = 1;
  if (!rc) rc = dev_norm(&G, db, &bnorm);
  int total = 0, outer = 0;
  double one[2] = {1, 0}, mone[2] = {-1, 0};
  while (!rc && total < (int)maxIter) {
    /* r = b - A x in FP64 */
    rc = op_apply(&G, dx, dt, 0);
    if (!rc) rc = bfk_copy(db, dr, (size_t)n * 16);
    if (!rc) rc = bfk_axpy_multi(n, 1, dt, n, one, 1.0, dr);
    double rn = 0;
    if (!rc) rc = dev_norm(&G, dr, &rn);
    if (rc) break;
    rel = bnorm > 0 ? rn / bnorm : 0;
    if (rel <= tol) break;
    /* inner solve to what the FP32 apply can resolve */
    const double inner = fmax(tol / rel, 1e-5);
    int it = 0;
    rc = bfk_memset(dd, (size_t)n * 16);
    if (!rc) rc = gmres_solve(&G, dr, dd, inner, (int)maxIter - total, 1, &it, NULL);
    total += it;
    ++outer;
    if (!rc) rc = bfk_axpy_multi(n, 1, dd, n, mone, 1.0, dx); /* x = x - d (-1) */
    if (it == 0) break;
  }
  if (!rc) rc = bfk_d2h(x, dx, (size_t)n * 16);
  if (numIter) *numIter = (BfSize)total;
  if (numRefine) *numRefine = (BfSize)outer;
  if (relRes) *relRes = rel;
  gmres_free(&G);
  bfk_free(db); bfk_free(dx); bfk_free(dr); bfk_free(dd); bfk_free(dt);
  if (rc) {
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return rc;
  }
  return rel <= tol ? 0 : 1;
}
