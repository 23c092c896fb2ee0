/*
 * Streaming adaptive butterfly factorization (fac_streamer.c:54-572,
 * fac.c:84-1294, fac_span.c:1-157, linalg.c:26-35 and :1040-1120).
 *
 * The matrix arrives one column block at a time, in the post-order of the
 * column tree.  A fed block is cut along the row tree into row blocks that
 * compress (truncated SVD, sigma_k >= tol sigma_0, fewer terms than the
 * block's smaller side): Psi = diag(U_k), W0 = [S_k V_k^H; ...].  When all
 * children of a column node are in, their partial factorizations are
 * merged: along the coarsest common row cut, the children's Psi blocks
 * are set side by side (Psi*) and cut again into compressing row blocks;
 * the children's W0 rows for each cut node form the next factor W1, and
 * their trailing factors are set on a block diagonal.  The result for the
 * whole matrix is A ~ Psi W0 W1 ... W_L.
 *
 * Here the block bookkeeping is plain C on the host, every SVD of one
 * round (all row blocks of one level of one feed or merge) runs in one
 * batched launch of the one-sided Jacobi kernel, and the finished product
 * is packed into one device arena and applied by the stage kernels of
 * the butterfly apply.
 */
#include <math.h>

#include "bf_host.h"

/* ------------------------------------------------ block-sparse matrices */

typedef struct {
  int64_t i0, j0, m, n;
  BfComplex *data; /* column-major m x n; NULL = identity (m == n) */
} SBlock;

typedef struct {
  int64_t rows, cols;
  SBlock *b;
  size_t nb, cap;
} SMat;

static void smat_init(SMat *s, int64_t rows, int64_t cols) {
  memset(s, 0, sizeof(*s));
  s->rows = rows;
  s->cols = cols;
}

/* takes ownership of data; empty blocks are kept so that the blocks of a
 * Psi factor stay one per row node */
static void smat_push(SMat *s, int64_t i0, int64_t j0, int64_t m, int64_t n, BfComplex *data) {
  if (s->nb == s->cap) {
    s->cap = s->cap ? 2 * s->cap : 8;
    s->b = realloc(s->b, s->cap * sizeof(SBlock));
    if (!s->b) abort();
  }
  SBlock blk = {i0, j0, m, n, data};
  s->b[s->nb++] = blk;
}

static void smat_free(SMat *s) {
  for (size_t i = 0; i < s->nb; ++i) free(s->b[i].data);
  free(s->b);
  memset(s, 0, sizeof(*s));
}

/* an allocation count of at least one */
static size_t nz(int64_t n) { return n > 0 ? (size_t)n : 1; }

static BfComplex *cdup(const BfComplex *a, int64_t n) {
  BfComplex *d = malloc((size_t)(n ? n : 1) * sizeof(BfComplex));
  if (!d) abort();
  if (n) memcpy(d, a, (size_t)n * sizeof(BfComplex));
  return d;
}

/* append all blocks of `src` shifted by (di, dj), copying their data */
static void smat_append_shifted(SMat *dst, const SMat *src, int64_t di, int64_t dj) {
  for (size_t i = 0; i < src->nb; ++i) {
    const SBlock *b = &src->b[i];
    smat_push(dst, b->i0 + di, b->j0 + dj, b->m, b->n, b->data ? cdup(b->data, b->m * b->n) : NULL);
  }
}

/* rows [r0, r1) of `src` as blocks of `dst`, shifted so row r0 lands on
 * row di and column 0 on column dj */
static void smat_append_rows(SMat *dst, const SMat *src, int64_t r0, int64_t r1, int64_t di, int64_t dj) {
  for (size_t i = 0; i < src->nb; ++i) {
    const SBlock *b = &src->b[i];
    int64_t a = b->i0 > r0 ? b->i0 : r0, e = b->i0 + b->m < r1 ? b->i0 + b->m : r1;
    if (a >= e) continue;
    if (!b->data) {
      /* rows [a, e) of an identity block: a rectangular piece of I */
      int64_t h = e - a;
      BfComplex *d = calloc((size_t)(h * b->n), sizeof(BfComplex));
      for (int64_t r = 0; r < h; ++r) d[(a - b->i0 + r) * h + r] = 1;
      smat_push(dst, a - r0 + di, b->j0 + dj, h, b->n, d);
      continue;
    }
    if (a == b->i0 && e == b->i0 + b->m) {
      smat_push(dst, a - r0 + di, b->j0 + dj, b->m, b->n, cdup(b->data, b->m * b->n));
      continue;
    }
    int64_t h = e - a;
    BfComplex *d = malloc((size_t)(h * b->n) * sizeof(BfComplex));
    for (int64_t c = 0; c < b->n; ++c)
      for (int64_t r = 0; r < h; ++r) d[c * h + r] = b->data[c * b->m + (a - b->i0) + r];
    smat_push(dst, a - r0 + di, b->j0 + dj, h, b->n, d);
  }
}

static int64_t smat_entries(const SMat *s) {
  int64_t e = 0;
  for (size_t i = 0; i < s->nb; ++i)
    if (s->b[i].data) e += s->b[i].m * s->b[i].n;
  return e;
}

/* ----------------------------------------------------------- the trees */

typedef struct {
  const BfQuadtreeNode *nodes;
} TreeView;

static TreeView tree_view(const BfTree *t) {
  TreeView v = {t->root};
  return v;
}
static int64_t nd_start(TreeView t, int32_t i) { return (int64_t)t.nodes[i].offset[0]; }
static int64_t nd_end(TreeView t, int32_t i) { return (int64_t)t.nodes[i].offset[4]; }
static int nd_leaf(TreeView t, int32_t i) {
  for (int q = 0; q < 4; ++q)
    if (t.nodes[i].child[q] >= 0) return 0;
  return 1;
}

/* ------------------------------------------------------------ the facs */

struct BfFac {
  const BfQuadtreeNode *colNode;
  int32_t colIndex;
  int32_t *rowNodes; /* row-tree node indices of the row cut, in row order */
  size_t numRowNodes;
  SMat Psi;
  size_t numW;
  SMat *W;
};

struct BfFacSpan {
  BfSize numFacs;
  BfFac **fac;
};

static void fac_free(BfFac *f) {
  if (!f) return;
  free(f->rowNodes);
  smat_free(&f->Psi);
  for (size_t i = 0; i < f->numW; ++i) smat_free(&f->W[i]);
  free(f->W);
  free(f);
}

void bfFacDeinitAndDealloc(BfFac **fac) {
  if (!fac) return;
  fac_free(*fac);
  *fac = NULL;
}

BfSize bfFacGetNumBytes(BfFac const *fac) {
  if (!fac) return 0;
  int64_t e = smat_entries(&fac->Psi);
  for (size_t i = 0; i < fac->numW; ++i) e += smat_entries(&fac->W[i]);
  return (BfSize)e * sizeof(BfComplex);
}

/* ------------------------------------------------ batched truncated SVD */

/* One truncated SVD request: the m x n column-major matrix `a`. */
typedef struct {
  int64_t m, n;
  BfComplex *a;
  /* results */
  int64_t k;          /* retained terms */
  int truncated;      /* k < min(m, n) */
  BfComplex *U;       /* m x k column-major */
  BfComplex *W;       /* k x n column-major: S_k V_k^H */
} SvdJob;

typedef struct { double s; int32_t j; } SigIdx;
static int cmp_sig_desc(const void *a, const void *b) {
  const SigIdx *x = a, *y = b;
  if (x->s != y->s) return x->s > y->s ? -1 : 1;
  return x->j - y->j;
}

/* the largest smaller side of a matrix the batched SVD takes (its cost
 * grows as the square of the column count) */
#define SVD_MAX_COLS 1024
/* matrices of at least this many entries get the whole chip per SVD */
#define SVD_BIG_ENTRIES (1 << 15)

/* One-sided Jacobi SVDs of a batch whose tables (host copies h*, device
 * copies d*) are in place: large matrices take the per-step launches over
 * the whole chip, the rest one workgroup each in one launch.  The device
 * tables are overwritten (with the small matrices' entries). */
static int jacobi_dispatch(int64_t nj, const int64_t *aOff, const int32_t *M, const int32_t *N, const int64_t *vOff,
                           const int64_t *sOff, void *dAo, void *dM, void *dN, void *dVo, void *dSo, void *dA,
                           void *dV, void *dS) {
  int32_t *bigIdx = malloc(nz(nj) * sizeof(int32_t)), *tallIdx = malloc(nz(nj) * sizeof(int32_t));
  int64_t *sA = malloc(nz(nj) * 8), *sV = malloc(nz(nj) * 8), *sS = malloc(nz(nj) * 8);
  int32_t *sM = malloc(nz(nj) * 4), *sN = malloc(nz(nj) * 4);
  int64_t nbig = 0, nsmall = 0, ntall = 0, tallEnd = 0;
  int maxN = 0, maxTallN = 0, rc = 0;
  const int tallMax = bfk_svd_tall_max_n();
  for (int64_t i = 0; i < nj; ++i) {
    if ((int64_t)M[i] * N[i] >= SVD_BIG_ENTRIES && N[i] >= 2 && N[i] <= tallMax && M[i] >= 2 * N[i]) {
      /* tall: QR first, the Jacobi on the n x n triangle in LDS */
      tallIdx[ntall++] = (int32_t)i;
      if (N[i] > maxTallN) maxTallN = N[i];
      if (aOff[i] + (int64_t)M[i] * N[i] > tallEnd) tallEnd = aOff[i] + (int64_t)M[i] * N[i];
    } else if ((int64_t)M[i] * N[i] >= SVD_BIG_ENTRIES && N[i] >= 2) {
      bigIdx[nbig++] = (int32_t)i;
      if (N[i] > maxN) maxN = N[i];
    } else {
      sA[nsmall] = aOff[i];
      sV[nsmall] = vOff[i];
      sS[nsmall] = sOff[i];
      sM[nsmall] = M[i];
      sN[nsmall] = N[i];
      ++nsmall;
    }
  }
  if (ntall) {
    void *dT = bfk_malloc((size_t)ntall * 4), *dScr = bfk_malloc((size_t)tallEnd * 16);
    if (!dT || !dScr) rc = -2;
    if (!rc) rc = bfk_h2d(dT, tallIdx, (size_t)ntall * 4);
    if (!rc) rc = bfk_svd_jacobi_tall(ntall, dT, dAo, dM, dN, dA, dScr, dVo, dV, dSo, dS, 60, maxTallN);
    bfk_free(dT);
    bfk_free(dScr);
  }
  if (nbig && !rc) {
    void *dB = bfk_malloc((size_t)nbig * 4);
    if (!dB) rc = -2;
    if (!rc) rc = bfk_h2d(dB, bigIdx, (size_t)nbig * 4);
    if (!rc) rc = bfk_svd_jacobi_big(nbig, dB, dAo, dM, dN, dA, dVo, dV, dSo, dS, 60, maxN);
    bfk_free(dB);
  }
  if ((nbig || ntall) && !rc) {
    if (nsmall) { /* the small ones' tables */
      rc = bfk_h2d(dAo, sA, (size_t)nsmall * 8);
      if (!rc) rc = bfk_h2d(dVo, sV, (size_t)nsmall * 8);
      if (!rc) rc = bfk_h2d(dSo, sS, (size_t)nsmall * 8);
      if (!rc) rc = bfk_h2d(dM, sM, (size_t)nsmall * 4);
      if (!rc) rc = bfk_h2d(dN, sN, (size_t)nsmall * 4);
    }
  }
  if (!rc && nsmall) rc = bfk_svd_jacobi(nsmall, dAo, dM, dN, dA, dVo, dV, dSo, dS, 60);
  free(bigIdx); free(tallIdx); free(sA); free(sV); free(sS); free(sM); free(sN);
  return rc;
}

/* All jobs in one launch.  Wide blocks (m < n) are decomposed through
 * their conjugate transpose so the kernel always sees n <= m columns. */
static int svd_batch(SvdJob *jobs, size_t nj, double tol) {
  if (nj == 0) return 0;
  int64_t *aOff = malloc(nj * sizeof(int64_t)), *vOff = malloc(nj * sizeof(int64_t)),
          *sOff = malloc(nj * sizeof(int64_t));
  int32_t *M = malloc(nj * sizeof(int32_t)), *N = malloc(nj * sizeof(int32_t));
  int64_t totA = 0, totV = 0, totS = 0;
  for (size_t i = 0; i < nj; ++i) {
    int wide = jobs[i].m < jobs[i].n;
    int64_t m = wide ? jobs[i].n : jobs[i].m, n = wide ? jobs[i].m : jobs[i].n;
    if (n > SVD_MAX_COLS) {
      free(aOff); free(vOff); free(sOff); free(M); free(N);
      return -3;
    }
    M[i] = (int32_t)m;
    N[i] = (int32_t)n;
    aOff[i] = totA;
    vOff[i] = totV;
    sOff[i] = totS;
    totA += m * n;
    totV += n * n;
    totS += n;
  }
  BfComplex *hA = malloc((size_t)(totA ? totA : 1) * sizeof(BfComplex));
  BfComplex *hV = malloc((size_t)(totV ? totV : 1) * sizeof(BfComplex));
  double *hS = malloc((size_t)(totS ? totS : 1) * sizeof(double));
  for (size_t i = 0; i < nj; ++i) {
    SvdJob *J = &jobs[i];
    BfComplex *dst = hA + aOff[i];
    if (J->m < J->n) { /* store A^H (n x m) column-major */
      for (int64_t c = 0; c < J->m; ++c)
        for (int64_t r = 0; r < J->n; ++r) dst[c * J->n + r] = conj(J->a[r * J->m + c]);
    } else {
      memcpy(dst, J->a, (size_t)(J->m * J->n) * sizeof(BfComplex));
    }
  }
  int rc = 0;
  void *dA = NULL, *dV = NULL, *dS = NULL, *dAo = NULL, *dVo = NULL, *dSo = NULL, *dM = NULL, *dN = NULL;
  dA = bfk_malloc((size_t)(totA ? totA : 1) * 16);
  dV = bfk_malloc((size_t)(totV ? totV : 1) * 16);
  dS = bfk_malloc((size_t)(totS ? totS : 1) * 8);
  dAo = bfk_malloc(nj * 8);
  dVo = bfk_malloc(nj * 8);
  dSo = bfk_malloc(nj * 8);
  dM = bfk_malloc(nj * 4);
  dN = bfk_malloc(nj * 4);
  if (!dA || !dV || !dS || !dAo || !dVo || !dSo || !dM || !dN) rc = -2;
  if (!rc) rc = bfk_h2d(dA, hA, (size_t)totA * 16);
  if (!rc) rc = bfk_h2d(dAo, aOff, nj * 8);
  if (!rc) rc = bfk_h2d(dVo, vOff, nj * 8);
  if (!rc) rc = bfk_h2d(dSo, sOff, nj * 8);
  if (!rc) rc = bfk_h2d(dM, M, nj * 4);
  if (!rc) rc = bfk_h2d(dN, N, nj * 4);
  if (!rc) rc = jacobi_dispatch((int64_t)nj, aOff, M, N, vOff, sOff, dAo, dM, dN, dVo, dSo, dA, dV, dS);
  if (!rc) rc = bfk_d2h(hA, dA, (size_t)totA * 16);
  if (!rc) rc = bfk_d2h(hV, dV, (size_t)totV * 16);
  if (!rc) rc = bfk_d2h(hS, dS, (size_t)totS * 8);
  bfk_free(dA); bfk_free(dV); bfk_free(dS); bfk_free(dAo); bfk_free(dVo); bfk_free(dSo);
  bfk_free(dM); bfk_free(dN);
  for (size_t i = 0; i < nj && !rc; ++i) {
    SvdJob *J = &jobs[i];
    const int64_t m = M[i], n = N[i];
    const BfComplex *Wc = hA + aOff[i]; /* m x n: columns sigma_j u_j */
    const BfComplex *Vc = hV + vOff[i]; /* n x n: columns v_j */
    const double *s = hS + sOff[i];
    SigIdx *si = malloc((size_t)n * sizeof(SigIdx));
    for (int64_t j = 0; j < n; ++j) { si[j].s = s[j]; si[j].j = (int32_t)j; }
    qsort(si, (size_t)n, sizeof(SigIdx), cmp_sig_desc);
    int64_t k = 0; /* linalg.c:26-35 */
    while (k < n && si[k].s >= tol * si[0].s) ++k;
    J->k = k;
    J->truncated = k < n;
    /* for a block A = sum_j sigma_j u_j v_j^H (tall form):
     *   U = [u_j] (m x k), W = [sigma_j v_j^H] (k x n).
     * For a wide block the kernel saw B = A^H = sum sigma u' v'^H, so
     *   A = sum sigma v' u'^H: U = [v'_j], W = [sigma_j u'_j^H]. */
    const int64_t Am = J->m, An = J->n;
    J->U = malloc((size_t)(Am * (k ? k : 1)) * sizeof(BfComplex));
    J->W = malloc((size_t)((k ? k : 1) * An) * sizeof(BfComplex));
    for (int64_t q = 0; q < k; ++q) {
      const int32_t j = si[q].j;
      const double sg = si[q].s;
      if (J->m >= J->n) {
        for (int64_t r = 0; r < m; ++r) J->U[q * Am + r] = sg > 0 ? Wc[(int64_t)j * m + r] / sg : 0;
        for (int64_t c = 0; c < n; ++c) J->W[c * k + q] = sg * conj(Vc[(int64_t)j * n + c]);
      } else {
        for (int64_t r = 0; r < n; ++r) J->U[q * Am + r] = Vc[(int64_t)j * n + r];
        for (int64_t c = 0; c < m; ++c) J->W[c * k + q] = conj(Wc[(int64_t)j * m + c]);
      }
    }
    free(si);
  }
  free(hA); free(hV); free(hS); free(aOff); free(vOff); free(sOff); free(M); free(N);
  return rc;
}

static void svd_job_free(SvdJob *J) {
  free(J->a);
  free(J->U);
  free(J->W);
  J->a = J->U = J->W = NULL;
}

int bfhipSvdTallMaxN(void) {
  if (bfk_init()) {
    bfSetError(BF_ERROR_RUNTIME_ERROR);
    return -1;
  }
  return bfk_svd_tall_max_n();
}

int bfhipSvdBatched(BfSize batch, BfSize const *m, BfSize const *n, BfComplex const *A, BfReal tol, BfSize *k,
                    BfComplex *U, BfComplex *W) {
  if (!m || !n || !A || !k || !U || !W) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  SvdJob *jobs = calloc(batch ? batch : 1, sizeof(SvdJob));
  int64_t off = 0;
  for (BfSize i = 0; i < batch; ++i) {
    jobs[i].m = (int64_t)m[i];
    jobs[i].n = (int64_t)n[i];
    jobs[i].a = cdup(A + off, jobs[i].m * jobs[i].n);
    off += jobs[i].m * jobs[i].n;
  }
  int rc = svd_batch(jobs, batch, tol);
  int64_t uo = 0, wo = 0;
  for (BfSize i = 0; i < batch; ++i) {
    if (!rc) {
      k[i] = (BfSize)jobs[i].k;
      memcpy(U + uo, jobs[i].U, (size_t)(jobs[i].m * jobs[i].k) * sizeof(BfComplex));
      memcpy(W + wo, jobs[i].W, (size_t)(jobs[i].k * jobs[i].n) * sizeof(BfComplex));
    }
    /* the caller sizes U and W for full rank */
    int64_t full = jobs[i].m < jobs[i].n ? jobs[i].m : jobs[i].n;
    uo += jobs[i].m * full;
    wo += full * jobs[i].n;
    svd_job_free(&jobs[i]);
  }
  free(jobs);
  if (rc) bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
  return rc;
}

/* ------------------------------------------- randomized truncated SVD */

/* one matrix of a sketch batch (the kernels' BfkSketchJob) */
typedef struct {
  int32_t m, n, w, pad;
  int64_t aOff, oOff, yOff, bOff;
} SketchJob;

/* Truncated SVDs through a randomized range finder (Halko, Martinsson &
 * Tropp 2011, Alg. 4.1 then 5.1): Y = A Omega (n x w Gaussian Omega),
 * Q = orth(Y) from the one-sided Jacobi of Y, B^H = A^H Q, then the Jacobi
 * SVD of B^H gives A ~ Q B = sum_j (Q y_j)(sigma_j x_j^H).  A sketch is
 * accepted when ||A - Q B||_F <= tol sigma_0 (so every discarded singular
 * value of A is below the cut) and fewer than w terms pass the cut (the
 * width did not saturate); a rejected sketch (fell[i] = 1) and every job
 * with width 0, width >= min(m, n) or width > SVD_MAX_COLS take the full SVD.  The
 * rule for k is svd_batch's (linalg.c:26-35). */
static int svd_batch_sketch(SvdJob *jobs, size_t nj, double tol, const int64_t *width, uint64_t seed, int *fell) {
  if (nj == 0) return 0;
  if (bfk_sketch_jobs_bytes() != (int)sizeof(SketchJob)) return -1;
  size_t *sk = malloc(nz((int64_t)nj) * sizeof(size_t)), *full = malloc(nz((int64_t)nj) * sizeof(size_t));
  size_t ns = 0, nf = 0;
  for (size_t i = 0; i < nj; ++i) {
    fell[i] = 0;
    const int64_t mn = jobs[i].m < jobs[i].n ? jobs[i].m : jobs[i].n;
    if (width[i] > 0 && width[i] < mn && width[i] <= SVD_MAX_COLS) sk[ns++] = i;
    else full[nf++] = i;
  }
  int rc = 0;
  if (ns) {
    SketchJob *sj = calloc(ns, sizeof(SketchJob));
    int64_t *yOff = malloc(ns * 8), *bOff = malloc(ns * 8), *v1 = malloc(ns * 8), *v2 = malloc(ns * 8),
            *s1 = malloc(ns * 8), *s2 = malloc(ns * 8);
    int32_t *M = malloc(ns * 4), *N = malloc(ns * 4), *Wd = malloc(ns * 4);
    int64_t tA = 0, tO = 0, tY = 0, tB = 0, tV = 0, tS = 0, maxMW = 0, maxNW = 0, maxMN = 0;
    for (size_t q = 0; q < ns; ++q) {
      const SvdJob *J = &jobs[sk[q]];
      const int64_t w = width[sk[q]];
      if (J->m * w > maxMW) maxMW = J->m * w;
      if (J->n * w > maxNW) maxNW = J->n * w;
      if (J->m * J->n > maxMN) maxMN = J->m * J->n;
      SketchJob x = {(int32_t)J->m, (int32_t)J->n, (int32_t)w, 0, tA, tO, tY, tB};
      sj[q] = x;
      M[q] = (int32_t)J->m;
      N[q] = (int32_t)J->n;
      Wd[q] = (int32_t)w;
      yOff[q] = tY;
      bOff[q] = tB;
      v1[q] = tV;
      v2[q] = tV + w * w;
      s1[q] = tS;
      s2[q] = tS + w;
      tA += J->m * J->n;
      tO += J->n * w;
      tY += J->m * w;
      tB += J->n * w;
      tV += 2 * w * w;
      tS += 2 * w;
    }
    BfComplex *hA = malloc(nz(tA) * sizeof(BfComplex)), *hY = malloc(nz(tY) * sizeof(BfComplex));
    BfComplex *hB = malloc(nz(tB) * sizeof(BfComplex)), *hV = malloc(nz(tV) * sizeof(BfComplex));
    const int chunks = bfk_sketch_resid_chunks(maxMN);
    double *hS = malloc(nz(tS) * sizeof(double)), *hR = malloc(ns * (size_t)chunks * sizeof(double));
    for (size_t q = 0; q < ns; ++q) {
      const SvdJob *J = &jobs[sk[q]];
      memcpy(hA + sj[q].aOff, J->a, (size_t)(J->m * J->n) * sizeof(BfComplex));
    }
    void *dA = bfk_malloc(nz(tA) * 16), *dO = bfk_malloc(nz(tO) * 16), *dY = bfk_malloc(nz(tY) * 16),
         *dB = bfk_malloc(nz(tB) * 16), *dV = bfk_malloc(nz(tV) * 16), *dS = bfk_malloc(nz(tS) * 8),
         *dR = bfk_malloc(ns * (size_t)chunks * 8), *dJ = bfk_malloc(ns * sizeof(SketchJob)), *dYo = bfk_malloc(ns * 8),
         *dBo = bfk_malloc(ns * 8), *dV1 = bfk_malloc(ns * 8), *dV2 = bfk_malloc(ns * 8), *dS1 = bfk_malloc(ns * 8),
         *dS2 = bfk_malloc(ns * 8), *dM = bfk_malloc(ns * 4), *dN = bfk_malloc(ns * 4), *dW = bfk_malloc(ns * 4);
    if (!dA || !dO || !dY || !dB || !dV || !dS || !dR || !dJ || !dYo || !dBo || !dV1 || !dV2 || !dS1 || !dS2 ||
        !dM || !dN || !dW)
      rc = -2;
    if (!rc) rc = bfk_h2d(dA, hA, (size_t)tA * 16);
    if (!rc) rc = bfk_h2d(dJ, sj, ns * sizeof(SketchJob));
    if (!rc) rc = bfk_h2d(dYo, yOff, ns * 8);
    if (!rc) rc = bfk_h2d(dBo, bOff, ns * 8);
    if (!rc) rc = bfk_h2d(dV1, v1, ns * 8);
    if (!rc) rc = bfk_h2d(dV2, v2, ns * 8);
    if (!rc) rc = bfk_h2d(dS1, s1, ns * 8);
    if (!rc) rc = bfk_h2d(dS2, s2, ns * 8);
    if (!rc) rc = bfk_h2d(dM, M, ns * 4);
    if (!rc) rc = bfk_h2d(dN, N, ns * 4);
    if (!rc) rc = bfk_h2d(dW, Wd, ns * 4);
    if (!rc) rc = bfk_sketch_range((int64_t)ns, dJ, seed, dA, dO, dY, maxMW);
    if (!rc) rc = jacobi_dispatch((int64_t)ns, yOff, M, Wd, v1, s1, dYo, dM, dW, dV1, dS1, dY, dV, dS);
    /* the dispatcher left the small matrices' tables in place: the
     * projection reads the sigma offsets of every matrix */
    if (!rc) rc = bfk_h2d(dS1, s1, ns * 8);
    if (!rc) rc = bfk_h2d(dW, Wd, ns * 4);
    if (!rc) rc = bfk_sketch_project((int64_t)ns, dJ, dS1, dS, dA, dY, dB, dR, maxMW, maxNW, maxMN);
    if (!rc) rc = jacobi_dispatch((int64_t)ns, bOff, N, Wd, v2, s2, dBo, dN, dW, dV2, dS2, dB, dV, dS);
    if (!rc) rc = bfk_d2h(hY, dY, (size_t)tY * 16);
    if (!rc) rc = bfk_d2h(hB, dB, (size_t)tB * 16);
    if (!rc) rc = bfk_d2h(hV, dV, (size_t)tV * 16);
    if (!rc) rc = bfk_d2h(hS, dS, (size_t)tS * 8);
    if (!rc) rc = bfk_d2h(hR, dR, ns * (size_t)chunks * 8);
    for (size_t q = 0; q < ns && !rc; ++q) { /* the residual partials, summed in order */
      const int used = bfk_sketch_resid_chunks(sj[q].m * (int64_t)sj[q].n);
      double t = 0;
      for (int c = 0; c < used; ++c) t += hR[q * (size_t)chunks + c];
      hR[q] = t;
    }
    bfk_free(dA); bfk_free(dO); bfk_free(dY); bfk_free(dB); bfk_free(dV); bfk_free(dS); bfk_free(dR);
    bfk_free(dJ); bfk_free(dYo); bfk_free(dBo); bfk_free(dV1); bfk_free(dV2); bfk_free(dS1); bfk_free(dS2);
    bfk_free(dM); bfk_free(dN); bfk_free(dW);
    for (size_t q = 0; q < ns && !rc; ++q) {
      SvdJob *J = &jobs[sk[q]];
      const int64_t m = J->m, n = J->n, w = Wd[q];
      const double *sg = hS + s2[q];
      SigIdx *si = malloc((size_t)w * sizeof(SigIdx));
      for (int64_t j = 0; j < w; ++j) { si[j].s = sg[j]; si[j].j = (int32_t)j; }
      qsort(si, (size_t)w, sizeof(SigIdx), cmp_sig_desc);
      const double s0 = si[0].s;
      int64_t k = 0;
      while (k < w && si[k].s >= tol * s0) ++k;
      if (!(s0 > 0) || k >= w || !(hR[q] <= tol * s0 * tol * s0)) {
        fell[sk[q]] = 1;
        full[nf++] = sk[q];
        free(si);
        continue;
      }
      const BfComplex *Q = hY + yOff[q], *BH = hB + bOff[q], *V = hV + v2[q];
      J->k = k;
      J->truncated = 1;
      J->U = malloc((size_t)(m * k) * sizeof(BfComplex));
      J->W = malloc((size_t)(k * n) * sizeof(BfComplex));
      for (int64_t t = 0; t < k; ++t) {
        const int32_t j = si[t].j;
        const BfComplex *vj = V + (int64_t)j * w;
        for (int64_t r = 0; r < m; ++r) {
          BfComplex u = 0;
          for (int64_t c = 0; c < w; ++c) u += Q[c * m + r] * vj[c];
          J->U[t * m + r] = u;
        }
        for (int64_t c = 0; c < n; ++c) J->W[c * k + t] = conj(BH[(int64_t)j * n + c]);
      }
      free(si);
    }
    free(hA); free(hY); free(hB); free(hV); free(hS); free(hR);
    free(sj); free(yOff); free(bOff); free(v1); free(v2); free(s1); free(s2); free(M); free(N); free(Wd);
  }
  if (!rc && nf) { /* the full SVDs, rejected sketches included */
    SvdJob *fj = malloc(nf * sizeof(SvdJob));
    for (size_t x = 0; x < nf; ++x) fj[x] = jobs[full[x]];
    rc = svd_batch(fj, nf, tol);
    for (size_t x = 0; x < nf; ++x) {
      SvdJob *J = &jobs[full[x]];
      J->k = fj[x].k;
      J->truncated = fj[x].truncated;
      J->U = fj[x].U;
      J->W = fj[x].W;
    }
    free(fj);
  }
  free(sk);
  free(full);
  return rc;
}

int bfhipSvdBatchedSketch(BfSize batch, BfSize const *m, BfSize const *n, BfComplex const *A, BfReal tol,
                          BfSize const *width, uint64_t seed, BfSize *k, BfComplex *U, BfComplex *W, int *fellBack) {
  if (!m || !n || !A || !width || !k || !U || !W || !fellBack) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  SvdJob *jobs = calloc(batch ? batch : 1, sizeof(SvdJob));
  int64_t *wd = malloc((batch ? batch : 1) * sizeof(int64_t));
  int64_t off = 0;
  for (BfSize i = 0; i < batch; ++i) {
    jobs[i].m = (int64_t)m[i];
    jobs[i].n = (int64_t)n[i];
    jobs[i].a = cdup(A + off, jobs[i].m * jobs[i].n);
    wd[i] = (int64_t)width[i];
    off += jobs[i].m * jobs[i].n;
  }
  int rc = svd_batch_sketch(jobs, batch, tol, wd, seed, fellBack);
  int64_t uo = 0, wo = 0;
  for (BfSize i = 0; i < batch; ++i) {
    if (!rc) {
      k[i] = (BfSize)jobs[i].k;
      memcpy(U + uo, jobs[i].U, (size_t)(jobs[i].m * jobs[i].k) * sizeof(BfComplex));
      memcpy(W + wo, jobs[i].W, (size_t)(jobs[i].k * jobs[i].n) * sizeof(BfComplex));
    }
    int64_t fullr = jobs[i].m < jobs[i].n ? jobs[i].m : jobs[i].n;
    uo += jobs[i].m * fullr;
    wo += fullr * jobs[i].n;
    svd_job_free(&jobs[i]);
  }
  free(jobs);
  free(wd);
  if (rc) bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : rc == -3 ? BF_ERROR_NOT_IMPLEMENTED : BF_ERROR_RUNTIME_ERROR);
  return rc;
}

/* ------------------------------------------------ rank model in a stream */

/* the stream's rank model, its samples and counters */
typedef struct {
  const BfhipRankModel *model;
  double k;
  int64_t oversample;
  int on;              /* record samples (set by bfhipFacStreamerSetRankModel) */
  TreeView ct;
  BF_VEC(float) feat;  /* 6 per sample */
  BF_VEC(float) rank;
  int64_t numSvd, numSketched, numFell;
  uint64_t seed;
} RankCtx;

/* the SVDs of one round: with a model, widths from its predictions;
 * samples recorded afterwards */
static int svd_round(RankCtx *rk, const BfFacSpec *spec, TreeView rt, int32_t colNode, const int32_t *rowNodes,
                     SvdJob *jobs, size_t nj) {
  if (!rk || !rk->on || nj == 0) return svd_batch(jobs, nj, spec->tol);
  float *f = malloc(nj * 6 * sizeof(float));
  const BfQuadtreeNode *cn = &rk->ct.nodes[colNode];
  const BfReal cb[4] = {cn->bboxMin[0], cn->bboxMin[1], cn->bboxMax[0], cn->bboxMax[1]};
  for (size_t i = 0; i < nj; ++i) {
    const BfQuadtreeNode *rn = &rt.nodes[rowNodes[i]];
    const BfReal rb[4] = {rn->bboxMin[0], rn->bboxMin[1], rn->bboxMax[0], rn->bboxMax[1]};
    bfhipRankFeatures(cb, rb, rk->k, (BfSize)rn->depth, spec->tol, f + 6 * i);
  }
  int rc = 0;
  if (rk->model) {
    float *pr = malloc(nj * sizeof(float));
    int64_t *wd = malloc(nj * sizeof(int64_t));
    int *fell = malloc(nj * sizeof(int));
    rc = bfhipRankModelPredict(rk->model, nj, f, pr);
    for (size_t i = 0; i < nj && !rc; ++i) {
      /* a sketch costs two Jacobi SVDs of w columns (on m and n rows) and
       * three m n w products: it pays only for the large matrices (one
       * launch per Jacobi step: w - 1 steps twice against min(m, n) - 1)
       * and only when w is well below min(m, n) */
      const double p = pr[i] > 0 ? ceil((double)pr[i]) : 0;
      const int64_t mn = jobs[i].m < jobs[i].n ? jobs[i].m : jobs[i].n;
      wd[i] = (int64_t)p + rk->oversample;
      if (2 * wd[i] > mn || jobs[i].m * jobs[i].n < SVD_BIG_ENTRIES) wd[i] = 0;
    }
    if (!rc) rc = svd_batch_sketch(jobs, nj, spec->tol, wd, rk->seed++, fell);
    for (size_t i = 0; i < nj && !rc; ++i) {
      const int64_t mn = jobs[i].m < jobs[i].n ? jobs[i].m : jobs[i].n;
      if (wd[i] > 0 && wd[i] < mn && wd[i] <= SVD_MAX_COLS) rk->numSketched += 1;
      rk->numFell += fell[i];
    }
    free(pr);
    free(wd);
    free(fell);
  } else {
    rc = svd_batch(jobs, nj, spec->tol);
  }
  for (size_t i = 0; i < nj && !rc; ++i) {
    for (int q = 0; q < 6; ++q) BF_VEC_PUSH(rk->feat, f[6 * i + q]);
    BF_VEC_PUSH(rk->rank, (float)jobs[i].k);
  }
  rk->numSvd += (int64_t)nj;
  free(f);
  return rc;
}

/* --------------------------------------------- epsilon-rank row cuts */

/* One row block being cut: node `nd` of the row tree, the dense rows of
 * its parent matrix relative to row `base`. */
typedef struct {
  int32_t nd;
  int32_t owner;   /* which cut (feed: 0, merge: merge-cut node) */
  SvdJob svd;
  int kind;        /* 0 svd, 1 few rows (Psi = I), 2 few columns (W = I), 3 leaf pass-through */
} CutItem;

typedef struct {
  int32_t nd;
  int32_t owner;
  int64_t k;        /* columns of its Psi block = rows of its W block */
  BfComplex *Psi;   /* rows x k column-major, NULL = identity */
  BfComplex *W;     /* k x cols column-major, NULL = identity */
} CutOut;

typedef BF_VEC(CutItem) CutItemVec;
typedef BF_VEC(CutOut) CutOutVec;

/* column-major copy of rows [r0, r1) of the dense column-major A (lda rows) */
static BfComplex *rows_copy(const BfComplex *A, int64_t lda, int64_t cols, int64_t r0, int64_t r1) {
  int64_t h = r1 - r0;
  BfComplex *d = malloc(nz(h * cols) * sizeof(BfComplex));
  for (int64_t c = 0; c < cols; ++c) memcpy(d + c * h, A + c * lda + r0, (size_t)h * sizeof(BfComplex));
  return d;
}

/* Cut the dense matrices mats[o] (rows of node roots[o], `cols[o]`
 * columns, column-major with leading dimension = the root's row count)
 * along the row tree, all owners level by level with one SVD launch per
 * level.  `mergeRule` selects the merge test (also require k < rows).
 * Returns 0, or 1 when a leaf would not compress and `bail` is set. */
static int eps_cut(const BfFacSpec *spec, TreeView rt, size_t numOwners, int32_t *const *frontiers,
                   const size_t *frontierLen, const int32_t *roots, BfComplex *const *mats, const int64_t *cols,
                   int mergeRule, int bail, CutOutVec *out, int *rcOut, RankCtx *rk, int32_t colNode) {
  CutItemVec cur, nxt;
  memset(&cur, 0, sizeof(cur));
  memset(&nxt, 0, sizeof(nxt));
  for (size_t o = 0; o < numOwners; ++o)
    for (size_t i = 0; i < frontierLen[o]; ++i) {
      CutItem it;
      memset(&it, 0, sizeof(it));
      it.nd = frontiers[o][i];
      it.owner = (int32_t)o;
      BF_VEC_PUSH(cur, it);
    }
  int bailed = 0;
  *rcOut = 0;
  while (cur.n && !bailed) {
    /* classify, build the SVD batch */
    BF_VEC(SvdJob) jobs;
    memset(&jobs, 0, sizeof(jobs));
    BF_VEC(int32_t) jobNodes;
    memset(&jobNodes, 0, sizeof(jobNodes));
    for (size_t i = 0; i < cur.n; ++i) {
      CutItem *it = &cur.a[i];
      const int o = it->owner;
      const int64_t base = nd_start(rt, roots[o]), lda = nd_end(rt, roots[o]) - base;
      const int64_t r0 = nd_start(rt, it->nd) - base, r1 = nd_end(rt, it->nd) - base;
      if (!mergeRule && cols[o] < (int64_t)spec->minNumCols) it->kind = 2;
      else if (r1 - r0 < (int64_t)spec->minNumRows) it->kind = 1;
      else if (mergeRule && cols[o] < (int64_t)spec->minNumCols) it->kind = 2;
      else {
        it->kind = 0;
        SvdJob J;
        memset(&J, 0, sizeof(J));
        J.m = r1 - r0;
        J.n = cols[o];
        J.a = rows_copy(mats[o], lda, cols[o], r0, r1);
        BF_VEC_PUSH(jobs, J);
        BF_VEC_PUSH(jobNodes, it->nd);
      }
    }
    int rc = svd_round(rk, spec, rt, colNode, jobNodes.a, jobs.a, jobs.n);
    BF_VEC_FREE(jobNodes);
    if (rc) {
      *rcOut = rc;
      for (size_t j = 0; j < jobs.n; ++j) svd_job_free(&jobs.a[j]);
      BF_VEC_FREE(jobs);
      break;
    }
    size_t jq = 0;
    nxt.n = 0;
    for (size_t i = 0; i < cur.n; ++i) {
      CutItem *it = &cur.a[i];
      const int o = it->owner;
      const int64_t base = nd_start(rt, roots[o]), lda = nd_end(rt, roots[o]) - base;
      const int64_t r0 = nd_start(rt, it->nd) - base, r1 = nd_end(rt, it->nd) - base;
      CutOut co;
      memset(&co, 0, sizeof(co));
      co.nd = it->nd;
      co.owner = it->owner;
      if (it->kind == 1) { /* few rows: Psi = I, W = the rows (fac.c:746-760, :934-945) */
        co.k = r1 - r0;
        co.W = rows_copy(mats[o], lda, cols[o], r0, r1);
        BF_VEC_PUSH(*out, co);
        continue;
      }
      if (it->kind == 2) { /* few columns: Psi = the rows, W = I (fac.c:649-676, :952-966) */
        co.k = cols[o];
        co.Psi = rows_copy(mats[o], lda, cols[o], r0, r1);
        BF_VEC_PUSH(*out, co);
        continue;
      }
      SvdJob *J = &jobs.a[jq++];
      int ok = J->truncated && (!mergeRule || J->k < J->m);
      if (ok) {
        co.k = J->k;
        co.Psi = J->U;
        co.W = J->W;
        J->U = J->W = NULL;
        BF_VEC_PUSH(*out, co);
      } else if (nd_leaf(rt, it->nd)) {
        if (bail && !mergeRule) {
          bailed = 1;
        } else { /* a leaf that does not compress passes through */
          co.k = r1 - r0;
          co.W = rows_copy(mats[o], lda, cols[o], r0, r1);
          BF_VEC_PUSH(*out, co);
        }
      } else {
        for (int q = 0; q < 4; ++q) {
          int32_t c = rt.nodes[it->nd].child[q];
          if (c < 0) continue;
          CutItem ch;
          memset(&ch, 0, sizeof(ch));
          ch.nd = c;
          ch.owner = it->owner;
          BF_VEC_PUSH(nxt, ch);
        }
      }
      svd_job_free(J);
    }
    for (; jq < jobs.n; ++jq) svd_job_free(&jobs.a[jq]);
    BF_VEC_FREE(jobs);
    CutItemVec t = cur;
    cur = nxt;
    nxt = t;
  }
  BF_VEC_FREE(cur);
  BF_VEC_FREE(nxt);
  return bailed;
}

/* sort CutOuts by (owner, row start) with an insertion-friendly merge:
 * items are few per owner, so a simple stable sort suffices */
static void sort_cutouts(CutOutVec *v, TreeView rt) {
  for (size_t i = 1; i < v->n; ++i) {
    CutOut x = v->a[i];
    size_t j = i;
    while (j > 0 && (v->a[j - 1].owner > x.owner ||
                     (v->a[j - 1].owner == x.owner && nd_start(rt, v->a[j - 1].nd) > nd_start(rt, x.nd)))) {
      v->a[j] = v->a[j - 1];
      --j;
    }
    v->a[j] = x;
  }
}

/* --------------------------------------------------------- the streamer */

struct BfFacStreamer {
  BfFacSpec spec;
  TreeView rt, ct;
  BfPerm *rowRevPerm;
  int32_t *post;
  size_t numPost, pos;
  BfFac **facs;
  size_t numFacs, capFacs;
  int32_t *initCut;
  size_t initCutLen;
  RankCtx rank;
};

BfFacStreamer *bfFacStreamerNew(void) {
  BfFacStreamer *s = calloc(1, sizeof(BfFacStreamer));
  if (!s) bfSetError(BF_ERROR_MEMORY_ERROR);
  return s;
}

static void post_order(TreeView t, int32_t i, int32_t *out, size_t *n) {
  for (int q = 0; q < 4; ++q)
    if (t.nodes[i].child[q] >= 0) post_order(t, t.nodes[i].child[q], out, n);
  out[(*n)++] = i;
}

static size_t count_nodes(TreeView t, int32_t i) {
  size_t c = 1;
  for (int q = 0; q < 4; ++q)
    if (t.nodes[i].child[q] >= 0) c += count_nodes(t, t.nodes[i].child[q]);
  return c;
}

/* the starting cut: nodes at the initial depth, and shallower leaves so
 * that every row is covered, in row order */
static void init_cut(TreeView t, int32_t i, int depth, int32_t *out, size_t *n) {
  if (t.nodes[i].depth >= depth || nd_leaf(t, i)) {
    out[(*n)++] = i;
    return;
  }
  for (int q = 0; q < 4; ++q)
    if (t.nodes[i].child[q] >= 0) init_cut(t, t.nodes[i].child[q], depth, out, n);
}

void bfFacStreamerInit(BfFacStreamer *s, BfFacSpec const *spec) {
  if (!s || !spec || !spec->rowTree || !spec->colTree || !spec->rowTree->root || !spec->colTree->root ||
      !(spec->tol > 0)) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return;
  }
  memset(s, 0, sizeof(*s));
  s->spec = *spec;
  s->rt = tree_view(spec->rowTree);
  s->ct = tree_view(spec->colTree);
  s->rowRevPerm = spec->rowTree->perm ? bfPermGetReversePerm(spec->rowTree->perm) : NULL;
  size_t nc = count_nodes(s->ct, 0);
  s->post = malloc(nc * sizeof(int32_t));
  s->numPost = 0;
  post_order(s->ct, 0, s->post, &s->numPost);
  size_t nr = count_nodes(s->rt, 0);
  s->initCut = malloc(nr * sizeof(int32_t));
  s->initCutLen = 0;
  init_cut(s->rt, 0, (int)spec->rowTreeInitDepth, s->initCut, &s->initCutLen);
}

void bfFacStreamerDeinit(BfFacStreamer *s) {
  if (!s) return;
  if (s->rowRevPerm) {
    free(s->rowRevPerm->index);
    free(s->rowRevPerm);
  }
  for (size_t i = 0; i < s->numFacs; ++i) fac_free(s->facs[i]);
  free(s->facs);
  free(s->post);
  free(s->initCut);
  BF_VEC_FREE(s->rank.feat);
  BF_VEC_FREE(s->rank.rank);
  memset(s, 0, sizeof(*s));
}

/* after bfFacStreamerInit (which clears it) */
int bfhipFacStreamerSetRankModel(BfFacStreamer *s, BfhipRankModel const *model, BfReal k, BfSize oversample) {
  if (!s || !s->post || !(k >= 0) || oversample > 256) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  s->rank.model = model;
  s->rank.k = k;
  s->rank.oversample = (int64_t)oversample;
  s->rank.on = 1;
  s->rank.ct = s->ct;
  s->rank.seed = 0x2545F4914F6CDD1Dull;
  return 0;
}

BfSize bfhipFacStreamerGetRankSamples(BfFacStreamer const *s, float *features, float *ranks, BfSize cap) {
  if (!s) return 0;
  const BfSize n = (BfSize)s->rank.rank.n;
  const BfSize c = n < cap ? n : cap;
  if (features && c) memcpy(features, s->rank.feat.a, c * 6 * sizeof(float));
  if (ranks && c) memcpy(ranks, s->rank.rank.a, c * sizeof(float));
  return n;
}

void bfhipFacStreamerGetSketchStats(BfFacStreamer const *s, BfSize *svds, BfSize *sketched, BfSize *fellBack) {
  if (svds) *svds = s ? (BfSize)s->rank.numSvd : 0;
  if (sketched) *sketched = s ? (BfSize)s->rank.numSketched : 0;
  if (fellBack) *fellBack = s ? (BfSize)s->rank.numFell : 0;
}

void bfFacStreamerDealloc(BfFacStreamer **s) {
  if (!s) return;
  free(*s);
  *s = NULL;
}

void bfFacStreamerDelete(BfFacStreamer **s) {
  if (!s || !*s) return;
  bfFacStreamerDeinit(*s);
  bfFacStreamerDealloc(s);
}

BfSize bfFacStreamerGetNumRows(BfFacStreamer const *s) {
  return (BfSize)(nd_end(s->rt, 0) - nd_start(s->rt, 0));
}

bool bfFacStreamerIsDone(BfFacStreamer const *s) { return s->pos >= s->numPost; }

BfQuadtreeNode const *bfFacStreamerGetCurrentColumnNode(BfFacStreamer const *s) {
  if (bfFacStreamerIsDone(s)) return NULL;
  return &s->ct.nodes[s->post[s->pos]];
}

BfPerm const *bfFacStreamerGetRowTreeReversePerm(BfFacStreamer const *s) { return s->rowRevPerm; }

static void push_fac(BfFacStreamer *s, BfFac *f) {
  if (s->numFacs == s->capFacs) {
    s->capFacs = s->capFacs ? 2 * s->capFacs : 16;
    s->facs = realloc(s->facs, s->capFacs * sizeof(BfFac *));
    if (!s->facs) abort();
  }
  s->facs[s->numFacs++] = f;
}

/* Merge the partial factorizations of the children of column node c
 * (fac.c:1080-1294). */
static BfFac *merge_children(BfFacStreamer *s, int32_t c, int *rc) {
  TreeView rt = s->rt;
  BfFac *kids[4];
  int nk = 0;
  for (int q = 0; q < 4; ++q) {
    int32_t ch = s->ct.nodes[c].child[q];
    if (ch < 0) continue;
    for (size_t i = 0; i < s->numFacs; ++i)
      if (s->facs[i]->colIndex == ch) { kids[nk++] = s->facs[i]; break; }
  }
  *rc = 0;
  /* the coarsest common row cut (fac.c:509-573) */
  BF_VEC(int32_t) cut;
  memset(&cut, 0, sizeof(cut));
  size_t ptr[4] = {0, 0, 0, 0};
  int64_t i = nd_start(rt, 0), end = nd_end(rt, 0);
  while (i < end) {
    int32_t best = -1;
    for (int k = 0; k < nk; ++k) {
      BfFac *f = kids[k];
      while (ptr[k] < f->numRowNodes && nd_start(rt, f->rowNodes[ptr[k]]) < i) ++ptr[k];
      if (ptr[k] >= f->numRowNodes || nd_start(rt, f->rowNodes[ptr[k]]) != i) { *rc = -1; break; }
      int32_t nd = f->rowNodes[ptr[k]];
      if (best < 0 || nd_end(rt, nd) > nd_end(rt, best)) best = nd;
    }
    if (*rc) break;
    BF_VEC_PUSH(cut, best);
    i = nd_end(rt, best);
  }
  if (*rc) { BF_VEC_FREE(cut); return NULL; }
  const size_t nr = cut.n;
  /* per cut node R: Psi* (dense) and the column range of every kid */
  BfComplex **star = calloc(nr, sizeof(BfComplex *));
  int64_t *starCols = calloc(nr, sizeof(int64_t));
  int64_t (*cj)[4][2] = calloc(nr, sizeof(*cj));
  for (int k = 0; k < nk; ++k) ptr[k] = 0;
  for (size_t r = 0; r < nr; ++r) {
    const int32_t R = cut.a[r];
    const int64_t R0 = nd_start(rt, R), R1 = nd_end(rt, R);
    /* kid k's Psi blocks with rows inside R: its row nodes in [R0, R1) */
    int64_t tot = 0;
    size_t first[4], last[4];
    for (int k = 0; k < nk; ++k) {
      BfFac *f = kids[k];
      first[k] = ptr[k];
      while (ptr[k] < f->numRowNodes && nd_end(rt, f->rowNodes[ptr[k]]) <= R1) ++ptr[k];
      last[k] = ptr[k];
      /* Psi blocks are stored one per row node, in row-node order */
      int64_t a = first[k] < last[k] ? f->Psi.b[first[k]].j0 : 0;
      int64_t e = first[k] < last[k] ? f->Psi.b[last[k] - 1].j0 + f->Psi.b[last[k] - 1].n : 0;
      cj[r][k][0] = a;
      cj[r][k][1] = e;
      tot += e - a;
    }
    starCols[r] = tot;
    const int64_t h = R1 - R0;
    BfComplex *P = calloc(nz(h * tot), sizeof(BfComplex));
    int64_t cb = 0;
    for (int k = 0; k < nk; ++k) {
      BfFac *f = kids[k];
      for (size_t q = first[k]; q < last[k]; ++q) {
        const SBlock *b = &f->Psi.b[q];
        const int64_t ro = b->i0 - (R0 - nd_start(rt, 0)), co = cb + b->j0 - cj[r][k][0];
        for (int64_t cc = 0; cc < b->n; ++cc)
          for (int64_t rr = 0; rr < b->m; ++rr)
            P[(co + cc) * h + ro + rr] = b->data ? b->data[cc * b->m + rr] : (rr == cc ? 1.0 : 0.0);
      }
      cb += cj[r][k][1] - cj[r][k][0];
    }
    star[r] = P;
  }
  /* cut every Psi* along the row tree, all cut nodes at once */
  int32_t **fr = malloc(nr * sizeof(int32_t *));
  size_t *frl = malloc(nr * sizeof(size_t));
  for (size_t r = 0; r < nr; ++r) { fr[r] = &cut.a[r]; frl[r] = 1; }
  CutOutVec outs;
  memset(&outs, 0, sizeof(outs));
  int svdRc = 0;
  eps_cut(&s->spec, rt, nr, fr, frl, cut.a, star, starCols, 1, 0, &outs, &svdRc, &s->rank, c);
  free(fr);
  free(frl);
  BfFac *m = NULL;
  if (svdRc) {
    *rc = svdRc;
  } else {
    sort_cutouts(&outs, rt);
    m = calloc(1, sizeof(BfFac));
    m->colNode = &s->ct.nodes[c];
    m->colIndex = c;
    m->rowNodes = malloc((outs.n ? outs.n : 1) * sizeof(int32_t));
    m->numRowNodes = outs.n;
    size_t maxW = 0;
    for (int k = 0; k < nk; ++k) if (kids[k]->numW > maxW) maxW = kids[k]->numW;
    m->numW = maxW + 1;
    m->W = calloc(m->numW, sizeof(SMat));
    /* sizes */
    int64_t psiCols = 0, starTot = 0, w0colsTot = 0;
    for (size_t q = 0; q < outs.n; ++q) psiCols += outs.a[q].k;
    for (size_t r = 0; r < nr; ++r) starTot += starCols[r];
    for (int k = 0; k < nk; ++k) w0colsTot += kids[k]->W[0].cols;
    const int64_t nrows = nd_end(rt, 0) - nd_start(rt, 0);
    smat_init(&m->Psi, nrows, psiCols);
    smat_init(&m->W[0], psiCols, starTot);
    smat_init(&m->W[1], starTot, w0colsTot);
    /* Psi and W0: block diagonal over the cut nodes of their eps cuts */
    int64_t pc = 0;
    int64_t *starBase = calloc(nr, sizeof(int64_t));
    for (size_t r = 1; r < nr; ++r) starBase[r] = starBase[r - 1] + starCols[r - 1];
    for (size_t q = 0; q < outs.n; ++q) {
      CutOut *co = &outs.a[q];
      m->rowNodes[q] = co->nd;
      const int64_t r0 = nd_start(rt, co->nd) - nd_start(rt, 0), h = nd_end(rt, co->nd) - nd_start(rt, co->nd);
      smat_push(&m->Psi, r0, pc, h, co->k, co->Psi);
      smat_push(&m->W[0], pc, starBase[co->owner], co->k, starCols[co->owner], co->W);
      co->Psi = co->W = NULL;
      pc += co->k;
    }
    /* W1: per cut node, the kids' W0 rows for its Psi* columns, block
     * diagonal over the kids */
    for (size_t r = 0; r < nr; ++r) {
      int64_t rowBase = starBase[r], colBase = 0;
      for (int k = 0; k < nk; ++k) {
        smat_append_rows(&m->W[1], &kids[k]->W[0], cj[r][k][0], cj[r][k][1], rowBase, colBase);
        rowBase += cj[r][k][1] - cj[r][k][0];
        colBase += kids[k]->W[0].cols;
      }
    }
    /* trailing factors: block diagonal over the kids (identity where a
     * kid has fewer factors) */
    for (size_t w = 1; w < maxW; ++w) {
      int64_t rows = 0, cols = 0;
      for (int k = 0; k < nk; ++k) {
        BfFac *f = kids[k];
        const SMat *src = w < f->numW ? &f->W[w] : NULL;
        int64_t sz = src ? src->rows : f->W[f->numW - 1].cols;
        rows += sz;
        cols += src ? src->cols : sz;
      }
      smat_init(&m->W[w + 1], rows, cols);
      int64_t ro = 0, co = 0;
      for (int k = 0; k < nk; ++k) {
        BfFac *f = kids[k];
        if (w < f->numW) {
          smat_append_shifted(&m->W[w + 1], &f->W[w], ro, co);
          ro += f->W[w].rows;
          co += f->W[w].cols;
        } else {
          int64_t sz = f->W[f->numW - 1].cols;
          smat_push(&m->W[w + 1], ro, co, sz, sz, NULL);
          ro += sz;
          co += sz;
        }
      }
    }
    free(starBase);
  }
  for (size_t q = 0; q < outs.n; ++q) { free(outs.a[q].Psi); free(outs.a[q].W); }
  BF_VEC_FREE(outs);
  for (size_t r = 0; r < nr; ++r) free(star[r]);
  free(star);
  free(starCols);
  free(cj);
  BF_VEC_FREE(cut);
  return m;
}

static int continue_factorizing(BfFacStreamer *s) {
  while (s->pos < s->numPost && !nd_leaf(s->ct, s->post[s->pos])) {
    const int32_t c = s->post[s->pos];
    int rc = 0;
    BfFac *m = merge_children(s, c, &rc);
    if (!m) return rc ? rc : -1;
    /* drop the children's partial factorizations */
    size_t w = 0;
    for (size_t i = 0; i < s->numFacs; ++i) {
      BfFac *f = s->facs[i];
      int isKid = 0;
      for (int q = 0; q < 4; ++q) isKid |= s->ct.nodes[c].child[q] == f->colIndex;
      if (isKid) fac_free(f);
      else s->facs[w++] = f;
    }
    s->numFacs = w;
    push_fac(s, m);
    ++s->pos;
  }
  return 0;
}

bool bfFacStreamerFeed(BfFacStreamer *s, BfMat *PhiMat) {
  if (!s || !PhiMat || bfFacStreamerIsDone(s) || bfMatGetType(PhiMat) != BF_TYPE_MAT_DENSE_COMPLEX) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return false;
  }
  const int32_t c = s->post[s->pos];
  if (!nd_leaf(s->ct, c)) {
    bfSetError(BF_ERROR_RUNTIME_ERROR);
    return false;
  }
  BfMatDenseComplex *Phi = (BfMatDenseComplex *)PhiMat;
  const int64_t nrows = nd_end(s->rt, 0) - nd_start(s->rt, 0);
  const int64_t ncols = nd_end(s->ct, c) - nd_start(s->ct, c);
  if ((int64_t)PhiMat->numRows != nrows || (int64_t)PhiMat->numCols != ncols) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return false;
  }
  /* column-major copy of Phi */
  BfComplex *A = malloc(nz(nrows * ncols) * sizeof(BfComplex));
  for (int64_t j = 0; j < ncols; ++j)
    for (int64_t i = 0; i < nrows; ++i) A[j * nrows + i] = Phi->data[i * Phi->rowStride + j * Phi->colStride];
  int32_t root = 0;
  int32_t *fr = s->initCut;
  size_t frl = s->initCutLen;
  CutOutVec outs;
  memset(&outs, 0, sizeof(outs));
  int svdRc = 0;
  int bailed = eps_cut(&s->spec, s->rt, 1, &fr, &frl, &root, &A, &ncols, 0, s->spec.bailIfBottomedOut, &outs,
                       &svdRc, &s->rank, c);
  free(A);
  if (svdRc || bailed) {
    for (size_t q = 0; q < outs.n; ++q) { free(outs.a[q].Psi); free(outs.a[q].W); }
    BF_VEC_FREE(outs);
    if (svdRc) bfSetError(svdRc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return false;
  }
  sort_cutouts(&outs, s->rt);
  /* a leaf partial factorization: Psi = diag(Psi_i), W0 = [W_i; ...]
   * (fac.c:84-121) */
  BfFac *f = calloc(1, sizeof(BfFac));
  f->colNode = &s->ct.nodes[c];
  f->colIndex = c;
  f->numRowNodes = outs.n;
  f->rowNodes = malloc((outs.n ? outs.n : 1) * sizeof(int32_t));
  f->numW = 1;
  f->W = calloc(1, sizeof(SMat));
  int64_t tk = 0;
  for (size_t q = 0; q < outs.n; ++q) tk += outs.a[q].k;
  smat_init(&f->Psi, nrows, tk);
  smat_init(&f->W[0], tk, ncols);
  int64_t pc = 0;
  for (size_t q = 0; q < outs.n; ++q) {
    CutOut *co = &outs.a[q];
    f->rowNodes[q] = co->nd;
    const int64_t r0 = nd_start(s->rt, co->nd) - nd_start(s->rt, 0);
    const int64_t h = nd_end(s->rt, co->nd) - nd_start(s->rt, co->nd);
    smat_push(&f->Psi, r0, pc, h, co->k, co->Psi);
    smat_push(&f->W[0], pc, 0, co->k, ncols, co->W);
    pc += co->k;
  }
  BF_VEC_FREE(outs);
  push_fac(s, f);
  ++s->pos;
  int rc = continue_factorizing(s);
  if (rc) {
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return false;
  }
  return true;
}

BfFac *bfFacStreamerGetFac(BfFacStreamer const *s) {
  if (!s || !bfFacStreamerIsDone(s) || s->numFacs != 1) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  return s->facs[0];
}

BfFacSpan *bfFacStreamerGetFacSpan(BfFacStreamer const *s) {
  BfFacSpan *sp = calloc(1, sizeof(BfFacSpan));
  sp->numFacs = s->numFacs;
  sp->fac = malloc((s->numFacs ? s->numFacs : 1) * sizeof(BfFac *));
  for (size_t i = 0; i < s->numFacs; ++i) sp->fac[i] = s->facs[i];
  return sp;
}

void bfFacSpanDelete(BfFacSpan **sp) {
  if (!sp || !*sp) return;
  free((*sp)->fac); /* the facs stay with the streamer */
  free(*sp);
  *sp = NULL;
}

BfSize bfFacSpanGetNumFacs(BfFacSpan const *sp) { return sp ? sp->numFacs : 0; }

/* ------------------------------------------ introspection for checks */

BfSize bfFacGetNumW(BfFac const *f) { return f ? f->numW : 0; }

BfSize bfFacGetRowNodes(BfFac const *f, int32_t *out, BfSize cap) {
  if (!f) return 0;
  for (size_t i = 0; i < f->numRowNodes && i < cap; ++i) out[i] = f->rowNodes[i];
  return f->numRowNodes;
}

/* shape of factor `which` (0 = Psi, 1 + i = W[i]) and its block count */
int bfFacGetFactorShape(BfFac const *f, BfSize which, BfSize *rows, BfSize *cols, BfSize *numBlocks) {
  if (!f || which > f->numW) return -1;
  const SMat *s = which == 0 ? &f->Psi : &f->W[which - 1];
  *rows = (BfSize)s->rows;
  *cols = (BfSize)s->cols;
  *numBlocks = s->nb;
  return 0;
}

/* dense copy (row-major rows x cols) of factor `which` */
int bfFacGetFactorDense(BfFac const *f, BfSize which, BfComplex *out) {
  if (!f || which > f->numW) return -1;
  const SMat *s = which == 0 ? &f->Psi : &f->W[which - 1];
  memset(out, 0, (size_t)(s->rows * s->cols) * sizeof(BfComplex));
  for (size_t i = 0; i < s->nb; ++i) {
    const SBlock *b = &s->b[i];
    for (int64_t c = 0; c < b->n; ++c)
      for (int64_t r = 0; r < b->m; ++r)
        out[(b->i0 + r) * s->cols + b->j0 + c] += b->data ? b->data[c * b->m + r] : (r == c ? 1.0 : 0.0);
  }
  return 0;
}

/* ---------------------------------------- the product on the device */

typedef struct {
  int64_t start, len, blk, xOff;
} PContrib;

typedef struct BfMatHipProduct {
  BfMat super;
  BfSegVec segs;
  BfTermVec terms;
  BfI64Vec stageSeg;
  int numStages;
  int64_t bufRows;
  int64_t entries;
  double *dArena;
  BfkSeg *dSegs;
  BfkTerm *dTerms;
  void *dPing, *dPong;
  size_t bufBytes;
} BfMatHipProduct;

static BfType prod_type(BfMat const *m) { (void)m; return BF_TYPE_MAT_HIP_PRODUCT; }
static BfSize prod_bytes(BfMat const *m) { return (BfSize)((BfMatHipProduct const *)m)->entries * 16; }
static void prod_delete(BfMat **mp) {
  BfMatHipProduct *P = (BfMatHipProduct *)*mp;
  bfk_free(P->dArena);
  bfk_free(P->dSegs);
  bfk_free(P->dTerms);
  bfk_free(P->dPing);
  bfk_free(P->dPong);
  BF_VEC_FREE(P->segs);
  BF_VEC_FREE(P->terms);
  BF_VEC_FREE(P->stageSeg);
  free(P);
  *mp = NULL;
}
static BfMat *prod_mul(BfMat const *A, BfMat const *X);
static BfMatVtable prodVtable = {prod_type, prod_bytes, prod_delete, NULL, prod_mul, NULL};

static int cmp_i64s(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return x < y ? -1 : x > y;
}

/* Segments of one factor: the common refinement of its blocks' row
 * ranges; every block adds a term to each segment it covers. */
static void plan_factor(BfMatHipProduct *P, const SMat *F, const int64_t *dataOff, int xbuf, int ybuf) {
  BF_VEC_PUSH(P->stageSeg, (int64_t)P->segs.n);
  int64_t *bnd = malloc((2 * F->nb + 2) * sizeof(int64_t));
  size_t nb = 0;
  bnd[nb++] = 0;
  bnd[nb++] = F->rows;
  for (size_t i = 0; i < F->nb; ++i) {
    bnd[nb++] = F->b[i].i0;
    bnd[nb++] = F->b[i].i0 + F->b[i].m;
  }
  qsort(bnd, nb, sizeof(int64_t), cmp_i64s);
  size_t nu = 0;
  for (size_t i = 0; i < nb; ++i)
    if (nu == 0 || bnd[i] != bnd[nu - 1]) bnd[nu++] = bnd[i];
  for (size_t sgi = 0; sgi + 1 < nu; ++sgi) {
    const int64_t a = bnd[sgi], e = bnd[sgi + 1];
    BfkSeg sg;
    sg.yOff = a;
    sg.m = (int32_t)(e - a);
    sg.ybuf = ybuf;
    sg.t0 = (int32_t)P->terms.n;
    for (size_t i = 0; i < F->nb; ++i) {
      const SBlock *b = &F->b[i];
      if (b->i0 > a || b->i0 + b->m < e) continue;
      BfkTerm t;
      t.aOff = b->data ? dataOff[i] : -1;
      t.n = (int32_t)b->n;
      t.lda = (int32_t)b->m;
      t.rowOff = (int32_t)(a - b->i0);
      t.xOff = b->j0 + (b->data ? 0 : a - b->i0);
      t.xbuf = xbuf;
      BF_VEC_PUSH(P->terms, t);
    }
    sg.t1 = (int32_t)P->terms.n;
    BF_VEC_PUSH(P->segs, sg);
  }
  free(bnd);
}

/* y = F[0] F[1] ... F[nf-1] x on the device */
static BfMat *product_new(const SMat *const *F, size_t nf) {
  BfMatHipProduct *P = calloc(1, sizeof(BfMatHipProduct));
  bfMatInitBase(&P->super, &prodVtable, (BfSize)F[0]->rows, (BfSize)F[nf - 1]->cols);
  int64_t entries = 0;
  for (size_t f = 0; f < nf; ++f) entries += smat_entries(F[f]);
  P->entries = entries;
  BfComplex *h = malloc((size_t)(entries ? entries : 1) * sizeof(BfComplex));
  int64_t **offs = malloc(nf * sizeof(int64_t *));
  int64_t pos = 0;
  for (size_t f = 0; f < nf; ++f) {
    offs[f] = malloc((F[f]->nb ? F[f]->nb : 1) * sizeof(int64_t));
    for (size_t i = 0; i < F[f]->nb; ++i) {
      const SBlock *b = &F[f]->b[i];
      offs[f][i] = -1;
      if (!b->data) continue;
      offs[f][i] = pos;
      memcpy(h + pos, b->data, (size_t)(b->m * b->n) * sizeof(BfComplex));
      pos += b->m * b->n;
    }
  }
  /* stages: the last factor first; buffers X -> ping -> pong ... -> Y */
  P->numStages = (int)nf;
  P->bufRows = 0;
  for (size_t s = 0; s < nf; ++s) {
    size_t f = nf - 1 - s;
    int xbuf = s == 0 ? BFK_BUF_X : (s % 2 ? BFK_BUF_PING : BFK_BUF_PONG);
    int ybuf = s == nf - 1 ? BFK_BUF_Y : (s % 2 ? BFK_BUF_PONG : BFK_BUF_PING);
    if (s != nf - 1 && F[f]->rows > P->bufRows) P->bufRows = F[f]->rows;
    plan_factor(P, F[f], offs[f], xbuf, ybuf);
  }
  BF_VEC_PUSH(P->stageSeg, (int64_t)P->segs.n);
  int rc = 0;
  P->dArena = bfk_malloc((size_t)(entries ? entries : 1) * 16);
  P->dSegs = bfk_malloc((P->segs.n ? P->segs.n : 1) * sizeof(BfkSeg));
  P->dTerms = bfk_malloc((P->terms.n ? P->terms.n : 1) * sizeof(BfkTerm));
  if (!P->dArena || !P->dSegs || !P->dTerms) rc = -2;
  if (!rc) rc = bfk_h2d(P->dArena, h, (size_t)entries * 16);
  if (!rc) rc = bfk_h2d(P->dSegs, P->segs.a, P->segs.n * sizeof(BfkSeg));
  if (!rc) rc = bfk_h2d(P->dTerms, P->terms.a, P->terms.n * sizeof(BfkTerm));
  for (size_t f = 0; f < nf; ++f) free(offs[f]);
  free(offs);
  free(h);
  if (rc) {
    BfMat *m = &P->super;
    prod_delete(&m);
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  return &P->super;
}

static BfMat *prod_mul(BfMat const *A, BfMat const *Xm) {
  BfMatHipProduct *P = (BfMatHipProduct *)A;
  if (!Xm || bfMatGetType(Xm) != BF_TYPE_MAT_DENSE_COMPLEX || Xm->numRows != A->numCols) {
    bfSetError(BF_ERROR_INCOMPATIBLE_SHAPES);
    return NULL;
  }
  const BfMatDenseComplex *X = (const BfMatDenseComplex *)Xm;
  const int64_t nrhs = (int64_t)Xm->numCols, n = (int64_t)A->numCols, m = (int64_t)A->numRows;
  BfComplex *hx = malloc(nz(n * nrhs) * sizeof(BfComplex));
  for (int64_t i = 0; i < n; ++i)
    for (int64_t j = 0; j < nrhs; ++j) hx[i * nrhs + j] = X->data[i * X->rowStride + j * X->colStride];
  size_t need = (size_t)(P->bufRows ? P->bufRows : 1) * (size_t)nrhs * 16;
  int rc = 0;
  if (need > P->bufBytes) {
    bfk_free(P->dPing);
    bfk_free(P->dPong);
    P->dPing = bfk_malloc(need);
    P->dPong = bfk_malloc(need);
    P->bufBytes = need;
    if (!P->dPing || !P->dPong) rc = -2;
  }
  void *dx = bfk_malloc(nz(n * nrhs) * 16);
  void *dy = bfk_malloc(nz(m * nrhs) * 16);
  if (!dx || !dy) rc = -2;
  if (!rc) rc = bfk_h2d(dx, hx, (size_t)(n * nrhs) * 16);
  if (!rc)
    rc = bfk_apply(P->dSegs, P->stageSeg.a, P->numStages, P->dTerms, P->dArena, dx, P->dPing, P->dPong, dy,
                   (int)nrhs, NULL);
  BfMatDenseComplex *Y = NULL;
  if (!rc) {
    Y = bfMatDenseComplexNew();
    bfMatDenseComplexInit(Y, (BfSize)m, (BfSize)nrhs);
    rc = bfk_d2h(Y->data, dy, (size_t)(m * nrhs) * 16);
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

/* a device product from caller-described factors: factor f is rows[f] x
 * cols[f] with nb[f] column-major blocks (a block without data is an
 * identity block); the data are copied into the product's arena */
BfMat *bfk_product_from_blocks(size_t nf, const int64_t *rows, const int64_t *cols, const size_t *nb,
                               const BfkBlockRef *const *blocks) {
  if (nf == 0) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  SMat *S = calloc(nf, sizeof(SMat));
  const SMat **F = malloc(nf * sizeof(SMat *));
  for (size_t f = 0; f < nf; ++f) {
    smat_init(&S[f], rows[f], cols[f]);
    S[f].b = malloc((nb[f] ? nb[f] : 1) * sizeof(SBlock));
    S[f].nb = S[f].cap = nb[f];
    for (size_t i = 0; i < nb[f]; ++i) {
      const BfkBlockRef *r = &blocks[f][i];
      SBlock blk = {r->i0, r->j0, r->m, r->n, r->data};
      S[f].b[i] = blk;
    }
    F[f] = &S[f];
  }
  BfMat *m = product_new(F, nf);
  for (size_t f = 0; f < nf; ++f) free(S[f].b); /* the data belong to the caller */
  free(S);
  free(F);
  return m;
}

/* Psi W0 ... W_L as one device product (fac.c:40-75) */
BfMat *bfFacGetMat(BfFac const *fac, BfPolicy policy) {
  (void)policy;
  if (!fac) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  const SMat **F = malloc((fac->numW + 1) * sizeof(SMat *));
  F[0] = &fac->Psi;
  for (size_t i = 0; i < fac->numW; ++i) F[i + 1] = &fac->W[i];
  BfMat *m = product_new(F, fac->numW + 1);
  free(F);
  return m;
}

/* ------------------------------------- the product as host containers */

static int cmp_i64(const void *a, const void *b) {
  const int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return (x > y) - (x < y);
}

/* sorted distinct values of v[0..n) in place; returns their count */
static size_t unique_i64(int64_t *v, size_t n) {
  qsort(v, n, sizeof(int64_t), cmp_i64);
  size_t k = 0;
  for (size_t i = 0; i < n; ++i)
    if (k == 0 || v[i] != v[k - 1]) v[k++] = v[i];
  return k;
}

static size_t find_i64(const int64_t *v, size_t n, int64_t x) {
  size_t lo = 0, hi = n;
  while (hi - lo > 1) {
    const size_t mid = (lo + hi) / 2;
    if (v[mid] <= x) lo = mid; else hi = mid;
  }
  return lo;
}

/* one factor as a BfMatBlockCoo on the grid of its blocks' row and column
 * boundaries: a block over several grid cells becomes one dense block per
 * cell (identity blocks as dense ones); blocks on one cell are summed by
 * the container's apply, as overlapping blocks are by the factor's */
static BfMat *smat_host(const SMat *s) {
  int64_t *R = malloc((2 * s->nb + 2) * sizeof(int64_t)), *Cb = malloc((2 * s->nb + 2) * sizeof(int64_t));
  size_t nr = 0, nc = 0;
  R[nr++] = 0;
  R[nr++] = s->rows;
  Cb[nc++] = 0;
  Cb[nc++] = s->cols;
  for (size_t i = 0; i < s->nb; ++i) {
    const SBlock *b = &s->b[i];
    if (b->m == 0 || b->n == 0) continue;
    R[nr++] = b->i0;
    R[nr++] = b->i0 + b->m;
    Cb[nc++] = b->j0;
    Cb[nc++] = b->j0 + b->n;
  }
  nr = unique_i64(R, nr);
  nc = unique_i64(Cb, nc);
  size_t cells = 0;
  for (size_t i = 0; i < s->nb; ++i) {
    const SBlock *b = &s->b[i];
    if (b->m == 0 || b->n == 0) continue;
    cells += (find_i64(R, nr, b->i0 + b->m) - find_i64(R, nr, b->i0)) *
             (find_i64(Cb, nc, b->j0 + b->n) - find_i64(Cb, nc, b->j0));
  }
  BfMatBlockCoo *c = bfMatBlockCooNew();
  bfMatBlockCooInit(c, nr - 1, nc - 1, cells);
  for (size_t i = 0; i < nr; ++i) c->super.rowOffset[i] = (BfSize)R[i];
  for (size_t j = 0; j < nc; ++j) c->super.colOffset[j] = (BfSize)Cb[j];
  size_t k = 0;
  for (size_t i = 0; i < s->nb; ++i) {
    const SBlock *b = &s->b[i];
    if (b->m == 0 || b->n == 0) continue;
    const size_t r0 = find_i64(R, nr, b->i0), r1 = find_i64(R, nr, b->i0 + b->m);
    const size_t c0 = find_i64(Cb, nc, b->j0), c1 = find_i64(Cb, nc, b->j0 + b->n);
    for (size_t ri = r0; ri < r1; ++ri)
      for (size_t ci = c0; ci < c1; ++ci, ++k) {
        const int64_t m = R[ri + 1] - R[ri], n = Cb[ci + 1] - Cb[ci];
        const int64_t di = R[ri] - b->i0, dj = Cb[ci] - b->j0;
        BfMatDenseComplex *D = bfMatDenseComplexNew();
        bfMatDenseComplexInit(D, (BfSize)m, (BfSize)n);
        for (int64_t r = 0; r < m; ++r)
          for (int64_t q = 0; q < n; ++q)
            D->data[r * n + q] = b->data ? b->data[(dj + q) * b->m + di + r] : (di + r == dj + q ? 1.0 : 0.0);
        c->rowInd[k] = ri;
        c->colInd[k] = ci;
        c->super.block[k] = bfMatDenseComplexToMat(D);
        c->ownsBlock[k] = true;
      }
  }
  free(R);
  free(Cb);
  return bfMatBlockCooToMat(c);
}

BfMatProduct *bfFacGetMatProduct(BfFac const *fac, BfPolicy policy) {
  (void)policy;
  if (!fac) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  BfMatProduct *P = bfMatProductNew();
  bfMatProductInit(P);
  bfMatProductPostMultiply(P, smat_host(&fac->Psi));
  for (size_t i = 0; i < fac->numW; ++i) bfMatProductPostMultiply(P, smat_host(&fac->W[i]));
  return P;
}

/* the partial factorizations side by side: [F_1 F_2 ...] as one chain,
 * [Psi_1 Psi_2 ...] then block diagonals of the W factors, shorter
 * chains padded with identities (fac_span.c:128-157) */
BfMat *bfFacSpanGetMat(BfFacSpan const *sp, BfPolicy policy) {
  (void)policy;
  if (!sp || sp->numFacs == 0) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  size_t L = 0;
  for (size_t i = 0; i < sp->numFacs; ++i) if (sp->fac[i]->numW > L) L = sp->fac[i]->numW;
  SMat *T = calloc(L + 1, sizeof(SMat));
  int64_t rows = sp->fac[0]->Psi.rows, cols = 0;
  for (size_t i = 0; i < sp->numFacs; ++i) cols += sp->fac[i]->Psi.cols;
  smat_init(&T[0], rows, cols);
  int64_t co = 0;
  for (size_t i = 0; i < sp->numFacs; ++i) {
    smat_append_shifted(&T[0], &sp->fac[i]->Psi, 0, co);
    co += sp->fac[i]->Psi.cols;
  }
  for (size_t w = 0; w < L; ++w) {
    int64_t r = 0, c = 0;
    for (size_t i = 0; i < sp->numFacs; ++i) {
      BfFac *f = sp->fac[i];
      int64_t sz = f->W[f->numW - 1].cols;
      r += w < f->numW ? f->W[w].rows : sz;
      c += w < f->numW ? f->W[w].cols : sz;
    }
    smat_init(&T[w + 1], r, c);
    int64_t ro = 0, cc = 0;
    for (size_t i = 0; i < sp->numFacs; ++i) {
      BfFac *f = sp->fac[i];
      if (w < f->numW) {
        smat_append_shifted(&T[w + 1], &f->W[w], ro, cc);
        ro += f->W[w].rows;
        cc += f->W[w].cols;
      } else {
        int64_t sz = f->W[f->numW - 1].cols;
        smat_push(&T[w + 1], ro, cc, sz, sz, NULL);
        ro += sz;
        cc += sz;
      }
    }
  }
  const SMat **F = malloc((L + 1) * sizeof(SMat *));
  for (size_t i = 0; i <= L; ++i) F[i] = &T[i];
  BfMat *m = product_new(F, L + 1);
  free(F);
  for (size_t i = 0; i <= L; ++i) smat_free(&T[i]);
  free(T);
  return m;
}
