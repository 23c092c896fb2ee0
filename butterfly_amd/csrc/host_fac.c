/*
 * The device factorization object: builds the plan, runs the sampling and
 * compression kernels into one HBM arena, lays out the apply stages and
 * applies.  Replaces bfFacHelm2MakeMultilevel (fac_helm2.c:943-1002) and,
 * for its result, bfMatMul (mat.c:183-185).
 */
#define _DEFAULT_SOURCE 1
#include <math.h>

#include <complex.h>
#include <pthread.h>

#include "bf_host.h"

static BfType hipbf_get_type(BfMat const *m) {
  (void)m;
  return BF_TYPE_MAT_HIP_BUTTERFLY;
}

BfMatVtable bfMatHipButterflyVtable = {
    .GetType = hipbf_get_type,
    .NumBytes = bfMatHipButterflyNumBytes,
    .Delete = bfMatHipButterflyDelete,
    .Copy = NULL,
    .Mul = bfMatHipButterflyMul,
    .PermuteRows = NULL,
};

void bfFacHelm2OptionsInitDefault(BfFacHelm2Options *opts) {
  opts->tol = 1e-15;
  opts->rankConst = 1;
  opts->maxDenseSize = 128 * 128;
  opts->passThrough = false;
  opts->pinvRcond = 0;
}

static void free_plan(BfMatHipButterfly *F) {
  BF_VEC_FREE(F->leaves);
  BF_VEC_FREE(F->factors);
  BF_VEC_FREE(F->blocks);
  BF_VEC_FREE(F->offs);
  BF_VEC_FREE(F->lstsqJobs);
  BF_VEC_FREE(F->kernelJobs);
  BF_VEC_FREE(F->segs);
  BF_VEC_FREE(F->terms);
  BF_VEC_FREE(F->stageSeg);
}

static void free_device(BfMatHipButterfly *F) {
  bfk_free(F->dArena);
  bfk_free(F->dArena32);
  bfk_free(F->dArena32lo);
  bfk_free(F->dSrcXY);
  bfk_free(F->dTgtXY);
  bfk_free(F->dSrcN);
  bfk_free(F->dTgtN);
  bfk_graph_destroy(F->graphExec);
  F->graphExec = NULL;
  for (int p = 0; p < 3; ++p)
    for (int c = 0; c < BF_NRHS_CLASSES; ++c) {
      free(F->tuneCfg[p][c]);
      F->tuneCfg[p][c] = NULL;
    }
  bfk_free(F->dSegs);
  bfk_free(F->dTerms);
  bfk_free(F->dPing);
  bfk_free(F->dPong);
  F->dArena = NULL;
  F->dArena32 = F->dArena32lo = NULL;
  F->dSrcXY = F->dTgtXY = F->dSrcN = F->dTgtN = NULL;
  F->dSegs = NULL;
  F->dTerms = NULL;
  F->dPing = F->dPong = NULL;
}

void bfMatHipButterflyDelete(BfMat **mat) {
  if (!mat || !*mat) return;
  BfMatHipButterfly *F = (BfMatHipButterfly *)*mat;
  free_device(F);
  free_plan(F);
  free(F);
  *mat = NULL;
}

BfSize bfMatHipButterflyNumBytes(BfMat const *mat) {
  BfMatHipButterfly const *F = (BfMatHipButterfly const *)mat;
  return (BfSize)F->arenaEntries * 16;
}

static int64_t *offs_at(BfMatHipButterfly *F, int64_t start) { return &F->offs.a[start]; }

typedef struct {
  int64_t start, len;
  int64_t blk;   /* block index, or -1 - leafIdx for a dense leaf */
  int64_t xOff;
  int32_t xbuf;
} Contrib;

static int cmp_i64(const void *a, const void *b) {
  int64_t x = *(const int64_t *)a, y = *(const int64_t *)b;
  return x < y ? -1 : x > y;
}

static int32_t stage_buf(int s) { return (s % 2 == 0) ? BFK_BUF_PING : BFK_BUF_PONG; }

/* Stage layout: butterfly b with nf factors runs factor f at stage
 * Fmax - nf + f, so every last factor and every dense leaf lands in the
 * final stage, which writes Y once per row segment (no atomics, a fixed
 * summation order).  Earlier stages write ping (even) or pong (odd). */
int bfApplyPlanBuild(BfMatHipButterfly *F) {
  int Fmax = 1;
  for (size_t l = 0; l < F->leaves.n; ++l)
    if (F->leaves.a[l].isButterfly && F->leaves.a[l].numFactors > Fmax) Fmax = F->leaves.a[l].numFactors;
  F->numStages = Fmax;
  int64_t *fill = calloc(Fmax, sizeof(int64_t));
  /* region of every non-last factor */
  int64_t *region = malloc((F->factors.n + 1) * sizeof(int64_t));
  for (size_t l = 0; l < F->leaves.n; ++l) {
    BfHipLeaf *lf = &F->leaves.a[l];
    if (!lf->isButterfly) continue;
    for (int f = 0; f < lf->numFactors - 1; ++f) {
      int s = Fmax - lf->numFactors + f;
      BfHipFactor *fa = &F->factors.a[lf->firstFactor + f];
      region[lf->firstFactor + f] = fill[s];
      fill[s] += offs_at(F, fa->rowOffStart)[fa->numBlockRows];
    }
  }
  F->bufRows[0] = F->bufRows[1] = 0;
  for (int s = 0; s < Fmax; ++s) {
    int b = s % 2;
    if (fill[s] > F->bufRows[b]) F->bufRows[b] = fill[s];
  }
  double facBytes = 0, vecBytes = 0;
  F->segs.n = F->terms.n = F->stageSeg.n = 0;
  BF_VEC(int64_t) rowFirst, colCount, rowOrder;
  memset(&rowFirst, 0, sizeof(rowFirst));
  memset(&colCount, 0, sizeof(colCount));
  memset(&rowOrder, 0, sizeof(rowOrder));
  for (int s = 0; s < Fmax - 1; ++s) {
    BF_VEC_PUSH(F->stageSeg, (int64_t)F->segs.n);
    for (size_t l = 0; l < F->leaves.n; ++l) {
      BfHipLeaf *lf = &F->leaves.a[l];
      if (!lf->isButterfly) continue;
      int f = s - (Fmax - lf->numFactors);
      if (f < 0 || f >= lf->numFactors - 1) continue;
      int64_t fi = lf->firstFactor + f;
      BfHipFactor *fa = &F->factors.a[fi];
      int64_t *ro = offs_at(F, fa->rowOffStart), *co = offs_at(F, fa->colOffStart);
      vecBytes += 16.0 * (double)(ro[fa->numBlockRows] + co[fa->numBlockCols]);
      const int64_t bbeg = fa->firstBlock, bend = fa->firstBlock + fa->numBlocks;
      /* rows in the order of their first input column block (counting
       * sort, stable): the rows that read the same inputs become adjacent
       * segments, which the many-RHS apply keeps on one XCD */
      rowFirst.n = 0;
      for (int64_t b = bbeg; b < bend; ++b)
        if (b == bbeg || F->blocks.a[b].rowInd != F->blocks.a[b - 1].rowInd) BF_VEC_PUSH(rowFirst, b);
      const int64_t nrow = (int64_t)rowFirst.n, ncol = fa->numBlockCols;
      colCount.n = 0;
      for (int64_t c = 0; c <= ncol; ++c) BF_VEC_PUSH(colCount, 0);
      for (int64_t r = 0; r < nrow; ++r) colCount.a[F->blocks.a[rowFirst.a[r]].colInd + 1]++;
      for (int64_t c = 0; c < ncol; ++c) colCount.a[c + 1] += colCount.a[c];
      rowOrder.n = 0;
      for (int64_t r = 0; r < nrow; ++r) BF_VEC_PUSH(rowOrder, 0);
      for (int64_t r = 0; r < nrow; ++r) rowOrder.a[colCount.a[F->blocks.a[rowFirst.a[r]].colInd]++] = rowFirst.a[r];
      for (int64_t ri = 0; ri < nrow; ++ri) {
        int64_t b = rowOrder.a[ri];
        int32_t row = F->blocks.a[b].rowInd;
        BfkSeg sg;
        sg.yOff = region[fi] + ro[row];
        sg.m = (int32_t)(ro[row + 1] - ro[row]);
        sg.ybuf = stage_buf(s);
        sg.t0 = (int32_t)F->terms.n;
        for (; b < bend && F->blocks.a[b].rowInd == row; ++b) {
          BfHipBlock *bl = &F->blocks.a[b];
          BfkTerm t;
          t.aOff = bl->dataOff;
          t.n = bl->n;
          t.lda = bl->m;
          t.rowOff = 0;
          if (f == 0) {
            t.xOff = lf->col0 + co[bl->colInd];
            t.xbuf = BFK_BUF_X;
          } else {
            t.xOff = region[fi - 1] + co[bl->colInd];
            t.xbuf = stage_buf(s - 1);
          }
          if (bl->dataOff >= 0) facBytes += 16.0 * bl->m * bl->n;
          BF_VEC_PUSH(F->terms, t);
        }
        sg.t1 = (int32_t)F->terms.n;
        BF_VEC_PUSH(F->segs, sg);
      }
    }
  }
  BF_VEC_FREE(rowFirst);
  BF_VEC_FREE(colCount);
  BF_VEC_FREE(rowOrder);
  /* final stage */
  BF_VEC_PUSH(F->stageSeg, (int64_t)F->segs.n);
  BF_VEC(Contrib) cons;
  memset(&cons, 0, sizeof(cons));
  for (size_t l = 0; l < F->leaves.n; ++l) {
    BfHipLeaf *lf = &F->leaves.a[l];
    if (!lf->isButterfly) {
      Contrib c = {lf->row0, lf->m, -1 - (int64_t)l, lf->col0, BFK_BUF_X};
      BF_VEC_PUSH(cons, c);
      facBytes += 16.0 * lf->m * lf->n;
      vecBytes += 16.0 * (lf->m + lf->n);
      continue;
    }
    int64_t fi = lf->firstFactor + lf->numFactors - 1;
    BfHipFactor *fa = &F->factors.a[fi];
    int64_t *ro = offs_at(F, fa->rowOffStart), *co = offs_at(F, fa->colOffStart);
    vecBytes += 16.0 * (double)(ro[fa->numBlockRows] + co[fa->numBlockCols]);
    for (int64_t b = fa->firstBlock; b < fa->firstBlock + fa->numBlocks; ++b) {
      BfHipBlock *bl = &F->blocks.a[b];
      Contrib c = {lf->row0 + ro[bl->rowInd], bl->m, b, region[fi - 1] + co[bl->colInd], stage_buf(Fmax - 2)};
      BF_VEC_PUSH(cons, c);
      if (bl->dataOff >= 0) facBytes += 16.0 * bl->m * bl->n;
    }
  }
  int64_t *bnd = malloc((2 * cons.n + 2) * sizeof(int64_t));
  size_t nb = 0;
  bnd[nb++] = 0;
  bnd[nb++] = (int64_t)F->numTgt;
  for (size_t i = 0; i < cons.n; ++i) {
    bnd[nb++] = cons.a[i].start;
    bnd[nb++] = cons.a[i].start + cons.a[i].len;
  }
  qsort(bnd, nb, sizeof(int64_t), cmp_i64);
  size_t nu = 0;
  for (size_t i = 0; i < nb; ++i)
    if (nu == 0 || bnd[i] != bnd[nu - 1]) bnd[nu++] = bnd[i];
  size_t numSeg = nu - 1;
  int64_t *cnt = calloc(numSeg + 1, sizeof(int64_t));
  for (size_t i = 0; i < cons.n; ++i) {
    int64_t a = cons.a[i].start, e = a + cons.a[i].len;
    size_t lo = 0, hi = numSeg; /* first boundary == a */
    while (lo < hi) { size_t mid = (lo + hi) / 2; if (bnd[mid] < a) lo = mid + 1; else hi = mid; }
    for (size_t sgi = lo; sgi < numSeg && bnd[sgi] < e; ++sgi) cnt[sgi + 1]++;
  }
  for (size_t i = 0; i < numSeg; ++i) cnt[i + 1] += cnt[i];
  size_t termBase = F->terms.n;
  size_t totalTerms = termBase + (size_t)cnt[numSeg];
  if (F->terms.cap < totalTerms) {
    void *na = realloc(F->terms.a, totalTerms * sizeof(BfkTerm));
    if (!na) abort();
    F->terms.a = na;
    F->terms.cap = totalTerms;
  }
  F->terms.n = totalTerms;
  int64_t *pos = malloc(numSeg * sizeof(int64_t));
  for (size_t i = 0; i < numSeg; ++i) pos[i] = cnt[i];
  for (size_t i = 0; i < cons.n; ++i) {
    Contrib *c = &cons.a[i];
    int64_t a = c->start, e = a + c->len;
    size_t lo = 0, hi = numSeg;
    while (lo < hi) { size_t mid = (lo + hi) / 2; if (bnd[mid] < a) lo = mid + 1; else hi = mid; }
    for (size_t sgi = lo; sgi < numSeg && bnd[sgi] < e; ++sgi) {
      int64_t rOff = bnd[sgi] - a;
      BfkTerm t;
      if (c->blk < 0) {
        BfHipLeaf *lf = &F->leaves.a[-1 - c->blk];
        t.aOff = lf->denseOff;
        t.n = (int32_t)lf->n;
        t.lda = (int32_t)lf->m;
        t.rowOff = (int32_t)rOff;
        t.xOff = c->xOff;
      } else {
        BfHipBlock *bl = &F->blocks.a[c->blk];
        t.aOff = bl->dataOff;
        t.n = bl->n;
        t.lda = bl->m;
        t.rowOff = (int32_t)rOff;
        t.xOff = bl->dataOff >= 0 ? c->xOff : c->xOff + rOff;
      }
      t.xbuf = c->xbuf;
      F->terms.a[termBase + pos[sgi]++] = t;
    }
  }
  for (size_t i = 0; i < numSeg; ++i) {
    BfkSeg sg;
    sg.yOff = bnd[i];
    sg.m = (int32_t)(bnd[i + 1] - bnd[i]);
    sg.ybuf = BFK_BUF_Y;
    sg.t0 = (int32_t)(termBase + cnt[i]);
    sg.t1 = (int32_t)(termBase + cnt[i + 1]);
    BF_VEC_PUSH(F->segs, sg);
  }
  BF_VEC_PUSH(F->stageSeg, (int64_t)F->segs.n);
  F->applyFactorBytes = facBytes;
  F->stats[11] = facBytes;
  F->stats[12] = vecBytes;
  free(pos);
  free(cnt);
  free(bnd);
  BF_VEC_FREE(cons);
  free(fill);
  free(region);
  if (F->terms.n > (size_t)INT32_MAX) return -1;
  return 0;
}

static int upload(void **dst, const void *src, size_t bytes) {
  *dst = bfk_malloc(bytes);
  if (!*dst) return -2;
  if (bytes && bfk_h2d(*dst, src, bytes)) return -1;
  return 0;
}

/* the same on the copy stream: does not wait for the queued kernels */
static int upload_side(void **dst, const void *src, size_t bytes) {
  *dst = bfk_malloc(bytes);
  if (!*dst) return -2;
  if (bytes && bfk_h2d_side(*dst, src, bytes)) return -1;
  return 0;
}

/* the apply of a factorization whose plan and arena are in place (a loaded
 * one): stage segments and terms, uploaded */
int bfMatHipButterflyFinish(BfMatHipButterfly *F) {
  double t0 = bfWallTime();
  if (bfApplyPlanBuild(F)) return -1;
  int rc = upload((void **)&F->dSegs, F->segs.a, F->segs.n * sizeof(BfkSeg));
  if (!rc) rc = upload((void **)&F->dTerms, F->terms.a, F->terms.n * sizeof(BfkTerm));
  int64_t nL = 0, nK = 0, nI = 0, nD = 0, nB = 0;
  for (size_t b = 0; b < F->blocks.n; ++b) {
    nL += F->blocks.a[b].kind == BF_BLOCK_LSTSQ;
    nK += F->blocks.a[b].kind == BF_BLOCK_KERNEL;
    nI += F->blocks.a[b].kind == BF_BLOCK_IDENTITY;
  }
  for (size_t l = 0; l < F->leaves.n; ++l) {
    nB += F->leaves.a[l].isButterfly;
    nD += !F->leaves.a[l].isButterfly;
  }
  F->stats[3] = (double)F->arenaEntries;
  F->stats[4] = (double)nL;
  F->stats[5] = (double)nK;
  F->stats[6] = (double)nI;
  F->stats[7] = (double)nD;
  F->stats[8] = (double)nB;
  F->stats[9] = (double)F->numStages;
  F->stats[15] = bfWallTime() - t0;
  F->statsCounted = 1;
  return rc;
}

static double *tree_coords(BfQuadtree const *t) {
  BfSize n = t->points->size;
  double *xy = malloc(2 * n * sizeof(double));
  for (BfSize i = 0; i < n; ++i) {
    BfSize j = t->super.perm->index[i];
    xy[2 * i] = t->points->data[j][0];
    xy[2 * i + 1] = t->points->data[j][1];
  }
  return xy;
}

/* unit normals of the tree's points in tree order (NULL without normals) */
static double *tree_normals(BfQuadtree const *t) {
  if (!t->unitNormals) return NULL;
  const BfSize n = t->points->size;
  double *nv = malloc(2 * (n ? n : 1) * sizeof(double));
  for (BfSize i = 0; i < n; ++i) {
    const BfSize j = t->super.perm->index[i];
    nv[2 * i] = t->unitNormals->data[j][0];
    nv[2 * i + 1] = t->unitNormals->data[j][1];
  }
  return nv;
}

/* the layers that read source normals (D and the combined field) and the
 * one that reads target normals (S'), layer_pot.h:40-48 */
static int uses_src_normals(BfLayerPotential lp) {
  return lp == BF_LAYER_POTENTIAL_PV_DOUBLE || lp == BF_LAYER_POTENTIAL_COMBINED_FIELD;
}
static int uses_tgt_normals(BfLayerPotential lp) { return lp == BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE; }

/* the kernels' form of a layer potential: the kernel blocks (last factors,
 * values blocks, dense leaves) sample the layer itself */
static BfkLayer layer_of(BfHelm2 const *h) {
  BfkLayer L;
  memset(&L, 0, sizeof(L));
  switch (h->layerPot) {
  case BF_LAYER_POTENTIAL_PV_DOUBLE: L.layer = 1; break;
  case BF_LAYER_POTENTIAL_COMBINED_FIELD: L.layer = 2; break;
  case BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE: L.layer = 3; break;
  default: L.layer = 0;
  }
  L.ar = creal(h->alpha);
  L.ai = cimag(h->alpha);
  L.br = creal(h->beta);
  L.bi = cimag(h->beta);
  return L;
}

/* the layer of the re-expansions (BF_PROXY_LAYER_POT, layer_pot.h:63-69):
 * the layer itself for S, D and the combined field; S' drops its target
 * normal derivative and re-expands with the single layer (its proxies hold
 * single-layer charges, fac_helm2.c:58-59, :230-231), so the target normals
 * enter only through the kernel blocks */
static BfkLayer proxy_layer_of(BfHelm2 const *h) {
  BfkLayer L = layer_of(h);
  if (L.layer == 3) L.layer = 0;
  return L;
}

typedef BF_VEC(size_t) BfSizeVec;

/* Stable counting sort of jobs by LDS class (class c launches with c * 4 KiB):
 * order[] lists the jobs class by class, in plan order inside a class;
 * gs gets each non-empty class's first position (plus a final nj) and gb
 * its byte size. */
static void group_by_class(const uint32_t *cls, size_t nj, int32_t *order, BfI64Vec *gs, BfSizeVec *gb,
                           int descending) {
  uint32_t maxCls = 0;
  for (size_t i = 0; i < nj; ++i)
    if (cls[i] > maxCls) maxCls = cls[i];
  /* key: the class, or its mirror for descending order */
  size_t *start = calloc((size_t)maxCls + 2, sizeof(size_t));
#define KEY(c) (descending ? maxCls - (c) : (c))
  for (size_t i = 0; i < nj; ++i) ++start[KEY(cls[i]) + 1];
  for (size_t c = 0; c <= maxCls; ++c) start[c + 1] += start[c];
  gs->n = gb->n = 0;
  for (size_t q = 0; q <= maxCls; ++q)
    if (start[q + 1] > start[q]) {
      BF_VEC_PUSH(*gs, (int64_t)start[q]);
      BF_VEC_PUSH(*gb, (size_t)KEY(q) << 12);
    }
  BF_VEC_PUSH(*gs, (int64_t)nj);
  for (size_t i = 0; i < nj; ++i) order[start[KEY(cls[i])]++] = (int32_t)i;
#undef KEY
  free(start);
}

/* the apply plan, built on a helper thread while the main thread uploads
 * and launches the compression */
typedef struct {
  BfMatHipButterfly *F;
  int rc;
  double seconds;
} ApplyPlanTask;

static void *apply_plan_task(void *arg) {
  ApplyPlanTask *t = arg;
  double t0 = bfWallTime();
  t->rc = bfApplyPlanBuild(t->F) ? -1 : 0;
  t->seconds = bfWallTime() - t0;
  return NULL;
}

/* Parallel stable counting sort of the re-expansion jobs by LDS class,
 * largest class first, writing the jobs themselves (no order array): each
 * of NT threads classes a contiguous range and counts it, the class starts
 * are summed per thread, then each thread scatters its range.  Within a
 * class the jobs keep plan order, as group_by_class(..., 1) with a gather
 * would give. */
enum { CLASS_NT = 16 };

typedef struct {
  const BfkLstsqJob *src;
  BfkLstsqJob *dst;
  uint32_t *cls;
  size_t i0, i1;
  uint32_t maxCls;
  size_t *pos; /* per class: the next slot of this thread (phase 3), counts (phase 2) */
  int phase;
} ClassTask;

static void *class_task(void *arg) {
  ClassTask *t = arg;
  if (t->phase == 1) {
    uint32_t mx = 0;
    for (size_t i = t->i0; i < t->i1; ++i) {
      const uint32_t c = (uint32_t)((BF_LSTSQ_UNITS(t->src[i].p, t->src[i].n) * 16 + 4095) >> 12);
      t->cls[i] = c;
      if (c > mx) mx = c;
    }
    t->maxCls = mx;
  } else if (t->phase == 2) {
    for (size_t i = t->i0; i < t->i1; ++i) ++t->pos[t->cls[i]];
  } else {
    for (size_t i = t->i0; i < t->i1; ++i) t->dst[t->pos[t->cls[i]]++] = t->src[i];
  }
  return NULL;
}

static void class_phase(ClassTask *t, int nt, int phase) {
  pthread_t th[CLASS_NT];
  int ok[CLASS_NT] = {0};
  for (int q = 0; q < nt; ++q) t[q].phase = phase;
  for (int q = 1; q < nt; ++q) {
    ok[q] = pthread_create(&th[q], NULL, class_task, &t[q]) == 0;
    if (!ok[q]) class_task(&t[q]);
  }
  class_task(&t[0]);
  for (int q = 1; q < nt; ++q)
    if (ok[q]) pthread_join(th[q], NULL);
}

/* dst gets src's jobs class by class (largest first); gs gets each
 * non-empty class's first position plus a final nj, gb its byte size */
static int class_sort_jobs(const BfkLstsqJob *src, size_t nj, BfkLstsqJob *dst, BfI64Vec *gs, BfSizeVec *gb) {
  ClassTask t[CLASS_NT];
  const int nt = nj < ((size_t)1 << 16) ? 1 : CLASS_NT;
  uint32_t *cls = malloc((nj ? nj : 1) * sizeof(uint32_t));
  if (!cls) return -2;
  bfHugePages(cls, nj * sizeof(uint32_t));
  memset(t, 0, sizeof(t));
  for (int q = 0; q < nt; ++q) {
    t[q].src = src;
    t[q].dst = dst;
    t[q].cls = cls;
    t[q].i0 = nj * (size_t)q / (size_t)nt;
    t[q].i1 = nj * (size_t)(q + 1) / (size_t)nt;
  }
  class_phase(t, nt, 1);
  uint32_t maxCls = 0;
  for (int q = 0; q < nt; ++q)
    if (t[q].maxCls > maxCls) maxCls = t[q].maxCls;
  size_t *pos = calloc((size_t)nt * (maxCls + 1), sizeof(size_t));
  if (!pos) {
    free(cls);
    return -2;
  }
  for (int q = 0; q < nt; ++q) t[q].pos = pos + (size_t)q * (maxCls + 1);
  class_phase(t, nt, 2);
  /* counts -> each thread's first slot per class, classes descending */
  gs->n = gb->n = 0;
  size_t run = 0;
  for (int64_t c = maxCls; c >= 0; --c) {
    const size_t first = run;
    for (int q = 0; q < nt; ++q) {
      const size_t cnt = t[q].pos[c];
      t[q].pos[c] = run;
      run += cnt;
    }
    if (run > first) {
      BF_VEC_PUSH(*gs, (int64_t)first);
      BF_VEC_PUSH(*gb, (size_t)c << 12);
    }
  }
  BF_VEC_PUSH(*gs, (int64_t)nj);
  class_phase(t, nt, 3);
  free(pos);
  free(cls);
  return 0;
}

/* Sampling, then the compression: the re-expansion jobs are grouped by
 * the LDS they need (4 KiB classes, launched largest first over four
 * streams) and uploaded in chunks that are launched as they arrive; the
 * apply plan is built on a helper thread meanwhile and its tables go up
 * on the copy stream.  Returns with the kernels running (bfk_sync waits). */
static int run_factor_kernels(BfMatHipButterfly *F, double *tApplyPlan) {
  int rc;
  *tApplyPlan = 0;
  const double tk = bfWallTime();
  size_t arenaBytes = (size_t)F->arenaEntries * 16;
  F->dArena = bfk_malloc(arenaBytes);
  if (!F->dArena) return -2;
  bfTrace("factor: arena allocated", tk);
  void *dK = NULL;
  if ((rc = upload(&dK, F->kernelJobs.a, F->kernelJobs.n * sizeof(BfkKernelJob)))) return rc;
  const BfkLayer layer = layer_of(&F->helm), proxyLayer = proxy_layer_of(&F->helm);
  rc = bfk_sample(dK, (int64_t)F->kernelJobs.n, F->dSrcXY, F->dTgtXY, F->dSrcN, F->dTgtN, layer, F->dArena);
  bfk_free(dK);
  if (rc) return rc;
  bfTrace("factor: sampling launched", tk);
  const size_t nj = F->lstsqJobs.n;
  /* the jobs in class order (largest first) */
  BfI64Vec gs;
  BfSizeVec gb;
  memset(&gs, 0, sizeof(gs));
  memset(&gb, 0, sizeof(gb));
  BfkLstsqJob *sorted = malloc((nj ? nj : 1) * sizeof(BfkLstsqJob));
  bfHugePages(sorted, nj * sizeof(BfkLstsqJob));
  if (!sorted || class_sort_jobs(F->lstsqJobs.a, nj, sorted, &gs, &gb)) {
    free(sorted);
    BF_VEC_FREE(gs);
    BF_VEC_FREE(gb);
    return -2;
  }
  bfTrace("factor: jobs sorted by class", tk);
  /* the apply plan is host work: build it on a helper thread while this
   * thread uploads the jobs and the kernels run */
  ApplyPlanTask task = {F, 0, 0};
  pthread_t th;
  const int threaded = pthread_create(&th, NULL, apply_plan_task, &task) == 0;
  if (!threaded) apply_plan_task(&task);
  /* uploaded in chunks on the copy stream; each uploaded range is launched
   * at once, so the upload (several GB for large or batched problems)
   * overlaps the kernels */
  BfkLstsqJob *dL = bfk_malloc((nj ? nj : 1) * sizeof(BfkLstsqJob));
  if (!dL) rc = -2;
  /* jobs above the LDS capacity share one scratch buffer */
  size_t bigBytes = 0;
  int64_t numBig = 0;
  for (size_t g = 0; g + 1 < gs.n; ++g)
    if (gb.a[g] > bfk_lds_cap()) {
      if (gb.a[g] > bigBytes) bigBytes = gb.a[g];
      numBig += gs.a[g + 1] - gs.a[g];
    }
  if (!rc) rc = bfk_lstsq_begin(bigBytes, numBig);
  const int begun = rc == 0;
  /* jobs per upload: small at first so that the kernels start early, then
   * doubling up to 4M */
  size_t chunk = (size_t)1 << 18;
  size_t up = 0, launched = 0, g = 0;
  while (!rc && up < nj) {
    const size_t n = nj - up < chunk ? nj - up : chunk;
    if (chunk < ((size_t)1 << 22)) chunk *= 2;
    rc = bfk_h2d_side(dL + up, sorted + up, n * sizeof(BfkLstsqJob));
    up += n;
    /* what is uploaded: the rest of the current class, the classes that
     * are complete, and the uploaded start of the next */
    while (!rc && g + 1 < gs.n && launched < up) {
      /* classes [g, gEnd): one class, or all the classes above the LDS
       * capacity (they come first and share the scratch, whose stride fits
       * the largest, so they go out in one launch: per class they were ~40
       * launches of a few dozen workgroups each) */
      size_t gEnd = g + 1, lds = gb.a[g];
      if (gb.a[g] > bfk_lds_cap()) {
        while (gEnd + 1 < gs.n && gb.a[gEnd] > bfk_lds_cap()) ++gEnd;
        lds = bigBytes;
      }
      const size_t classEnd = (size_t)gs.a[gEnd];
      const size_t end = classEnd < up ? classEnd : up;
      rc = bfk_lstsq_launch(dL + launched, NULL, (int64_t)(end - launched), lds, F->dSrcXY, F->dSrcN, proxyLayer,
                            F->dArena);
      launched = end;
      if (launched == classEnd) g = gEnd;
    }
  }
  bfTrace("factor: compression launched", tk);
  /* close the launch sequence even after a failed launch, so the side
   * streams are joined and the scratch is released */
  if (begun) {
    const int rc2 = bfk_lstsq_end();
    if (!rc) rc = rc2;
  }
  free(sorted);
  if (threaded) pthread_join(th, NULL);
  bfTrace("factor: apply plan joined", tk);
  *tApplyPlan = task.seconds;
  if (!rc && task.rc) rc = task.rc;
  /* the apply tables go up while the compression runs */
  if (!rc) rc = upload_side((void **)&F->dSegs, F->segs.a, F->segs.n * sizeof(BfkSeg));
  if (!rc) rc = upload_side((void **)&F->dTerms, F->terms.a, F->terms.n * sizeof(BfkTerm));
  bfTrace("factor: apply tables up", tk);
  /* the launches are still running: release their tables after the next sync */
  bfk_free_async(dL);
  BF_VEC_FREE(gs);
  BF_VEC_FREE(gb);
  return rc;
}

/* Everything after planning: coordinates up, sampling and compression
 * launched, the apply plan built on the host and its tables uploaded
 * meanwhile, then the statistics.  Takes ownership of sxy / txy (tree-ordered
 * coordinates), snrm and tnrm (tree-ordered source / target normals or
 * NULL); on failure deletes F and returns NULL. */
static BfMat *factor_from_plan(BfMatHipButterfly *F, double *sxy, double *txy, double *snrm, double *tnrm,
                               double t0, double t1) {
  /* the normal-dependent layers read the normals in every sampling and
   * compression job: no kernel runs without them */
  const BfLayerPotential lp = F->helm.layerPot;
  int rc = ((uses_src_normals(lp) && !snrm) || (uses_tgt_normals(lp) && !tnrm)) ? -2 : 0;
  bfTrace("factor: coordinates (since plan)", t1);
  if (!rc) rc = upload((void **)&F->dSrcXY, sxy, 2 * F->numSrc * sizeof(double));
  if (!rc) rc = upload((void **)&F->dTgtXY, txy, 2 * F->numTgt * sizeof(double));
  if (!rc && snrm) rc = upload((void **)&F->dSrcN, snrm, 2 * F->numSrc * sizeof(double));
  if (!rc && tnrm) rc = upload((void **)&F->dTgtN, tnrm, 2 * F->numTgt * sizeof(double));
  free(sxy);
  free(txy);
  free(snrm);
  free(tnrm);
  double tApplyPlan = 0;
  if (!rc) rc = run_factor_kernels(F, &tApplyPlan);
  if (!rc) rc = bfk_sync();
  double t2 = bfWallTime();
  bfTrace("factor: kernels done (since plan)", t1);
  if (rc) {
    bfMatHipButterflyDelete((BfMat **)&F);
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  double t3 = bfWallTime();
  bfTrace("factor: total", t0);
  double ktimes[8];
  bfk_get_times(ktimes);
  /* the block and evaluation counts (stats 4-8, 10) are taken when the
   * statistics are first read: passes over tens of millions of entries
   * that the factorization itself does not need */
  F->statsCounted = 0;
  F->stats[0] = t3 - t0;
  F->stats[1] = t1 - t0;
  F->stats[2] = t2 - t1;
  F->stats[3] = (double)F->arenaEntries;
  F->stats[9] = (double)F->numStages;
  F->stats[13] = ktimes[0];
  F->stats[14] = ktimes[1];
  F->stats[15] = tApplyPlan + (t3 - t2);
  return &F->super;
}

static int problem_ok(BfHelm2 const *helm, BfQuadtree const *srcTree, BfQuadtree const *tgtTree,
                      BfFacHelm2Options const *opts) {
  if (!helm || !srcTree || !tgtTree || !opts || !(helm->k > 0) || !srcTree->nodes || !tgtTree->nodes) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return 0;
  }
  /* S, D, the combined field and S' (re-expanded with the single layer,
   * layer_pot.h:63-69); the normal derivative of the double layer is the
   * hypersingular operator, which the reference does not sample either
   * (helm2.c:282-319 has no case for it) */
  const BfLayerPotential lp = helm->layerPot;
  if (lp != BF_LAYER_POTENTIAL_SINGLE && !uses_src_normals(lp) && !uses_tgt_normals(lp)) {
    bfSetError(BF_ERROR_NOT_IMPLEMENTED);
    return 0;
  }
  if (uses_src_normals(lp) && (!srcTree->unitNormals || srcTree->unitNormals->size != srcTree->points->size)) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return 0;
  }
  if (uses_tgt_normals(lp) && (!tgtTree->unitNormals || tgtTree->unitNormals->size != tgtTree->points->size)) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return 0;
  }
  return 1;
}

BfMat *bfFacHelm2MakeMultilevelWithOptions(BfHelm2 const *helm, BfQuadtree const *srcTree, BfQuadtree const *tgtTree,
                                           BfFacHelm2Options const *opts) {
  if (!problem_ok(helm, srcTree, tgtTree, opts)) return NULL;
  if (bfk_init()) {
    bfSetError(BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  double t0 = bfWallTime();
  BfMatHipButterfly *F = calloc(1, sizeof(BfMatHipButterfly));
  if (!F) { bfSetError(BF_ERROR_MEMORY_ERROR); return NULL; }
  F->helm = *helm;
  F->opts = *opts;
  F->numSrc = srcTree->points->size;
  F->numTgt = tgtTree->points->size;
  bfMatInitBase(&F->super, &bfMatHipButterflyVtable, F->numTgt, F->numSrc);
  if (bfPlanBuild(F, srcTree, tgtTree)) {
    bfMatHipButterflyDelete((BfMat **)&F);
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  double t1 = bfWallTime();
  bfTrace("factor: plan", t0);
  return factor_from_plan(F, tree_coords(srcTree), tree_coords(tgtTree),
                          uses_src_normals(helm->layerPot) ? tree_normals(srcTree) : NULL,
                          uses_tgt_normals(helm->layerPot) ? tree_normals(tgtTree) : NULL, t0, t1);
}

BfMat *bfhipFacHelm2MakeBatched(BfSize count, BfHelm2 const *helms, BfQuadtree const *const *srcTrees,
                                BfQuadtree const *const *tgtTrees, BfFacHelm2Options const *opts) {
  if (count == 0 || !helms || !srcTrees || !tgtTrees || !opts) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  for (BfSize i = 0; i < count; ++i) {
    if (!problem_ok(&helms[i], srcTrees[i], tgtTrees[i], &opts[i])) return NULL;
    /* one layer potential per launch */
    if (helms[i].layerPot != helms[0].layerPot || helms[i].alpha != helms[0].alpha ||
        helms[i].beta != helms[0].beta) {
      bfSetError(BF_ERROR_NOT_IMPLEMENTED);
      return NULL;
    }
  }
  const int withNormals = uses_src_normals(helms[0].layerPot), withTgtNormals = uses_tgt_normals(helms[0].layerPot);
  if (bfk_init()) {
    bfSetError(BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  double t0 = bfWallTime();
  BfMatHipButterfly *F = calloc(1, sizeof(BfMatHipButterfly));
  if (!F) { bfSetError(BF_ERROR_MEMORY_ERROR); return NULL; }
  F->helm = helms[0];
  F->opts = opts[0];
  for (BfSize i = 0; i < count; ++i) {
    F->numSrc += srcTrees[i]->points->size;
    F->numTgt += tgtTrees[i]->points->size;
  }
  bfMatInitBase(&F->super, &bfMatHipButterflyVtable, F->numTgt, F->numSrc);
  double *sxy = malloc(2 * F->numSrc * sizeof(double)), *txy = malloc(2 * F->numTgt * sizeof(double));
  double *snrm = withNormals ? malloc(2 * F->numSrc * sizeof(double)) : NULL;
  double *tnrm = withTgtNormals ? malloc(2 * F->numTgt * sizeof(double)) : NULL;
  int64_t sBase = 0, tBase = 0;
  int bad = !sxy || !txy || (withNormals && !snrm) || (withTgtNormals && !tnrm);
  for (BfSize i = 0; i < count && !bad; ++i) {
    /* the planner reads the problem's wavenumber and options from F */
    F->helm = helms[i];
    F->opts = opts[i];
    if (bfPlanBuildAt(F, srcTrees[i], tgtTrees[i], sBase, tBase)) {
      bad = 1;
    } else {
      const BfSize ns = srcTrees[i]->points->size, nt = tgtTrees[i]->points->size;
      double *a = tree_coords(srcTrees[i]), *b = tree_coords(tgtTrees[i]);
      memcpy(sxy + 2 * sBase, a, 2 * ns * sizeof(double));
      memcpy(txy + 2 * tBase, b, 2 * nt * sizeof(double));
      free(a);
      free(b);
      if (snrm) {
        double *c = tree_normals(srcTrees[i]);
        memcpy(snrm + 2 * sBase, c, 2 * ns * sizeof(double));
        free(c);
      }
      if (tnrm) {
        double *c = tree_normals(tgtTrees[i]);
        memcpy(tnrm + 2 * tBase, c, 2 * nt * sizeof(double));
        free(c);
      }
      sBase += (int64_t)ns;
      tBase += (int64_t)nt;
    }
  }
  F->helm = helms[0];
  F->opts = opts[0];
  if (bad) {
    free(sxy);
    free(txy);
    free(snrm);
    free(tnrm);
    bfMatHipButterflyDelete((BfMat **)&F);
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  double t1 = bfWallTime();
  return factor_from_plan(F, sxy, txy, snrm, tnrm, t0, t1);
}

BfMat *bfhipFacHelm2Plan(BfHelm2 const *helm, BfQuadtree const *srcTree, BfQuadtree const *tgtTree,
                         BfFacHelm2Options const *opts) {
  if (!helm || !srcTree || !tgtTree || !opts || !(helm->k > 0) || !srcTree->nodes || !tgtTree->nodes) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  BfMatHipButterfly *F = calloc(1, sizeof(BfMatHipButterfly));
  if (!F) { bfSetError(BF_ERROR_MEMORY_ERROR); return NULL; }
  F->helm = *helm;
  F->opts = *opts;
  F->numSrc = srcTree->points->size;
  F->numTgt = tgtTree->points->size;
  bfMatInitBase(&F->super, &bfMatHipButterflyVtable, F->numTgt, F->numSrc);
  double t0 = bfWallTime();
  if (bfPlanBuild(F, srcTree, tgtTree) || bfApplyPlanBuild(F)) {
    bfMatHipButterflyDelete((BfMat **)&F);
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  F->stats[1] = bfWallTime() - t0;
  F->stats[3] = (double)F->arenaEntries;
  F->stats[9] = (double)F->numStages;
  return &F->super;
}

BfMat *bfFacHelm2MakeMultilevel(BfHelm2 const *helm, BfQuadtree const *srcTree, BfQuadtree const *tgtTree) {
  BfFacHelm2Options opts;
  bfFacHelm2OptionsInitDefault(&opts);
  return bfFacHelm2MakeMultilevelWithOptions(helm, srcTree, tgtTree, &opts);
}

/* ---------------------------------------- one butterfly, host containers */

/* allRankEstimatesAreOK (fac_helm2.c:511-530) */
static int ranks_ok(BfHelm2 const *helm, BfQuadtreeNode const *tgtNode, BfPtrArray const *lv) {
  const BfCircle tc = bfQuadtreeNodeGetBoundingCircle(tgtNode);
  for (BfSize i = 0; i < lv->num_elts; ++i) {
    BfQuadtreeNode const *sn = lv->data[i];
    const BfCircle sc = bfQuadtreeNodeGetBoundingCircle(sn);
    const double R = hypot(sc.center[0] - tc.center[0], sc.center[1] - tc.center[1]);
    if (!(R - tc.r - sc.r > 0)) return 0; /* not separated: no estimate */
    if (bfHelm2RankEstForTwoCircles(helm, &tc, &sc, 1, 1e-15) > bfQuadtreeNodeGetNumPoints(sn)) return 0;
  }
  return 1;
}

BfSize bfFacHelm2Prepare(BfHelm2 const *helm, BfQuadtreeNode const *srcNode, BfQuadtreeNode const *tgtNode,
                         BfTreeLevelIter *srcLevelIter, BfTreeLevelIter *tgtLevelIter) {
  if (!helm || !(helm->k > 0) || !srcNode || !tgtNode || !srcLevelIter || !tgtLevelIter) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return 0;
  }
  const BfSize numSrc = bfQuadtreeNodeGetNumPoints(srcNode);
  bfTreeLevelIterInit(srcLevelIter, BF_TREE_TRAVERSAL_LR_REVERSE_LEVEL_ORDER, (BfTreeNode *)srcNode);
  bfTreeLevelIterInit(tgtLevelIter, BF_TREE_TRAVERSAL_LR_LEVEL_ORDER, (BfTreeNode *)tgtNode);
  if (!srcLevelIter->aux || !tgtLevelIter->aux) {
    bfTreeLevelIterDeinit(srcLevelIter);
    bfTreeLevelIterDeinit(tgtLevelIter);
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return 0;
  }
  /* the deepest fully internal level below the target node */
  BfSize maxAllow = bfTreeLevelIterCurrentDepth(tgtLevelIter);
  bfTreeLevelIterNext(tgtLevelIter);
  while (!bfTreeLevelIterIsDone(tgtLevelIter) && bfTreeLevelIterCurrentLevelIsInternal(tgtLevelIter)) {
    ++maxAllow;
    bfTreeLevelIterNext(tgtLevelIter);
  }
  bfTreeLevelIterDeinit(tgtLevelIter);
  bfTreeLevelIterInit(tgtLevelIter, BF_TREE_TRAVERSAL_LR_LEVEL_ORDER, (BfTreeNode *)tgtNode);
  const BfSize top = (BfSize)srcNode->depth, tgtDepth = (BfSize)tgtNode->depth;
  BfSize depth = bfTreeLevelIterCurrentDepth(srcLevelIter);
  /* up to that depth, then to a complete, then to an internal level, then
   * until every rank estimate fits its node's points (never above the
   * source node itself) */
  while (depth > maxAllow && depth > top) { bfTreeLevelIterNext(srcLevelIter); --depth; }
  while (bfTreeLevelIterGetNumPoints(srcLevelIter) != numSrc && depth > top) {
    bfTreeLevelIterNext(srcLevelIter);
    --depth;
  }
  while (!bfTreeLevelIterCurrentLevelIsInternal(srcLevelIter) && depth > top) {
    bfTreeLevelIterNext(srcLevelIter);
    --depth;
  }
  while (depth > tgtDepth && depth > top && !ranks_ok(helm, tgtNode, srcLevelIter->levelNodes)) {
    bfTreeLevelIterNext(srcLevelIter);
    --depth;
  }
  if (depth < tgtDepth || !ranks_ok(helm, tgtNode, srcLevelIter->levelNodes)) return 0;
  return depth - tgtDepth + 2;
}

/* the nodes of an iterator's current level as indices into its tree */
static int32_t *level_indices(BfTreeLevelIter const *it, BfQuadtree const *t, size_t *n) {
  *n = it->levelNodes->num_elts;
  int32_t *out = malloc((*n ? *n : 1) * sizeof(int32_t));
  for (size_t i = 0; i < *n; ++i) {
    BfQuadtreeNode const *v = it->levelNodes->data[i];
    if (v < t->nodes || v >= t->nodes + t->numNodes) {
      free(out);
      return NULL;
    }
    out[i] = (int32_t)(v - t->nodes);
  }
  return out;
}

/* m x n dense complex (row-major) from a column-major block; NULL data is
 * the identity */
static BfMat *host_block(const BfComplex *a, int64_t m, int64_t n) {
  BfMatDenseComplex *D = bfMatDenseComplexNew();
  bfMatDenseComplexInit(D, (BfSize)m, (BfSize)n);
  for (int64_t i = 0; i < m; ++i)
    for (int64_t j = 0; j < n; ++j) D->data[i * n + j] = a ? a[j * m + i] : (i == j ? 1.0 : 0.0);
  return bfMatDenseComplexToMat(D);
}

/* leaf `leaf` of F as the reference's containers: a BfMatProduct of its
 * factors, last factor first (fac_helm2.c:693-696), each a BfMatBlockDiag
 * or BfMatBlockCoo of dense blocks copied from the host copy of the arena */
static BfMatProduct *host_product(BfMatHipButterfly const *F, const BfComplex *arena, size_t leaf) {
  const BfHipLeaf *L = &F->leaves.a[leaf];
  BfMatProduct *P = bfMatProductNew();
  bfMatProductInit(P);
  for (int f = L->numFactors - 1; f >= 0; --f) {
    const BfHipFactor *fa = &F->factors.a[L->firstFactor + f];
    BfMatBlock *B;
    bool *owns;
    BfMat *m;
    if (fa->isDiag) {
      BfMatBlockDiag *d = bfMatBlockDiagNew();
      bfMatBlockDiagInit(d, (BfSize)fa->numBlockRows, (BfSize)fa->numBlockCols);
      B = &d->super;
      owns = d->ownsBlock;
      m = bfMatBlockDiagToMat(d);
    } else {
      BfMatBlockCoo *c = bfMatBlockCooNew();
      bfMatBlockCooInit(c, (BfSize)fa->numBlockRows, (BfSize)fa->numBlockCols, (BfSize)fa->numBlocks);
      for (int64_t i = 0; i < fa->numBlocks; ++i) {
        c->rowInd[i] = (BfSize)F->blocks.a[fa->firstBlock + i].rowInd;
        c->colInd[i] = (BfSize)F->blocks.a[fa->firstBlock + i].colInd;
      }
      B = &c->super;
      owns = c->ownsBlock;
      m = &c->super.super;
    }
    for (int64_t i = 0; i <= fa->numBlockRows; ++i) B->rowOffset[i] = (BfSize)F->offs.a[fa->rowOffStart + i];
    for (int64_t j = 0; j <= fa->numBlockCols; ++j) B->colOffset[j] = (BfSize)F->offs.a[fa->colOffStart + j];
    for (int64_t i = 0; i < fa->numBlocks; ++i) {
      const BfHipBlock *b = &F->blocks.a[fa->firstBlock + i];
      B->block[i] = host_block(b->dataOff >= 0 ? arena + b->dataOff : NULL, b->m, b->n);
      owns[i] = true;
    }
    m->numRows = B->rowOffset[fa->numBlockRows];
    m->numCols = B->colOffset[fa->numBlockCols];
    bfMatProductPostMultiply(P, m);
  }
  return P;
}

BfMatProduct *bfFacHelm2Make(BfHelm2 const *helm, BfQuadtree const *srcTree, BfQuadtree const *tgtTree,
                             BfTreeLevelIter *srcLevelIter, BfTreeLevelIter *tgtLevelIter, BfSize numFactors) {
  BfFacHelm2Options opts;
  bfFacHelm2OptionsInitDefault(&opts);
  if (numFactors < 2 || !srcLevelIter || !tgtLevelIter || !srcLevelIter->aux || !tgtLevelIter->aux ||
      !problem_ok(helm, srcTree, tgtTree, &opts)) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  if (bfk_init()) {
    bfSetError(BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  /* the levels, walked as the reference walks them: the source iterator
   * up one level before each middle factor, the target iterator down one
   * after it (fac_helm2.c:674-686) */
  const int D = (int)numFactors - 2;
  int32_t **ls = calloc((size_t)D + 1, sizeof(int32_t *)), **lt = calloc((size_t)D + 1, sizeof(int32_t *));
  size_t *nls = calloc((size_t)D + 1, sizeof(size_t)), *nlt = calloc((size_t)D + 1, sizeof(size_t));
  int bad = 0;
  ls[D] = level_indices(srcLevelIter, srcTree, &nls[D]);
  lt[0] = level_indices(tgtLevelIter, tgtTree, &nlt[0]);
  bad = !ls[D] || !lt[0];
  for (int i = 1; i <= D && !bad; ++i) {
    bfTreeLevelIterNext(srcLevelIter);
    ls[D - i] = level_indices(srcLevelIter, srcTree, &nls[D - i]);
    bfTreeLevelIterNext(tgtLevelIter);
    lt[i] = level_indices(tgtLevelIter, tgtTree, &nlt[i]);
    bad = !ls[D - i] || !lt[i];
  }
  BfMatHipButterfly *F = NULL;
  if (!bad && (nls[0] != 1 || nlt[0] != 1)) bad = 1; /* one source node on top, as makeLastFactor asserts */
  if (!bad) {
    F = calloc(1, sizeof(BfMatHipButterfly));
    F->helm = *helm;
    F->opts = opts;
    F->numSrc = srcTree->points->size;
    F->numTgt = tgtTree->points->size;
    bfMatInitBase(&F->super, &bfMatHipButterflyVtable, F->numTgt, F->numSrc);
    bad = bfPlanButterfly(F, srcTree, tgtTree, (const int32_t *const *)ls, nls, (const int32_t *const *)lt, nlt, D);
  }
  for (int i = 0; i <= D; ++i) {
    free(ls[i]);
    free(lt[i]);
  }
  free(ls); free(lt); free(nls); free(nlt);
  if (bad) {
    if (F) bfMatHipButterflyDelete((BfMat **)&F);
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  const double t0 = bfWallTime();
  BfMat *m = factor_from_plan(F, tree_coords(srcTree), tree_coords(tgtTree),
                              uses_src_normals(helm->layerPot) ? tree_normals(srcTree) : NULL,
                              uses_tgt_normals(helm->layerPot) ? tree_normals(tgtTree) : NULL, t0, t0);
  if (!m) return NULL;
  BfComplex *arena = malloc((size_t)(F->arenaEntries ? F->arenaEntries : 1) * sizeof(BfComplex));
  BfMatProduct *P = NULL;
  if (arena && !bfk_d2h(arena, F->dArena, (size_t)F->arenaEntries * sizeof(BfComplex)))
    P = host_product(F, arena, 0);
  else
    bfSetError(arena ? BF_ERROR_RUNTIME_ERROR : BF_ERROR_MEMORY_ERROR);
  free(arena);
  bfMatDelete(&m);
  return P;
}

BfMat *bfFacHelm2MakeSingleLevel(BfHelm2 const *helm, BfQuadtreeNode const *srcNode, BfQuadtreeNode const *tgtNode) {
  BfQuadtree const *srcTree = bfQuadtreeNodeGetQuadtreeConst(srcNode);
  BfQuadtree const *tgtTree = bfQuadtreeNodeGetQuadtreeConst(tgtNode);
  if (!srcTree || !tgtTree || srcNode->depth != tgtNode->depth) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  BfTreeLevelIter si, ti;
  const BfSize nf = bfFacHelm2Prepare(helm, srcNode, tgtNode, &si, &ti);
  BfMatProduct *P = NULL;
  if (nf == 0)
    bfSetError(BF_ERROR_RUNTIME_ERROR);
  else
    P = bfFacHelm2Make(helm, srcTree, tgtTree, &si, &ti, nf);
  bfTreeLevelIterDeinit(&si);
  bfTreeLevelIterDeinit(&ti);
  return bfMatProductToMat(P);
}

static int ensure_buffers(BfMatHipButterfly *F, size_t nrhs, size_t elemBytes) {
  size_t need = (size_t)(F->bufRows[0] > F->bufRows[1] ? F->bufRows[0] : F->bufRows[1]) * nrhs * elemBytes;
  if (need == 0) need = 16;
  if (need <= F->bufBytes) return 0;
  bfk_free(F->dPing);
  bfk_free(F->dPong);
  F->dPing = bfk_malloc(need);
  F->dPong = bfk_malloc(need);
  if (!F->dPing || !F->dPong) {
    F->bufBytes = 0;
    return -2;
  }
  F->bufBytes = need;
  return 0;
}

/* compression: the LDS classes (bytes per block) above which a
 * re-expansion block gets 2 and 4 waves */
void bfhipSetLstsqWaves(int64_t twoAbove, int64_t fourAbove, int64_t *oldTwo, int64_t *oldFour) {
  bfk_set_lstsq_waves(twoAbove, fourAbove, oldTwo, oldFour);
}

/* apply graphs: off (0, the default) or on */
static int g_applyGraphs = 0;

int bfhipSetApplyGraphs(int on) {
  g_applyGraphs = on ? 1 : 0;
  return 0;
}

/* the stage launches of one apply (into a graph while one is captured) */
static int enqueue_apply(BfMatHipButterfly *F, void const *dX, void *dY, BfSize nrhs, BfhipPrecision prec,
                         const int32_t *cfg) {
  if (prec == BFHIP_PRECISION_FP64)
    return bfk_apply(F->dSegs, F->stageSeg.a, F->numStages, F->dTerms, F->dArena, dX, F->dPing, F->dPong, dY,
                     (int)nrhs, cfg);
  return bfk_apply_f32(F->dSegs, F->stageSeg.a, F->numStages, F->dTerms, F->dArena32,
                       prec == BFHIP_PRECISION_FP32X2 ? F->dArena32lo : NULL, dX, F->dPing, F->dPong, dY, (int)nrhs,
                       cfg);
}

int bfhipMatMulDevice(BfMat const *fac, void const *dX, void *dY, BfSize nrhs, BfhipPrecision prec) {
  if (!fac || fac->vtbl != &bfMatHipButterflyVtable || nrhs == 0 || (unsigned)prec > 2u) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac;
  int rc;
  if (!F->dSegs || (prec == BFHIP_PRECISION_FP64 && !F->dArena) ||
      (prec != BFHIP_PRECISION_FP64 && (!F->dArena32 || (prec == BFHIP_PRECISION_FP32X2 && !F->dArena32lo)))) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  if ((rc = ensure_buffers(F, nrhs, prec == BFHIP_PRECISION_FP64 ? 16 : 8))) {
    bfSetError(BF_ERROR_MEMORY_ERROR);
    return rc;
  }
  const int32_t *cfg = F->tuneCfg[prec][bf_nrhs_class(nrhs)];
  if (!g_applyGraphs) {
    rc = enqueue_apply(F, dX, dY, nrhs, prec, cfg);
  } else {
    /* one captured chain per factorization, replayed while the vectors,
     * buffers, precision and launch shapes stay the same */
    const int stamp = bfk_apply_config_stamp();
    if (!F->graphExec || F->graphX != dX || F->graphY != dY || F->graphPing != F->dPing ||
        F->graphPong != F->dPong || F->graphNrhs != nrhs || F->graphPrec != (int)prec || F->graphCfg != cfg ||
        F->graphStamp != stamp) {
      bfk_graph_destroy(F->graphExec);
      F->graphExec = NULL;
      rc = bfk_graph_begin();
      if (!rc) {
        const int rcApply = enqueue_apply(F, dX, dY, nrhs, prec, cfg);
        rc = bfk_graph_end(&F->graphExec);
        if (!rc) rc = rcApply;
      }
      if (!rc) {
        F->graphX = dX;
        F->graphY = dY;
        F->graphPing = F->dPing;
        F->graphPong = F->dPong;
        F->graphNrhs = nrhs;
        F->graphPrec = (int)prec;
        F->graphCfg = cfg;
        F->graphStamp = stamp;
      }
    }
    if (!rc) rc = bfk_graph_launch(F->graphExec, prec == BFHIP_PRECISION_FP64 ? 2 : 3);
  }
  if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}

/* ------------------------------------------------------- the autotuner */

/* A stage's bucket: rounded log2 of its mean segment height and of its
 * mean columns per segment (the sum over the segment's terms). */
static void stage_bucket(BfMatHipButterfly const *F, int s, int *hb, int *cb) {
  const int64_t a = F->stageSeg.a[s], b = F->stageSeg.a[s + 1];
  double rows = 0, cols = 0;
  for (int64_t i = a; i < b; ++i) {
    rows += F->segs.a[i].m;
    for (int32_t t = F->segs.a[i].t0; t < F->segs.a[i].t1; ++t) cols += F->terms.a[t].aOff < 0 ? 1 : F->terms.a[t].n;
  }
  const double n = (double)(b > a ? b - a : 1);
  *hb = (int)lround(log2(rows / n > 1 ? rows / n : 1));
  *cb = (int)lround(log2(cols / n > 1 ? cols / n : 1));
}

typedef struct {
  int prec, cls, hb, cb, cfg;
} TuneEntry;
typedef BF_VEC(TuneEntry) TuneVec;

static void tune_load(TuneVec *v, char const *path) {
  FILE *fp = path ? fopen(path, "r") : NULL;
  if (!fp) return;
  TuneEntry e;
  while (fscanf(fp, "%d %d %d %d %d", &e.prec, &e.cls, &e.hb, &e.cb, &e.cfg) == 5) BF_VEC_PUSH(*v, e);
  fclose(fp);
}

static int tune_save(const TuneVec *v, char const *path) {
  FILE *fp = path ? fopen(path, "w") : NULL;
  if (!fp) return path ? -1 : 0;
  fprintf(fp, "%s", "");
  for (size_t i = 0; i < v->n; ++i)
    fprintf(fp, "%d %d %d %d %d\n", v->a[i].prec, v->a[i].cls, v->a[i].hb, v->a[i].cb, v->a[i].cfg);
  return fclose(fp) ? -1 : 0;
}

int bfhipFacTuneApply(BfMat *fac, BfSize nrhs, BfhipPrecision prec, char const *cachePath) {
  if (!fac || fac->vtbl != &bfMatHipButterflyVtable || nrhs == 0 || (unsigned)prec > 2u) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac;
  if (!F->dSegs || (prec == BFHIP_PRECISION_FP64 && !F->dArena) ||
      (prec != BFHIP_PRECISION_FP64 && (!F->dArena32 || (prec == BFHIP_PRECISION_FP32X2 && !F->dArena32lo)))) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  const int cls = bf_nrhs_class(nrhs);
  if (cls == 1) return 0; /* 2-15 right-hand sides: the vector-unit kernel only */
  if (!cachePath) cachePath = getenv("BF_AMD_TUNE_CACHE");
  /* candidates: waves per block x lane map (1 RHS); tiles per MFMA wave
   * x one or two row tiles, 0 = the vector-unit kernel (16 and more) */
  static const int one[] = {1, 2, 4, 1 | 16, 2 | 16, 4 | 16}, many[] = {1, 2, 4, 8, 1 | 16, 2 | 16, 4 | 16, 0};
  const int *cand = cls == 0 ? one : many;
  const int nc = cls == 0 ? 6 : 8;
  const size_t elem = prec == BFHIP_PRECISION_FP64 ? 16 : 8;
  int rc = ensure_buffers(F, nrhs, elem);
  void *dX = rc ? NULL : bfk_malloc(F->numSrc * nrhs * elem + 16);
  void *dY = rc ? NULL : bfk_malloc(F->numTgt * nrhs * elem + 16);
  if (!dX || !dY) rc = -2;
  if (!rc) rc = bfk_memset(dX, F->numSrc * nrhs * elem + 16);
  if (!rc) rc = enqueue_apply(F, dX, dY, nrhs, prec, NULL); /* finite values in every buffer */
  TuneVec cache;
  memset(&cache, 0, sizeof(cache));
  tune_load(&cache, cachePath);
  const size_t cached = cache.n;
  int32_t *cfg = calloc((size_t)(F->numStages ? F->numStages : 1), sizeof(int32_t));
  for (int s = 0; s < F->numStages && !rc; ++s) {
    int hb, cb, found = 0;
    stage_bucket(F, s, &hb, &cb);
    for (size_t i = 0; i < cache.n && !found; ++i)
      if (cache.a[i].prec == (int)prec && cache.a[i].cls == cls && cache.a[i].hb == hb && cache.a[i].cb == cb) {
        cfg[s] = cache.a[i].cfg;
        found = 1;
      }
    if (found || F->stageSeg.a[s + 1] == F->stageSeg.a[s]) continue;
    double best = 1e300;
    for (int c = 0; c < nc && !rc; ++c) {
      double ms = 0;
      rc = prec == BFHIP_PRECISION_FP64
               ? bfk_apply_stage_ms(F->dSegs, F->stageSeg.a, s, F->dTerms, F->dArena, dX, F->dPing, F->dPong, dY,
                                    (int)nrhs, cand[c], 3, &ms)
               : bfk_apply_f32_stage_ms(F->dSegs, F->stageSeg.a, s, F->dTerms, F->dArena32,
                                        prec == BFHIP_PRECISION_FP32X2 ? F->dArena32lo : NULL, dX, F->dPing,
                                        F->dPong, dY, (int)nrhs, cand[c], 3, &ms);
      if (!rc && ms < best) {
        best = ms;
        cfg[s] = cand[c];
      }
    }
    TuneEntry e = {(int)prec, cls, hb, cb, cfg[s]};
    if (!rc) BF_VEC_PUSH(cache, e);
  }
  if (!rc && cache.n > cached && tune_save(&cache, cachePath)) rc = -3;
  bfk_free(dX);
  bfk_free(dY);
  BF_VEC_FREE(cache);
  if (rc) {
    free(cfg);
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : rc == -3 ? BF_ERROR_FILE_ERROR : BF_ERROR_RUNTIME_ERROR);
    return rc;
  }
  free(F->tuneCfg[prec][cls]);
  F->tuneCfg[prec][cls] = cfg;
  return 0;
}

int bfhipFacGetApplyConfig(BfMat const *fac, BfSize nrhs, BfhipPrecision prec, int32_t *stageCfg) {
  if (!fac || fac->vtbl != &bfMatHipButterflyVtable || (unsigned)prec > 2u) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  const BfMatHipButterfly *F = (const BfMatHipButterfly *)fac;
  const int32_t *cfg = F->tuneCfg[prec][bf_nrhs_class(nrhs)];
  if (stageCfg)
    for (int s = 0; s < F->numStages; ++s) stageCfg[s] = cfg ? cfg[s] : -1;
  return F->numStages;
}

/* bfhipMatMulDevice with the intermediate vectors in the library's fixed
 * device scratch array (two halves, ping and pong) instead of buffers it
 * allocates: no device memory is allocated or freed.  Every stage writes
 * all the rows the next stage reads, so nothing needs clearing. */
int bfhipMatMulDeviceNoAlloc(BfMat const *fac, void const *dX, void *dY, BfSize nrhs, BfhipPrecision prec) {
  if (!fac || fac->vtbl != &bfMatHipButterflyVtable || nrhs == 0 || !dX || !dY) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac;
  if (!F->dSegs || (prec == BFHIP_PRECISION_FP64 && !F->dArena) ||
      (prec != BFHIP_PRECISION_FP64 && (!F->dArena32 || (prec == BFHIP_PRECISION_FP32X2 && !F->dArena32lo)))) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  void *ws = NULL;
  size_t wsBytes = 0;
  if (bfk_scratch(&ws, &wsBytes)) {
    bfSetError(BF_ERROR_RUNTIME_ERROR);
    return -1;
  }
  const size_t elem = prec == BFHIP_PRECISION_FP64 ? 16 : 8;
  const int64_t rows = F->bufRows[0] > F->bufRows[1] ? F->bufRows[0] : F->bufRows[1];
  const size_t half = (wsBytes / 2) & ~(size_t)255;
  if ((size_t)rows > half / elem / nrhs) { /* more than a half of the scratch per buffer */
    bfSetError(BF_ERROR_OUT_OF_RANGE);
    return -1;
  }
  void *ping = ws, *pong = (char *)ws + half;
  int rc;
  const int32_t *cfg = F->tuneCfg[prec][bf_nrhs_class(nrhs)];
  if (prec == BFHIP_PRECISION_FP64)
    rc = bfk_apply(F->dSegs, F->stageSeg.a, F->numStages, F->dTerms, F->dArena, dX, ping, pong, dY, (int)nrhs, cfg);
  else
    rc = bfk_apply_f32(F->dSegs, F->stageSeg.a, F->numStages, F->dTerms, F->dArena32,
                       prec == BFHIP_PRECISION_FP32X2 ? F->dArena32lo : NULL, dX, ping, pong, dY, (int)nrhs, cfg);
  if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}

int bfhipFacPrepareMixed(BfMat *fac, BfhipPrecision prec) {
  if (!fac || fac->vtbl != &bfMatHipButterflyVtable) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac;
  if (prec == BFHIP_PRECISION_FP64) return 0;
  size_t bytes = (size_t)F->arenaEntries * 8;
  if (!F->dArena32) {
    F->dArena32 = bfk_malloc(bytes);
    if (!F->dArena32) { bfSetError(BF_ERROR_MEMORY_ERROR); return -2; }
  }
  if (prec == BFHIP_PRECISION_FP32X2 && !F->dArena32lo) {
    F->dArena32lo = bfk_malloc(bytes);
    if (!F->dArena32lo) { bfSetError(BF_ERROR_MEMORY_ERROR); return -2; }
  }
  int rc = bfk_to_f32(F->dArena, F->arenaEntries, F->dArena32, F->dArena32lo);
  if (!rc) rc = bfk_sync();
  if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}

int bfhipFacReleaseFp64(BfMat *fac) {
  if (!fac || fac->vtbl != &bfMatHipButterflyVtable) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac;
  if (!F->dArena32) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  bfk_free(F->dArena);
  F->dArena = NULL;
  bfk_free(F->dPing);
  bfk_free(F->dPong);
  F->dPing = F->dPong = NULL;
  F->bufBytes = 0;
  return 0;
}

/* ---- orthogonalization
 *
 * For every butterfly leaf, from its last factor back to its first: the
 * blocks of one column of factor s (s >= 1) are stacked, multiplied on the
 * left by the triangular factors their rows received from factor s + 1,
 * and replaced by the Q of a thin QR; the R goes to the row of factor
 * s - 1 that feeds this column.  Then the map from every intermediate
 * vector to the output has orthonormal columns, so rounding an
 * intermediate (or a block) by eps changes the output by about eps
 * relative, instead of eps times the cancellation of equivalent charges.
 * Only columns fed by proxy rows (whose blocks are all re-expansions) are
 * rotated; identity and point-value passes stay as they are. */

typedef BF_VEC(BfkOrthoJob) OrthoJobVec;
typedef BF_VEC(BfkOrthoItem) OrthoItemVec;

static void ortho_push(OrthoJobVec *jobs, OrthoItemVec *items, const BfkOrthoItem *it, int nit, int n, int64_t rOut) {
  BfkOrthoJob j;
  j.rOut = rOut;
  j.n = n;
  j.rows = 0;
  j.i0 = (int32_t)items->n;
  for (int q = 0; q < nit; ++q) {
    BF_VEC_PUSH(*items, it[q]);
    j.rows += it[q].m;
  }
  j.i1 = (int32_t)items->n;
  BF_VEC_PUSH(*jobs, j);
}

/* The orthogonalization's jobs for leaves [l0, l1): per stage g the jobs
 * (item indices and R offsets local to this range), the items, and the R
 * entries each stage writes.  Ranges of leaves are independent: a leaf's
 * triangular factors pass only between its own stages. */
typedef struct {
  const BfMatHipButterfly *F;
  size_t l0, l1;
  int Fmax;
  size_t cap;
  OrthoJobVec *jobs; /* [Fmax] */
  OrthoItemVec items;
  int64_t *rFill;    /* [Fmax] */
  /* the merge: this range's first item, first job and first R entry per
   * stage in the concatenated tables, and those tables */
  int32_t itemBase;
  int64_t *jobBase, *rBase; /* [Fmax + 1] */
  OrthoJobVec *outJobs;     /* [Fmax] */
  BfkOrthoItem *outItems;
} OrthoPlanTask;

/* copy one range's jobs and items into the concatenated tables, rebasing
 * item indices and R offsets (the items of a stage-g job read the R that
 * stage g + 1 writes) */
static void *ortho_merge_range(void *arg) {
  OrthoPlanTask *t = arg;
  for (int g = 0; g < t->Fmax; ++g) {
    const OrthoJobVec *src = &t->jobs[g];
    BfkOrthoJob *dst = t->outJobs[g].a + t->jobBase[g];
    for (size_t i = 0; i < src->n; ++i) {
      BfkOrthoJob j = src->a[i];
      for (int32_t it = j.i0; it < j.i1; ++it) {
        BfkOrthoItem itm = t->items.a[it];
        if (itm.rIn >= 0) itm.rIn += t->rBase[g + 1];
        t->outItems[t->itemBase + it] = itm;
      }
      j.i0 += t->itemBase;
      j.i1 += t->itemBase;
      if (j.rOut >= 0) j.rOut += t->rBase[g];
      dst[i] = j;
    }
  }
  return NULL;
}

static void *ortho_plan_range(void *arg) {
  OrthoPlanTask *t = arg;
  const BfMatHipButterfly *F = t->F;
  const int Fmax = t->Fmax;
  const size_t cap = t->cap;
  BF_VEC(BfkOrthoItem) col;
  memset(&col, 0, sizeof(col));
  BF_VEC(int64_t) pend, npend, cstart, corder;
  memset(&pend, 0, sizeof(pend));
  memset(&npend, 0, sizeof(npend));
  memset(&cstart, 0, sizeof(cstart));
  memset(&corder, 0, sizeof(corder));
  BF_VEC(char) rowOK;
  memset(&rowOK, 0, sizeof(rowOK));
  for (size_t l = t->l0; l < t->l1; ++l) {
    BfHipLeaf *lf = &F->leaves.a[l];
    if (!lf->isButterfly) continue;
    const int nf = lf->numFactors;
    pend.n = 0;
    for (int s = nf - 1; s >= 0; --s) {
      BfHipFactor *fa = &F->factors.a[lf->firstFactor + s];
      const int g = Fmax - nf + s;
      const int64_t nbc = fa->numBlockCols, b0 = fa->firstBlock, b1 = fa->firstBlock + fa->numBlocks;
      /* rows of factor s - 1 that can take a triangular factor: all
       * re-expansion blocks, each small enough to be rotated alone */
      rowOK.n = 0;
      if (s >= 1) {
        BfHipFactor *fp = &F->factors.a[lf->firstFactor + s - 1];
        for (int64_t r = 0; r < fp->numBlockRows; ++r) BF_VEC_PUSH(rowOK, 1);
        for (int64_t b = fp->firstBlock; b < fp->firstBlock + fp->numBlocks; ++b) {
          BfHipBlock *bl = &F->blocks.a[b];
          if (bl->kind != BF_BLOCK_LSTSQ || bl->dataOff < 0 || bfk_ortho_units(bl->m, bl->n) > cap)
            rowOK.a[bl->rowInd] = 0;
        }
      }
      /* blocks of factor s grouped by column (counting sort) */
      cstart.n = 0;
      for (int64_t c = 0; c <= nbc; ++c) BF_VEC_PUSH(cstart, 0);
      for (int64_t b = b0; b < b1; ++b) cstart.a[F->blocks.a[b].colInd + 1]++;
      for (int64_t c = 0; c < nbc; ++c) cstart.a[c + 1] += cstart.a[c];
      corder.n = 0;
      for (int64_t b = b0; b < b1; ++b) BF_VEC_PUSH(corder, 0);
      for (int64_t b = b0; b < b1; ++b) corder.a[cstart.a[F->blocks.a[b].colInd]++] = b;
      for (int64_t c = nbc; c > 0; --c) cstart.a[c] = cstart.a[c - 1];
      cstart.a[0] = 0;
      npend.n = 0;
      for (int64_t c = 0; c < nbc; ++c) BF_VEC_PUSH(npend, -1);
      for (int64_t c = 0; c < nbc; ++c) {
        col.n = 0;
        int n = 0, rows = 0, allData = 1, anyPend = 0;
        for (int64_t q = cstart.a[c]; q < cstart.a[c + 1]; ++q) {
          BfHipBlock *bl = &F->blocks.a[corder.a[q]];
          BfkOrthoItem it;
          it.aOff = bl->dataOff;
          it.rIn = pend.n ? pend.a[bl->rowInd] : -1;
          it.m = bl->m;
          it.pad = 0;
          n = bl->n;
          rows += bl->m;
          if (bl->dataOff < 0) allData = 0;
          if (it.rIn >= 0) anyPend = 1;
          BF_VEC_PUSH(col, it);
        }
        if (col.n == 0) continue;
        const int fits = bfk_ortho_units(rows, n) <= cap;
        if (s >= 1 && rowOK.a[c] && allData && rows >= n && fits) {
          const int64_t rOut = t->rFill[g];
          t->rFill[g] += (int64_t)n * n;
          ortho_push(&t->jobs[g], &t->items, col.a, (int)col.n, n, rOut);
          npend.a[c] = rOut;
        } else if (anyPend) {
          /* only the blocks whose rows carry a triangular factor change */
          size_t k = 0;
          int prow = 0;
          for (size_t q = 0; q < col.n; ++q)
            if (col.a[q].rIn >= 0) { col.a[k++] = col.a[q]; prow += col.a[q].m; }
          if (bfk_ortho_units(prow, n) <= cap) ortho_push(&t->jobs[g], &t->items, col.a, (int)k, n, -1);
          else
            for (size_t q = 0; q < k; ++q) ortho_push(&t->jobs[g], &t->items, &col.a[q], 1, n, -1);
        }
      }
      /* the next (earlier) factor's rows are this factor's columns */
      pend.n = 0;
      if (s >= 1)
        for (int64_t c = 0; c < nbc; ++c) BF_VEC_PUSH(pend, npend.a[c]);
    }
  }
  BF_VEC_FREE(col);
  BF_VEC_FREE(pend);
  BF_VEC_FREE(npend);
  BF_VEC_FREE(cstart);
  BF_VEC_FREE(corder);
  BF_VEC_FREE(rowOK);
  return NULL;
}

int bfhipFacOrthogonalize(BfMat *fac) {
  if (!fac || fac->vtbl != &bfMatHipButterflyVtable) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac;
  if (!F->dArena) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  if (F->orthogonal) return 0;
  double t0 = bfWallTime();
  const int Fmax = F->numStages;
  const size_t cap = bfk_lds_cap() / 16;
  OrthoJobVec *jobs = calloc((size_t)Fmax, sizeof(OrthoJobVec));
  OrthoItemVec items;
  memset(&items, 0, sizeof(items));
  int64_t *rFill = calloc((size_t)Fmax, sizeof(int64_t));

  /* plan on threads over ranges of leaves of about equal block counts,
   * then concatenate the ranges in order, rebasing item indices and R
   * offsets: the same jobs, items and offsets as one pass over all leaves */
  enum { ORTHO_NT = 16 };
  int64_t totalBlocks = 0;
  for (size_t l = 0; l < F->leaves.n; ++l) {
    const BfHipLeaf *lf = &F->leaves.a[l];
    if (!lf->isButterfly) continue;
    for (int s = 0; s < lf->numFactors; ++s) totalBlocks += F->factors.a[lf->firstFactor + s].numBlocks;
  }
  const int nt = totalBlocks < (1 << 16) ? 1 : ORTHO_NT;
  OrthoPlanTask tasks[ORTHO_NT];
  memset(tasks, 0, sizeof(tasks));
  {
    size_t l = 0;
    int64_t acc = 0;
    for (int q = 0; q < nt; ++q) {
      tasks[q].F = F;
      tasks[q].Fmax = Fmax;
      tasks[q].cap = cap;
      tasks[q].l0 = l;
      const int64_t target = totalBlocks * (q + 1) / nt;
      while (l < F->leaves.n && (q == nt - 1 || acc < target)) {
        const BfHipLeaf *lf = &F->leaves.a[l];
        if (lf->isButterfly)
          for (int s = 0; s < lf->numFactors; ++s) acc += F->factors.a[lf->firstFactor + s].numBlocks;
        ++l;
      }
      tasks[q].l1 = l;
      tasks[q].jobs = calloc((size_t)Fmax, sizeof(OrthoJobVec));
      tasks[q].rFill = calloc((size_t)Fmax, sizeof(int64_t));
    }
  }
  {
    pthread_t th[ORTHO_NT];
    int ok[ORTHO_NT] = {0};
    for (int q = 1; q < nt; ++q) {
      ok[q] = pthread_create(&th[q], NULL, ortho_plan_range, &tasks[q]) == 0;
      if (!ok[q]) ortho_plan_range(&tasks[q]);
    }
    ortho_plan_range(&tasks[0]);
    for (int q = 1; q < nt; ++q)
      if (ok[q]) pthread_join(th[q], NULL);
  }
  {
    int64_t *bases = calloc((size_t)nt * 2 * (Fmax + 1), sizeof(int64_t));
    size_t nItems = 0;
    for (int q = 0; q < nt; ++q) {
      tasks[q].itemBase = (int32_t)nItems;
      nItems += tasks[q].items.n;
      tasks[q].jobBase = bases + (size_t)q * 2 * (Fmax + 1);
      tasks[q].rBase = tasks[q].jobBase + (Fmax + 1);
      tasks[q].outJobs = jobs;
    }
    items.a = malloc((nItems ? nItems : 1) * sizeof(BfkOrthoItem));
    bfHugePages(items.a, nItems * sizeof(BfkOrthoItem));
    items.cap = nItems ? nItems : 1;
    items.n = nItems;
    for (int g = 0; g < Fmax; ++g) {
      size_t nj = 0;
      for (int q = 0; q < nt; ++q) {
        tasks[q].jobBase[g] = (int64_t)nj;
        tasks[q].rBase[g] = rFill[g];
        rFill[g] += tasks[q].rFill[g];
        nj += tasks[q].jobs[g].n;
      }
      jobs[g].a = malloc((nj ? nj : 1) * sizeof(BfkOrthoJob));
      bfHugePages(jobs[g].a, nj * sizeof(BfkOrthoJob));
      jobs[g].cap = nj ? nj : 1;
      jobs[g].n = nj;
    }
    for (int q = 0; q < nt; ++q) tasks[q].outItems = items.a;
    pthread_t th[ORTHO_NT];
    int ok[ORTHO_NT] = {0};
    for (int q = 1; q < nt; ++q) {
      ok[q] = pthread_create(&th[q], NULL, ortho_merge_range, &tasks[q]) == 0;
      if (!ok[q]) ortho_merge_range(&tasks[q]);
    }
    ortho_merge_range(&tasks[0]);
    for (int q = 1; q < nt; ++q)
      if (ok[q]) pthread_join(th[q], NULL);
    for (int q = 0; q < nt; ++q) {
      for (int g = 0; g < Fmax; ++g) BF_VEC_FREE(tasks[q].jobs[g]);
      free(tasks[q].jobs);
      free(tasks[q].rFill);
      BF_VEC_FREE(tasks[q].items);
    }
    free(bases);
  }
  bfTrace("ortho: jobs planned", t0);

  /* two R buffers, alternating by stage parity: stage g writes buffer
   * g % 2, stage g - 1 reads it */
  int64_t rMax[2] = {0, 0};
  for (int g = 0; g < Fmax; ++g)
    if (rFill[g] > rMax[g % 2]) rMax[g % 2] = rFill[g];
  int rc = 0;
  double *dR[2] = {NULL, NULL};
  void *dItems = NULL, *dJobs = NULL, *dOrder = NULL;
  for (int q = 0; q < 2 && !rc; ++q) {
    dR[q] = bfk_malloc((size_t)(rMax[q] ? rMax[q] : 1) * 16);
    if (!dR[q]) rc = -2;
  }
  if (!rc) rc = upload(&dItems, items.a, items.n * sizeof(BfkOrthoItem));
  size_t maxJobs = 0;
  for (int g = 0; g < Fmax; ++g)
    if (jobs[g].n > maxJobs) maxJobs = jobs[g].n;
  if (!rc) {
    dJobs = bfk_malloc((maxJobs ? maxJobs : 1) * sizeof(BfkOrthoJob));
    dOrder = bfk_malloc((maxJobs ? maxJobs : 1) * sizeof(int32_t));
    if (!dJobs || !dOrder) rc = -2;
  }
  uint32_t *cls = malloc((maxJobs ? maxJobs : 1) * sizeof(uint32_t));
  int32_t *order = malloc((maxJobs ? maxJobs : 1) * sizeof(int32_t));
  BfI64Vec gs;
  BfSizeVec gb;
  memset(&gs, 0, sizeof(gs));
  memset(&gb, 0, sizeof(gb));
  const uint32_t capCls = (uint32_t)((cap * 16) >> 12);
  for (int g = Fmax - 1; g >= 0 && !rc; --g) {
    size_t nj = jobs[g].n;
    if (nj == 0) continue;
    for (size_t i = 0; i < nj; ++i) {
      uint32_t c = (uint32_t)((bfk_ortho_units(jobs[g].a[i].rows, jobs[g].a[i].n) * 16 + 4095) >> 12);
      cls[i] = c > capCls ? capCls : c;
    }
    group_by_class(cls, nj, order, &gs, &gb, 0);
    rc = bfk_h2d(dJobs, jobs[g].a, nj * sizeof(BfkOrthoJob));
    if (!rc) rc = bfk_h2d(dOrder, order, nj * sizeof(int32_t));
    if (!rc) rc = bfk_ortho(dJobs, dOrder, (int)gb.n, gs.a, gb.a, dItems, F->dArena, dR[(g + 1) % 2], dR[g % 2]);
    if (!rc) rc = bfk_sync();
  }
  bfTrace("ortho: stages done", t0);
  BF_VEC_FREE(gs);
  BF_VEC_FREE(gb);
  free(cls);
  free(order);
  bfk_free(dJobs);
  bfk_free(dOrder);
  bfk_free(dItems);
  bfk_free(dR[0]);
  bfk_free(dR[1]);
  for (int g = 0; g < Fmax; ++g) BF_VEC_FREE(jobs[g]);
  free(jobs);
  free(rFill);
  BF_VEC_FREE(items);
  if (rc) {
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return rc;
  }
  F->orthogonal = 1;
  F->orthoSeconds = bfWallTime() - t0;
  return 0;
}

double bfhipFacOrthogonalizeSeconds(BfMat const *fac) {
  if (!fac || fac->vtbl != &bfMatHipButterflyVtable) return -1;
  return ((BfMatHipButterfly const *)fac)->orthoSeconds;
}

BfMat *bfMatHipButterflyMul(BfMat const *A, BfMat const *Xm) {
  BfMatHipButterfly *F = (BfMatHipButterfly *)A;
  if (!Xm || bfMatGetType(Xm) != BF_TYPE_MAT_DENSE_COMPLEX || Xm->numRows != F->numSrc) {
    bfSetError(BF_ERROR_INCOMPATIBLE_SHAPES);
    return NULL;
  }
  BfMatDenseComplex const *X = (BfMatDenseComplex const *)Xm;
  BfSize nrhs = Xm->numCols;
  BfComplex *packed = malloc(F->numSrc * nrhs * sizeof(BfComplex));
  if (!packed) { bfSetError(BF_ERROR_MEMORY_ERROR); return NULL; }
  for (BfSize i = 0; i < F->numSrc; ++i)
    for (BfSize j = 0; j < nrhs; ++j) packed[i * nrhs + j] = X->data[i * X->rowStride + j * X->colStride];
  void *dX = bfk_malloc(F->numSrc * nrhs * 16), *dY = bfk_malloc(F->numTgt * nrhs * 16);
  BfMatDenseComplex *Y = NULL;
  int rc = (!dX || !dY) ? -2 : 0;
  if (!rc) rc = bfk_h2d(dX, packed, F->numSrc * nrhs * 16);
  if (!rc) rc = bfhipMatMulDevice(A, dX, dY, nrhs, BFHIP_PRECISION_FP64);
  if (!rc) {
    Y = bfMatDenseComplexNew();
    bfMatDenseComplexInit(Y, F->numTgt, nrhs);
    rc = bfk_d2h(Y->data, dY, F->numTgt * nrhs * 16);
  }
  bfk_free(dX);
  bfk_free(dY);
  free(packed);
  if (rc) {
    if (Y) bfMatDenseComplexDeinitAndDealloc(&Y);
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  return &Y->super;
}

BfSize bfhipFacLeafNumFactors(BfMat const *fac, BfSize leaf) {
  if (!fac || fac->vtbl != &bfMatHipButterflyVtable) {
    bfSetError(BF_ERROR_TYPE_ERROR);
    return BF_SIZE_BAD_VALUE;
  }
  BfMatHipButterfly const *F = (BfMatHipButterfly const *)fac;
  if (leaf >= F->leaves.n) {
    bfSetError(BF_ERROR_OUT_OF_RANGE);
    return BF_SIZE_BAD_VALUE;
  }
  return F->leaves.a[leaf].isButterfly ? (BfSize)F->leaves.a[leaf].numFactors : 0;
}

/* stats 4-8 (block kinds, leaf kinds) and 10 (kernel evaluations) */
static void count_stats(BfMatHipButterfly *F) {
  int64_t nL = 0, nK = 0, nI = 0, nD = 0, nB = 0;
  double h0 = 0;
  for (size_t b = 0; b < F->blocks.n; ++b) {
    const int kd = F->blocks.a[b].kind;
    nL += kd == BF_BLOCK_LSTSQ;
    nK += kd == BF_BLOCK_KERNEL;
    nI += kd == BF_BLOCK_IDENTITY;
  }
  for (size_t i = 0; i < F->lstsqJobs.n; ++i)
    h0 += (double)F->lstsqJobs.a[i].p * (F->lstsqJobs.a[i].p + F->lstsqJobs.a[i].n);
  for (size_t i = 0; i < F->kernelJobs.n; ++i) h0 += (double)F->kernelJobs.a[i].m * F->kernelJobs.a[i].n;
  for (size_t l = 0; l < F->leaves.n; ++l) {
    nB += F->leaves.a[l].isButterfly;
    nD += !F->leaves.a[l].isButterfly;
  }
  F->stats[4] = (double)nL;
  F->stats[5] = (double)nK;
  F->stats[6] = (double)nI;
  F->stats[7] = (double)nD;
  F->stats[8] = (double)nB;
  F->stats[10] = h0;
  F->statsCounted = 1;
}

void bfhipFacGetStats(BfMat const *fac, double out[16]) {
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac; /* the counts are cached on first read */
  if (!F->statsCounted) count_stats(F);
  memcpy(out, F->stats, 16 * sizeof(double));
}

void bfhipGetKernelTimes(double out[8]) { bfk_get_times(out); }

void bfhipGetKernelTimeSums(double sums[8], double counts[8]) { bfk_get_time_sums(sums, counts); }

void bfhipResetKernelTimes(void) { bfk_reset_times(); }

int bfhipSetApplyConfig(int wavesPerBlock, int unroll, int mfmaTiles) {
  int rc = bfk_set_apply_config(wavesPerBlock, unroll, mfmaTiles);
  if (rc) bfSetError(BF_ERROR_INVALID_ARGUMENTS);
  return rc;
}

int bfhipSetApplyCompensation(int on) { return bfk_set_compensation(on); }

void bfhipFacPlanSizes(BfMat const *fac, BfSize *numLeaves, BfSize *numFactors, BfSize *numBlocks) {
  BfMatHipButterfly const *F = (BfMatHipButterfly const *)fac;
  *numLeaves = F->leaves.n;
  *numFactors = F->factors.n;
  *numBlocks = F->blocks.n;
}

void bfhipFacPlanExport(BfMat const *fac, int64_t *leafTab, int64_t *facTab, int64_t *blkTab) {
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac;
  for (size_t l = 0; l < F->leaves.n; ++l) {
    BfHipLeaf *lf = &F->leaves.a[l];
    int64_t *L = &leafTab[6 * l];
    L[0] = lf->isButterfly;
    L[1] = lf->row0;
    L[2] = lf->col0;
    L[3] = lf->m;
    L[4] = lf->n;
    L[5] = lf->isButterfly ? lf->numFactors : 0;
  }
  for (size_t f = 0; f < F->factors.n; ++f) {
    BfHipFactor *fa = &F->factors.a[f];
    int64_t *R = &facTab[4 * f];
    R[0] = fa->isDiag;
    R[1] = fa->numBlockRows;
    R[2] = fa->numBlockCols;
    R[3] = fa->numBlocks;
    int64_t *ro = offs_at(F, fa->rowOffStart), *co = offs_at(F, fa->colOffStart);
    for (int64_t b = fa->firstBlock; b < fa->firstBlock + fa->numBlocks; ++b) {
      BfHipBlock *bl = &F->blocks.a[b];
      int64_t *B = &blkTab[7 * b];
      B[0] = bl->rowInd;
      B[1] = bl->colInd;
      B[2] = bl->m;
      B[3] = bl->n;
      B[4] = bl->kind;
      B[5] = ro[bl->rowInd];
      B[6] = co[bl->colInd];
    }
  }
}

static int copy_colmajor_to_rowmajor(int64_t off, int64_t m, int64_t n, const double *dArena, BfComplex *out) {
  BfComplex *tmp = malloc((size_t)(m * n) * sizeof(BfComplex));
  if (!tmp) return -2;
  int rc = bfk_d2h(tmp, dArena + 2 * off, (size_t)(m * n) * 16);
  if (!rc)
    for (int64_t i = 0; i < m; ++i)
      for (int64_t j = 0; j < n; ++j) out[i * n + j] = tmp[j * m + i];
  free(tmp);
  return rc;
}

int bfhipFacBlockData(BfMat const *fac, BfSize idx, BfComplex *out) {
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac;
  if (idx >= F->blocks.n) return -1;
  BfHipBlock *b = &F->blocks.a[idx];
  if (b->dataOff < 0) {
    for (int64_t i = 0; i < b->m; ++i)
      for (int64_t j = 0; j < b->n; ++j) out[i * b->n + j] = (i == j);
    return 0;
  }
  return copy_colmajor_to_rowmajor(b->dataOff, b->m, b->n, F->dArena, out);
}

int bfhipFacBlockGeometry(BfMat const *fac, BfSize idx, BfReal out[12]) {
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac;
  if (idx >= F->blocks.n) return -1;
  BfHipBlock *b = &F->blocks.a[idx];
  for (int i = 0; i < 12; ++i) out[i] = 0;
  if (b->kind == BF_BLOCK_LSTSQ) {
    BfkLstsqJob *J = &F->lstsqJobs.a[b->job];
    double v[12] = {J->ecx, J->ecy, J->er, J->tcx, J->tcy, J->tr, J->ocx, J->ocy, J->orr,
                    (double)J->origOff, (double)J->p, (double)J->n};
    memcpy(out, v, sizeof(v));
  } else if (b->kind == BF_BLOCK_KERNEL) {
    BfkKernelJob *J = &F->kernelJobs.a[b->job];
    double v[12] = {J->scx, J->scy, J->sr, 0, 0, 0, 0, 0, 0, (double)J->srcOff, (double)J->m, (double)J->n};
    memcpy(out, v, sizeof(v));
    out[3] = (double)J->tgtOff;
  }
  return 0;
}

int bfhipFacDenseData(BfMat const *fac, BfSize leafIdx, BfComplex *out) {
  BfMatHipButterfly *F = (BfMatHipButterfly *)fac;
  if (leafIdx >= F->leaves.n || F->leaves.a[leafIdx].isButterfly) return -1;
  BfHipLeaf *l = &F->leaves.a[leafIdx];
  return copy_colmajor_to_rowmajor(l->denseOff, l->m, l->n, F->dArena, out);
}
