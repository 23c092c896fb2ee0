/*
 * Saving and loading a device factorization, in the framing of the
 * reference's bfMatDump (mat.c:67-73: the type, then the type's payload):
 *   BF_TYPE_MAT_HIP_BUTTERFLY payload
 *     version, k, layer potential, numSrc, numTgt, the options,
 *     numLeaves, then per leaf: row0, col0, m, n and the leaf's matrix:
 *       a dense leaf    BF_TYPE_MAT_DENSE_COMPLEX: m, n, m*n complex row-major
 *       a butterfly     BF_TYPE_MAT_PRODUCT: numFactors, then the factors in
 *                       product order (the one applied last first,
 *                       mat_product.c:123-140), each
 *         block diag    BF_TYPE_MAT_BLOCK_DIAG: numBlocks, row offsets,
 *                       column offsets, the blocks (mat_block_diag.c:241-265)
 *         block coo     BF_TYPE_MAT_BLOCK_COO: row / column block counts,
 *                       numBlocks, row and column indices, offsets, the
 *                       blocks (mat_block_coo.c:260-290)
 *       and every block BF_TYPE_MAT_DENSE_COMPLEX as above or
 *                       BF_TYPE_MAT_IDENTITY: m, n
 *     then numBlocks and one byte per block giving how it was made
 *     (re-expansion, kernel, identity), which the orthogonalization reads.
 * Sizes are BfSize, reals FP64, complex interleaved FP64, all native
 * endian.  Loading rebuilds the plan, uploads the blocks and lays out the
 * apply; the factorization kernels do not run again.
 */
#include "bf_host.h"

#define BF_IO_VERSION 1

static int wr(FILE *fp, const void *p, size_t size, size_t n) { return fwrite(p, size, n, fp) == n ? 0 : -1; }
static int rd(FILE *fp, void *p, size_t size, size_t n) { return fread(p, size, n, fp) == n ? 0 : -1; }
static int wr_size(FILE *fp, BfSize v) { return wr(fp, &v, sizeof v, 1); }
static int wr_type(FILE *fp, BfType t) { return wr(fp, &t, sizeof t, 1); }

/* a column-major m x n arena block written as a row-major dense matrix */
static int dump_dense(FILE *fp, const BfComplex *colMajor, int64_t m, int64_t n) {
  int rc = wr_type(fp, BF_TYPE_MAT_DENSE_COMPLEX);
  rc |= wr_size(fp, (BfSize)m);
  rc |= wr_size(fp, (BfSize)n);
  BfComplex *row = malloc((size_t)(n ? n : 1) * sizeof(BfComplex));
  for (int64_t i = 0; i < m && !rc; ++i) {
    for (int64_t j = 0; j < n; ++j) row[j] = colMajor[j * m + i];
    rc |= wr(fp, row, sizeof(BfComplex), (size_t)n);
  }
  free(row);
  return rc;
}

static int dump_block(FILE *fp, const BfMatHipButterfly *F, const BfComplex *arena, const BfHipBlock *b) {
  if (b->dataOff < 0) {
    int rc = wr_type(fp, BF_TYPE_MAT_IDENTITY);
    rc |= wr_size(fp, (BfSize)b->m);
    rc |= wr_size(fp, (BfSize)b->n);
    return rc;
  }
  (void)F;
  return dump_dense(fp, arena + b->dataOff, b->m, b->n);
}

void bfMatDump(BfMat const *mat, FILE *fp) {
  if (!mat || !fp || mat->vtbl != &bfMatHipButterflyVtable) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return;
  }
  const BfMatHipButterfly *F = (const BfMatHipButterfly *)mat;
  if (!F->dArena) { /* FP64 factors were released */
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return;
  }
  BfComplex *arena = malloc((size_t)(F->arenaEntries ? F->arenaEntries : 1) * sizeof(BfComplex));
  if (!arena) { bfSetError(BF_ERROR_MEMORY_ERROR); return; }
  int rc = bfk_d2h(arena, F->dArena, (size_t)F->arenaEntries * 16);
  rc |= wr_type(fp, BF_TYPE_MAT_HIP_BUTTERFLY);
  rc |= wr_size(fp, BF_IO_VERSION);
  rc |= wr(fp, &F->helm.k, sizeof(double), 1);
  rc |= wr_size(fp, (BfSize)F->helm.layerPot);
  rc |= wr_size(fp, F->numSrc);
  rc |= wr_size(fp, F->numTgt);
  rc |= wr(fp, &F->opts.tol, sizeof(double), 1);
  rc |= wr(fp, &F->opts.rankConst, sizeof(double), 1);
  rc |= wr_size(fp, F->opts.maxDenseSize);
  rc |= wr_size(fp, (BfSize)F->opts.passThrough);
  rc |= wr(fp, &F->opts.pinvRcond, sizeof(double), 1);
  rc |= wr_size(fp, (BfSize)F->leaves.n);
  for (size_t l = 0; l < F->leaves.n && !rc; ++l) {
    const BfHipLeaf *lf = &F->leaves.a[l];
    rc |= wr_size(fp, (BfSize)lf->row0);
    rc |= wr_size(fp, (BfSize)lf->col0);
    rc |= wr_size(fp, (BfSize)lf->m);
    rc |= wr_size(fp, (BfSize)lf->n);
    if (!lf->isButterfly) {
      rc |= dump_dense(fp, arena + lf->denseOff, lf->m, lf->n);
      continue;
    }
    rc |= wr_type(fp, BF_TYPE_MAT_PRODUCT);
    rc |= wr_size(fp, (BfSize)lf->numFactors);
    for (int f = lf->numFactors - 1; f >= 0 && !rc; --f) {
      const BfHipFactor *fa = &F->factors.a[lf->firstFactor + f];
      const int64_t *ro = &F->offs.a[fa->rowOffStart], *co = &F->offs.a[fa->colOffStart];
      if (fa->isDiag) {
        rc |= wr_type(fp, BF_TYPE_MAT_BLOCK_DIAG);
        rc |= wr_size(fp, (BfSize)fa->numBlocks);
      } else {
        rc |= wr_type(fp, BF_TYPE_MAT_BLOCK_COO);
        rc |= wr_size(fp, (BfSize)fa->numBlockRows);
        rc |= wr_size(fp, (BfSize)fa->numBlockCols);
        rc |= wr_size(fp, (BfSize)fa->numBlocks);
        for (int64_t b = 0; b < fa->numBlocks; ++b) rc |= wr_size(fp, (BfSize)F->blocks.a[fa->firstBlock + b].rowInd);
        for (int64_t b = 0; b < fa->numBlocks; ++b) rc |= wr_size(fp, (BfSize)F->blocks.a[fa->firstBlock + b].colInd);
      }
      for (int64_t i = 0; i <= fa->numBlockRows; ++i) rc |= wr_size(fp, (BfSize)ro[i]);
      for (int64_t j = 0; j <= fa->numBlockCols; ++j) rc |= wr_size(fp, (BfSize)co[j]);
      for (int64_t b = 0; b < fa->numBlocks && !rc; ++b)
        rc |= dump_block(fp, F, arena, &F->blocks.a[fa->firstBlock + b]);
    }
  }
  rc |= wr_size(fp, (BfSize)F->blocks.n);
  for (size_t b = 0; b < F->blocks.n && !rc; ++b) {
    uint8_t kd = (uint8_t)F->blocks.a[b].kind;
    rc |= wr(fp, &kd, 1, 1);
  }
  free(arena);
  if (rc) bfSetError(BF_ERROR_FILE_ERROR);
}

void bfMatSave(BfMat const *mat, char const *path) {
  FILE *fp = path ? fopen(path, "wb") : NULL;
  if (!fp) {
    bfSetError(BF_ERROR_FILE_ERROR);
    return;
  }
  bfMatDump(mat, fp);
  if (fclose(fp)) bfSetError(BF_ERROR_FILE_ERROR);
}

/* ------------------------------------------------------------- loading */

typedef BF_VEC(BfComplex) CplxVec;

static int rd_size(FILE *fp, BfSize *v) { return rd(fp, v, sizeof *v, 1); }

static int expect_type(FILE *fp, BfType want) {
  BfType t;
  if (rd(fp, &t, sizeof t, 1)) return -1;
  return t == want ? 0 : -1;
}

/* a dense block into the host arena (column-major); returns its offset */
static int load_dense(FILE *fp, CplxVec *arena, int64_t m, int64_t n, int64_t *off) {
  BfSize mm, nn;
  if (rd_size(fp, &mm) || rd_size(fp, &nn) || (int64_t)mm != m || (int64_t)nn != n) return -1;
  *off = (int64_t)arena->n;
  BfComplex *row = malloc((size_t)(n ? n : 1) * sizeof(BfComplex));
  size_t need = arena->n + (size_t)(m * n);
  if (need > arena->cap) {
    size_t nc = arena->cap ? arena->cap : 1024;
    while (nc < need) nc *= 2;
    BfComplex *na = realloc(arena->a, nc * sizeof(BfComplex));
    if (!na) { free(row); return -2; }
    arena->a = na;
    arena->cap = nc;
  }
  BfComplex *dst = arena->a + arena->n;
  int rc = 0;
  for (int64_t i = 0; i < m && !rc; ++i) {
    rc = rd(fp, row, sizeof(BfComplex), (size_t)n);
    for (int64_t j = 0; j < n && !rc; ++j) dst[j * m + i] = row[j];
  }
  arena->n = need;
  free(row);
  return rc;
}

static int load_block(FILE *fp, CplxVec *arena, BfHipBlock *b) {
  BfType t;
  if (rd(fp, &t, sizeof t, 1)) return -1;
  if (t == BF_TYPE_MAT_IDENTITY) {
    BfSize mm, nn;
    if (rd_size(fp, &mm) || rd_size(fp, &nn) || (int64_t)mm != b->m || (int64_t)nn != b->n || mm != nn) return -1;
    b->dataOff = -1;
    b->kind = BF_BLOCK_IDENTITY;
    return 0;
  }
  if (t != BF_TYPE_MAT_DENSE_COMPLEX) return -1;
  b->kind = BF_BLOCK_KERNEL;
  return load_dense(fp, arena, b->m, b->n, &b->dataOff);
}



BfMat *bfhipFacLoad(FILE *fp) {
  if (!fp) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return NULL;
  }
  if (bfk_init()) {
    bfSetError(BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  BfMatHipButterfly *F = calloc(1, sizeof(BfMatHipButterfly));
  CplxVec arena;
  memset(&arena, 0, sizeof(arena));
  int rc = expect_type(fp, BF_TYPE_MAT_HIP_BUTTERFLY);
  BfSize version = 0, lp = 0, pt = 0, numLeaves = 0, v = 0;
  rc = rc || rd_size(fp, &version) || version != BF_IO_VERSION;
  rc = rc || rd(fp, &F->helm.k, sizeof(double), 1) || rd_size(fp, &lp);
  rc = rc || rd_size(fp, &F->numSrc) || rd_size(fp, &F->numTgt);
  rc = rc || rd(fp, &F->opts.tol, sizeof(double), 1) || rd(fp, &F->opts.rankConst, sizeof(double), 1);
  rc = rc || rd_size(fp, &F->opts.maxDenseSize) || rd_size(fp, &pt);
  rc = rc || rd(fp, &F->opts.pinvRcond, sizeof(double), 1) || rd_size(fp, &numLeaves);
  F->helm.layerPot = (BfLayerPotential)lp;
  F->opts.passThrough = pt != 0;
  for (BfSize l = 0; l < numLeaves && !rc; ++l) {
    BfHipLeaf lf;
    memset(&lf, 0, sizeof(lf));
    BfSize r0, c0, m, n;
    rc = rd_size(fp, &r0) || rd_size(fp, &c0) || rd_size(fp, &m) || rd_size(fp, &n);
    if (rc) break;
    lf.row0 = (int64_t)r0;
    lf.col0 = (int64_t)c0;
    lf.m = (int64_t)m;
    lf.n = (int64_t)n;
    BfType t;
    if (rd(fp, &t, sizeof t, 1)) { rc = -1; break; }
    if (t == BF_TYPE_MAT_DENSE_COMPLEX) {
      lf.isButterfly = 0;
      rc = load_dense(fp, &arena, lf.m, lf.n, &lf.denseOff);
      BF_VEC_PUSH(F->leaves, lf);
      continue;
    }
    if (t != BF_TYPE_MAT_PRODUCT) { rc = -1; break; }
    BfSize nf;
    if (rd_size(fp, &nf) || nf < 1 || nf > 4096) { rc = -1; break; }
    lf.isButterfly = 1;
    lf.numFactors = (int32_t)nf;
    /* the file holds the factors in product order; the plan keeps them in
     * the order they are applied */
    BfHipFactor *fs = calloc(nf, sizeof(BfHipFactor));
    BfHipBlockVec *bs = calloc(nf, sizeof(BfHipBlockVec));
    BfI64Vec *os = calloc(nf, sizeof(BfI64Vec));
    for (BfSize q = 0; q < nf && !rc; ++q) {
      const BfSize f = nf - 1 - q;
      BfHipFactor *fa = &fs[f];
      BfType ft;
      if (rd(fp, &ft, sizeof ft, 1)) { rc = -1; break; }
      BfSize nbr, nbc, nb;
      int32_t *ri = NULL, *ci = NULL;
      if (ft == BF_TYPE_MAT_BLOCK_DIAG) {
        if (rd_size(fp, &nb)) { rc = -1; break; }
        nbr = nbc = nb;
        fa->isDiag = 1;
      } else if (ft == BF_TYPE_MAT_BLOCK_COO) {
        if (rd_size(fp, &nbr) || rd_size(fp, &nbc) || rd_size(fp, &nb)) { rc = -1; break; }
        ri = malloc((nb ? nb : 1) * sizeof(int32_t));
        ci = malloc((nb ? nb : 1) * sizeof(int32_t));
        for (BfSize b = 0; b < nb && !rc; ++b) { rc = rd_size(fp, &v); ri[b] = (int32_t)v; }
        for (BfSize b = 0; b < nb && !rc; ++b) { rc = rd_size(fp, &v); ci[b] = (int32_t)v; }
        fa->isDiag = 0;
      } else {
        rc = -1;
        break;
      }
      fa->numBlockRows = (int64_t)nbr;
      fa->numBlockCols = (int64_t)nbc;
      fa->numBlocks = (int64_t)nb;
      for (BfSize i = 0; i <= nbr + nbc + 1 && !rc; ++i) {
        rc = rd_size(fp, &v);
        BF_VEC_PUSH(os[f], (int64_t)v);
      }
      const int64_t *ro = os[f].a, *co = os[f].a + nbr + 1;
      for (BfSize b = 0; b < nb && !rc; ++b) {
        BfHipBlock blk;
        memset(&blk, 0, sizeof(blk));
        blk.rowInd = fa->isDiag ? (int32_t)b : ri[b];
        blk.colInd = fa->isDiag ? (int32_t)b : ci[b];
        if (blk.rowInd < 0 || (BfSize)blk.rowInd >= nbr || blk.colInd < 0 || (BfSize)blk.colInd >= nbc) {
          rc = -1;
          break;
        }
        blk.m = (int32_t)(ro[blk.rowInd + 1] - ro[blk.rowInd]);
        blk.n = (int32_t)(co[blk.colInd + 1] - co[blk.colInd]);
        blk.job = -1;
        rc = load_block(fp, &arena, &blk);
        BF_VEC_PUSH(bs[f], blk);
      }
      free(ri);
      free(ci);
    }
    lf.firstFactor = (int64_t)F->factors.n;
    for (BfSize f = 0; f < nf && !rc; ++f) {
      BfHipFactor fa = fs[f];
      fa.firstBlock = (int64_t)F->blocks.n;
      fa.rowOffStart = (int64_t)F->offs.n;
      fa.colOffStart = fa.rowOffStart + fa.numBlockRows + 1;
      for (size_t i = 0; i < os[f].n; ++i) BF_VEC_PUSH(F->offs, os[f].a[i]);
      for (size_t b = 0; b < bs[f].n; ++b) BF_VEC_PUSH(F->blocks, bs[f].a[b]);
      BF_VEC_PUSH(F->factors, fa);
    }
    for (BfSize f = 0; f < nf; ++f) {
      BF_VEC_FREE(bs[f]);
      BF_VEC_FREE(os[f]);
    }
    free(fs);
    free(bs);
    free(os);
    BF_VEC_PUSH(F->leaves, lf);
  }
  BfSize nbk = 0;
  if (!rc && !rd_size(fp, &nbk) && nbk == F->blocks.n) {
    for (size_t b = 0; b < F->blocks.n && !rc; ++b) {
      uint8_t kd;
      rc = rd(fp, &kd, 1, 1);
      if (!rc && F->blocks.a[b].dataOff >= 0 && kd <= BF_BLOCK_IDENTITY && kd != BF_BLOCK_IDENTITY)
        F->blocks.a[b].kind = kd;
    }
  }
  if (!rc) {
    F->arenaEntries = (int64_t)arena.n;
    bfMatInitBase(&F->super, &bfMatHipButterflyVtable, F->numTgt, F->numSrc);
    F->dArena = bfk_malloc((size_t)(arena.n ? arena.n : 1) * 16);
    if (!F->dArena) rc = -2;
    if (!rc) rc = bfk_h2d(F->dArena, arena.a, arena.n * 16);
    if (!rc) rc = bfMatHipButterflyFinish(F);
  }
  BF_VEC_FREE(arena);
  if (rc) {
    BfMat *m = &F->super;
    F->super.vtbl = &bfMatHipButterflyVtable;
    bfMatHipButterflyDelete(&m);
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_FILE_ERROR);
    return NULL;
  }
  return &F->super;
}

BfMat *bfhipFacLoadFile(char const *path) {
  FILE *fp = path ? fopen(path, "rb") : NULL;
  if (!fp) {
    bfSetError(BF_ERROR_FILE_ERROR);
    return NULL;
  }
  BfMat *m = bfhipFacLoad(fp);
  fclose(fp);
  return m;
}
