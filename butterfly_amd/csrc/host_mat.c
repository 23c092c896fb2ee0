/*
 * Matrix objects: the BfMat dispatchers (mat.c), the host dense complex
 * matrix (mat_dense_complex.c), the dense kernel matrix sampled on the GPU
 * (helm2.c:282-319), and the thin device wrappers of bf.h.
 */
#define _DEFAULT_SOURCE 1
#include <complex.h>
#include <math.h>

#include "bf_host.h"

void bfMatInitBase(BfMat *mat, BfMatVtable *vtbl, BfSize numRows, BfSize numCols) {
  mat->vtbl = vtbl;
  mat->props = BF_MAT_PROPS_NONE;
  mat->numRows = numRows;
  mat->numCols = numCols;
}

/* ------------------------------------------------------ dispatchers */

#define NEED(mat, slot, ret)                                   \
  do {                                                         \
    if (!(mat) || !(mat)->vtbl || !(mat)->vtbl->slot) {        \
      bfSetError((mat) ? BF_ERROR_NOT_IMPLEMENTED : BF_ERROR_INVALID_ARGUMENTS); \
      return ret;                                              \
    }                                                          \
  } while (0)

BfType bfMatGetType(BfMat const *mat) {
  NEED(mat, GetType, BF_TYPE_NONE);
  return mat->vtbl->GetType(mat);
}

BfSize bfMatGetNumRows(BfMat const *mat) {
  if (!mat) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return BF_SIZE_BAD_VALUE; }
  bfk_mat_refresh_shape(mat);
  return mat->numRows;
}

BfSize bfMatGetNumCols(BfMat const *mat) {
  if (!mat) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return BF_SIZE_BAD_VALUE; }
  bfk_mat_refresh_shape(mat);
  return mat->numCols;
}

BfSize bfMatNumBytes(BfMat const *mat) {
  NEED(mat, NumBytes, 0);
  return mat->vtbl->NumBytes(mat);
}

BfMat *bfMatCopy(BfMat const *mat) {
  NEED(mat, Copy, NULL);
  return mat->vtbl->Copy(mat);
}

void bfMatDelete(BfMat **mat) {
  if (!mat || !*mat) return;
  if (!(*mat)->vtbl || !(*mat)->vtbl->Delete) { bfSetError(BF_ERROR_NOT_IMPLEMENTED); return; }
  (*mat)->vtbl->Delete(mat);
}

BfMat *bfMatMul(BfMat const *op1, BfMat const *op2) {
  NEED(op1, Mul, NULL);
  if (!op2) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  bfk_mat_refresh_shape(op1);
  bfk_mat_refresh_shape(op2);
  if (op1->numCols != op2->numRows) { bfSetError(BF_ERROR_INCOMPATIBLE_SHAPES); return NULL; }
  return op1->vtbl->Mul(op1, op2);
}

void bfMatPermuteRows(BfMat *mat, BfPerm const *perm) {
  NEED(mat, PermuteRows, );
  mat->vtbl->PermuteRows(mat, perm);
}

/* ------------------------------------------------- dense complex (host) */

static BfType dc_get_type(BfMat const *m) {
  (void)m;
  return BF_TYPE_MAT_DENSE_COMPLEX;
}

static BfSize dc_num_bytes(BfMat const *m) { return sizeof(BfMatDenseComplex) + m->numRows * m->numCols * 16; }

static void dc_delete(BfMat **m) {
  BfMatDenseComplex *d = (BfMatDenseComplex *)*m;
  bfMatDenseComplexDeinitAndDealloc(&d);
  *m = NULL;
}

static BfMat *dc_copy(BfMat const *m) {
  BfMatDenseComplex const *s = (BfMatDenseComplex const *)m;
  BfMatDenseComplex *d = bfMatDenseComplexNew();
  if (!d) return NULL;
  bfMatDenseComplexInit(d, m->numRows, m->numCols);
  if (!d->data) { free(d); return NULL; }
  for (BfSize i = 0; i < m->numRows; ++i)
    for (BfSize j = 0; j < m->numCols; ++j) d->data[i * m->numCols + j] = s->data[i * s->rowStride + j * s->colStride];
  return &d->super;
}

/* row-major complex GEMM on the host, C = A B (mat_dense_complex.c:1705-1766) */
static BfMat *dc_mul(BfMat const *a, BfMat const *b) {
  if (bfMatGetType(b) != BF_TYPE_MAT_DENSE_COMPLEX) { bfSetError(BF_ERROR_NOT_IMPLEMENTED); return NULL; }
  BfMatDenseComplex const *A = (BfMatDenseComplex const *)a, *B = (BfMatDenseComplex const *)b;
  BfSize m = a->numRows, k = a->numCols, n = b->numCols;
  BfMatDenseComplex *C = bfMatDenseComplexNew();
  if (!C) return NULL;
  bfMatDenseComplexInit(C, m, n);
  if (!C->data) { free(C); return NULL; }
  for (BfSize i = 0; i < m; ++i)
    for (BfSize j = 0; j < n; ++j) C->data[i * n + j] = 0;
  for (BfSize i = 0; i < m; ++i)
    for (BfSize l = 0; l < k; ++l) {
      BfComplex av = A->data[i * A->rowStride + l * A->colStride];
      for (BfSize j = 0; j < n; ++j) C->data[i * n + j] += av * B->data[l * B->rowStride + j * B->colStride];
    }
  return &C->super;
}

/* mat_dense_complex.c:741-774: row i moves to row perm->index[i] */
static void dc_permute_rows(BfMat *m, BfPerm const *perm) {
  BfMatDenseComplex *d = (BfMatDenseComplex *)m;
  if (!perm || perm->size != m->numRows) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  BfSize nr = m->numRows, nc = m->numCols;
  BfComplex *tmp = malloc(nr * nc * sizeof(BfComplex));
  if (!tmp) { bfSetError(BF_ERROR_MEMORY_ERROR); return; }
  for (BfSize i = 0; i < nr; ++i)
    for (BfSize j = 0; j < nc; ++j) tmp[i * nc + j] = d->data[i * d->rowStride + j * d->colStride];
  for (BfSize i = 0; i < nr; ++i) {
    BfSize r = perm->index[i];
    for (BfSize j = 0; j < nc; ++j) d->data[r * d->rowStride + j * d->colStride] = tmp[i * nc + j];
  }
  free(tmp);
}

static BfMatVtable dcVtable = {
    .GetType = dc_get_type,
    .NumBytes = dc_num_bytes,
    .Delete = dc_delete,
    .Copy = dc_copy,
    .Mul = dc_mul,
    .PermuteRows = dc_permute_rows,
};

BfMatDenseComplex *bfMatDenseComplexNew(void) {
  BfMatDenseComplex *m = calloc(1, sizeof(BfMatDenseComplex));
  if (!m) bfSetError(BF_ERROR_MEMORY_ERROR);
  return m;
}

void bfMatDenseComplexInit(BfMatDenseComplex *mat, BfSize numRows, BfSize numCols) {
  bfMatInitBase(&mat->super, &dcVtable, numRows, numCols);
  mat->rowStride = numCols;
  mat->colStride = 1;
  mat->data = malloc((numRows * numCols > 0 ? numRows * numCols : 1) * sizeof(BfComplex));
  mat->pyArray = NULL;
  if (!mat->data) bfSetError(BF_ERROR_MEMORY_ERROR);
}

void bfMatDenseComplexInitViewFromPtr(BfMatDenseComplex *mat, BfSize numRows, BfSize numCols, BfComplex *data) {
  bfMatInitBase(&mat->super, &dcVtable, numRows, numCols);
  mat->super.props = BF_MAT_PROPS_VIEW;
  mat->rowStride = numCols;
  mat->colStride = 1;
  mat->data = data;
  mat->pyArray = NULL;
}

BfMatDenseComplex *bfMatDenseComplexNewRandn(BfSize numRows, BfSize numCols) {
  BfMatDenseComplex *m = bfMatDenseComplexNew();
  if (!m) return NULL;
  bfMatDenseComplexInit(m, numRows, numCols);
  if (!m->data) { free(m); return NULL; }
  bfComplexRandn(numRows * numCols, m->data);
  return m;
}

void bfMatDenseComplexDeinit(BfMatDenseComplex *mat) {
  if (!(mat->super.props & BF_MAT_PROPS_VIEW)) free(mat->data);
  mat->data = NULL;
}

void bfMatDenseComplexDeinitAndDealloc(BfMatDenseComplex **mat) {
  if (!mat || !*mat) return;
  bfMatDenseComplexDeinit(*mat);
  free(*mat);
  *mat = NULL;
}

BfComplex *bfMatDenseComplexGetDataPtr(BfMatDenseComplex *mat) { return mat->data; }
BfMat *bfMatDenseComplexToMat(BfMatDenseComplex *mat) { return &mat->super; }
BfMatDenseComplex *bfMatToMatDenseComplex(BfMat *mat) {
  if (!mat || bfMatGetType(mat) != BF_TYPE_MAT_DENSE_COMPLEX) {
    bfSetError(BF_ERROR_TYPE_ERROR);
    return NULL;
  }
  return (BfMatDenseComplex *)mat;
}

/* --------------------------------------------- dense kernel matrix (GPU) */

/* helm2.c:282-319 for the single layer, the double layer and the combined
 * field (the last two with the source normals Nsrc) and the normal
 * derivative S' of the single layer (with the target normals Ntgt,
 * helm2.c:127-175): rows are targets, columns sources.  Sampled on the GPU
 * in row panels of 64 targets.  The hypersingular layer (normal derivative
 * of D) has no case in the reference either: BF_ERROR_NOT_IMPLEMENTED. */
BfMat *bfHelm2GetKernelMatrix(BfHelm2 const *helm, BfPoints2 const *Xsrc, BfPoints2 const *Xtgt,
                              BfVectors2 const *Nsrc, BfVectors2 const *Ntgt) {
  if (!helm || !Xsrc || !Xtgt) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  if (Nsrc && Nsrc->size != Xsrc->size) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  if (Ntgt && Ntgt->size != Xtgt->size) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  BfkLayer layer;
  memset(&layer, 0, sizeof(layer));
  if (helm->layerPot == BF_LAYER_POTENTIAL_PV_DOUBLE) {
    layer.layer = 1;
  } else if (helm->layerPot == BF_LAYER_POTENTIAL_COMBINED_FIELD) {
    layer.layer = 2;
    layer.ar = creal(helm->alpha);
    layer.ai = cimag(helm->alpha);
    layer.br = creal(helm->beta);
    layer.bi = cimag(helm->beta);
  } else if (helm->layerPot == BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE) {
    layer.layer = 3; /* S' with the target normals (helm2.c:127-175) */
  } else if (helm->layerPot != BF_LAYER_POTENTIAL_SINGLE) {
    bfSetError(BF_ERROR_NOT_IMPLEMENTED);
    return NULL;
  }
  if ((layer.layer == 1 || layer.layer == 2) && !Nsrc) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  if (layer.layer == 3 && !Ntgt) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  BfSize m = Xtgt->size, n = Xsrc->size;
  const BfSize panel = 64;
  BfSize np = (m + panel - 1) / panel;
  BfkKernelJob *jobs = calloc(np ? np : 1, sizeof(BfkKernelJob));
  for (BfSize p = 0; p < np; ++p) {
    jobs[p].k = helm->k;
    jobs[p].srcOff = 0;
    jobs[p].tgtOff = (int64_t)(p * panel);
    jobs[p].m = (int32_t)((p + 1) * panel <= m ? panel : m - p * panel);
    jobs[p].n = (int32_t)n;
    jobs[p].out = (int64_t)(p * panel * n);
  }
  void *dJobs = bfk_malloc(np * sizeof(BfkKernelJob) + 16);
  void *dS = bfk_malloc(n * 16 + 16), *dT = bfk_malloc(m * 16 + 16), *dOut = bfk_malloc(m * n * 16 + 16);
  void *dN = layer.layer == 1 || layer.layer == 2 ? bfk_malloc(n * 16 + 16) : NULL;
  void *dNt = layer.layer == 3 ? bfk_malloc(m * 16 + 16) : NULL;
  BfMatDenseComplex *K = NULL;
  BfComplex *tmp = malloc((m * n > 0 ? m * n : 1) * sizeof(BfComplex));
  int rc = (!dJobs || !dS || !dT || !dOut || !tmp || ((layer.layer == 1 || layer.layer == 2) && !dN) ||
            (layer.layer == 3 && !dNt)) ? -2 : 0;
  if (!rc && dN) rc = bfk_h2d(dN, Nsrc->data, n * 16);
  if (!rc && dNt) rc = bfk_h2d(dNt, Ntgt->data, m * 16);
  if (!rc) rc = bfk_h2d(dJobs, jobs, np * sizeof(BfkKernelJob));
  if (!rc) rc = bfk_h2d(dS, Xsrc->data, n * 16);
  if (!rc) rc = bfk_h2d(dT, Xtgt->data, m * 16);
  if (!rc) rc = bfk_sample(dJobs, (int64_t)np, dS, dT, (const double *)dN, (const double *)dNt, layer, dOut);
  if (!rc) rc = bfk_d2h(tmp, dOut, m * n * 16);
  if (!rc) {
    K = bfMatDenseComplexNew();
    bfMatDenseComplexInit(K, m, n);
    for (BfSize p = 0; p < np; ++p) {
      BfSize r0 = p * panel, pm = (BfSize)jobs[p].m;
      BfComplex const *blk = tmp + p * panel * n;
      for (BfSize j = 0; j < n; ++j)
        for (BfSize i = 0; i < pm; ++i) K->data[(r0 + i) * n + j] = blk[j * pm + i];
    }
  }
  bfk_free(dJobs);
  bfk_free(dS);
  bfk_free(dT);
  bfk_free(dOut);
  bfk_free(dN);
  bfk_free(dNt);
  free(tmp);
  free(jobs);
  if (rc) {
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  return &K->super;
}

/* ---------------------------------------------------- device wrappers */

void *bfhipMalloc(BfSize numBytes) {
  void *p = bfk_malloc(numBytes);
  if (!p) bfSetError(BF_ERROR_MEMORY_ERROR);
  return p;
}

void bfhipFree(void *ptr) { bfk_free(ptr); }

int bfhipMemcpyHtoD(void *dst, void const *src, BfSize numBytes) {
  int rc = bfk_h2d(dst, src, numBytes);
  if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}

int bfhipMemcpyDtoH(void *dst, void const *src, BfSize numBytes) {
  int rc = bfk_d2h(dst, src, numBytes);
  if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}

int bfhipSynchronize(void) {
  int rc = bfk_sync();
  if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}

int bfhipDeviceCount(void) { return bfk_device_count(); }

int bfhipDeviceCopy(void const *src, void *dst, BfSize numBytes) {
  int rc = bfk_copy(src, dst, numBytes);
  if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}

int bfhipHankel0(BfSize n, BfReal const *x, BfReal *out) {
  int rc = bfk_h0((int64_t)n, x, out);
  if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}

int bfhipLstSqBatched(BfSize batch, BfSize p, BfSize n, BfComplex const *A, BfComplex const *B, BfComplex *X) {
  /* the solver writes each X column-major (the apply's block layout) */
  BfComplex *colMajor = malloc((batch * p * n > 0 ? batch * p * n : 1) * sizeof(BfComplex));
  if (!colMajor) { bfSetError(BF_ERROR_MEMORY_ERROR); return -2; }
  int rc = bfk_lstsq_loaded((int64_t)batch, (int)p, (int)n, (const double *)A, (const double *)B,
                            (double *)colMajor);
  if (!rc)
    for (BfSize b = 0; b < batch; ++b)
      for (BfSize i = 0; i < p; ++i)
        for (BfSize j = 0; j < n; ++j) X[(b * p + i) * n + j] = colMajor[b * p * n + j * p + i];
  free(colMajor);
  if (rc == -3) bfSetError(BF_ERROR_NOT_IMPLEMENTED);
  else if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}
