/*
 * The reference's block matrix algebra (ptr_array.c, mat_product.c,
 * mat_block_diag.c, mat_block_coo.c, mat_block_dense.c): host containers
 * of matrices with the reference's names, fields and ownership rules.
 *
 * Applying one runs on the GPU: the leaves of the container (dense complex
 * matrices, possibly inside nested containers) are copied column-major
 * into a device product -- a block container is one factor whose leaves
 * are its blocks at their offsets, a product is one factor per factor --
 * and applied by the stage kernels of the butterfly apply
 * (bfk_product_from_blocks).  A product with a factor that is not made of
 * dense leaves (a device factorization, say) is applied factor by factor
 * through bfMatMul, last factor first (mat_product.c:211-245; the
 * reference returns NULL for a one-factor product there, here it returns
 * the product).
 */
#include "bf_host.h"

/* ------------------------------------------------------- pointer arrays */

void bfInitPtrArray(BfPtrArray *arr, BfSize capacity) {
  if (!arr) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  arr->capacity = capacity ? capacity : 1;
  arr->num_elts = 0;
  arr->isView = false;
  arr->data = malloc(arr->capacity * sizeof(void *));
  if (!arr->data) bfSetError(BF_ERROR_MEMORY_ERROR);
}

void bfInitPtrArrayWithDefaultCapacity(BfPtrArray *arr) { bfInitPtrArray(arr, 16); }

void bfPtrArrayDeinit(BfPtrArray *arr) {
  if (!arr) return;
  if (!arr->isView) free(arr->data);
  arr->data = NULL;
  arr->capacity = arr->num_elts = 0;
}

BfSize bfPtrArraySize(BfPtrArray const *arr) { return arr ? arr->num_elts : 0; }

bool bfPtrArrayIsEmpty(BfPtrArray const *arr) { return bfPtrArraySize(arr) == 0; }

void bfPtrArrayAppend(BfPtrArray *arr, void *ptr) {
  if (!arr) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  if (arr->num_elts == arr->capacity) {
    BfSize cap = arr->capacity ? 2 * arr->capacity : 16;
    void **d = realloc(arr->data, cap * sizeof(void *));
    if (!d) { bfSetError(BF_ERROR_MEMORY_ERROR); return; }
    arr->data = d;
    arr->capacity = cap;
  }
  arr->data[arr->num_elts++] = ptr;
}

void *bfPtrArrayGet(BfPtrArray const *arr, BfSize pos) {
  if (!arr || pos >= arr->num_elts) { bfSetError(BF_ERROR_OUT_OF_RANGE); return NULL; }
  return arr->data[pos];
}

void *bfPtrArrayPopLast(BfPtrArray *arr) {
  if (!arr || arr->num_elts == 0) { bfSetError(BF_ERROR_OUT_OF_RANGE); return NULL; }
  return arr->data[--arr->num_elts];
}

BfMat *bfMatGet(BfMat *mat, BfPolicy policy) {
  if (!mat) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  return policy == BF_POLICY_COPY ? bfMatCopy(mat) : mat;
}

/* ---------------------------------------------------- packing the leaves */

typedef BF_VEC(BfkBlockRef) RefVec;
typedef BF_VEC(BfComplex *) BufVec;
typedef BF_VEC(BfMat const *) MatVec;

static int is_block(BfType t) {
  return t == BF_TYPE_MAT_BLOCK_DIAG || t == BF_TYPE_MAT_BLOCK_COO || t == BF_TYPE_MAT_BLOCK_DENSE;
}

/* block b of a container: its block row / column (r, c) and whether set */
static int block_at(BfMat const *m, BfSize b, BfSize *r, BfSize *c) {
  switch (bfMatGetType(m)) {
    case BF_TYPE_MAT_BLOCK_DIAG: {
      const BfMatBlockDiag *d = (const BfMatBlockDiag *)m;
      *r = *c = b;
      return b < (d->numBlockRows < d->numBlockCols ? d->numBlockRows : d->numBlockCols);
    }
    case BF_TYPE_MAT_BLOCK_COO: {
      const BfMatBlockCoo *o = (const BfMatBlockCoo *)m;
      if (b >= o->numBlocks) return 0;
      *r = o->rowInd[b];
      *c = o->colInd[b];
      return *r < o->numBlockRows && *c < o->numBlockCols;
    }
    case BF_TYPE_MAT_BLOCK_DENSE: {
      const BfMatBlockDense *e = (const BfMatBlockDense *)m;
      if (b >= e->numBlockRows * e->numBlockCols) return 0;
      *r = b / e->numBlockCols;
      *c = b % e->numBlockCols;
      return 1;
    }
    default:
      return 0;
  }
}

static BfSize num_blocks(BfMat const *m) {
  switch (bfMatGetType(m)) {
    case BF_TYPE_MAT_BLOCK_DIAG: {
      const BfMatBlockDiag *d = (const BfMatBlockDiag *)m;
      return d->numBlockRows < d->numBlockCols ? d->numBlockRows : d->numBlockCols;
    }
    case BF_TYPE_MAT_BLOCK_COO: return ((const BfMatBlockCoo *)m)->numBlocks;
    case BF_TYPE_MAT_BLOCK_DENSE: {
      const BfMatBlockDense *e = (const BfMatBlockDense *)m;
      return e->numBlockRows * e->numBlockCols;
    }
    default: return 0;
  }
}

/* the dense leaves of `m` placed at (i0, j0), as column-major copies
 * (kept in `bufs` until the product is built); -1 for a leaf that is not
 * dense complex or a container of such (or a block whose shape is not its
 * cell's), -2 out of memory, -4 a block index outside the grid */
static int collect(BfMat const *m, int64_t i0, int64_t j0, RefVec *refs, BufVec *bufs) {
  if (!m) return -1;
  const BfType t = bfMatGetType(m);
  if (t == BF_TYPE_MAT_DENSE_COMPLEX) {
    const BfMatDenseComplex *d = (const BfMatDenseComplex *)m;
    const int64_t rows = (int64_t)m->numRows, cols = (int64_t)m->numCols;
    if (rows == 0 || cols == 0) return 0;
    BfComplex *c = malloc((size_t)(rows * cols) * sizeof(BfComplex));
    if (!c) return -2;
    for (int64_t j = 0; j < cols; ++j)
      for (int64_t i = 0; i < rows; ++i) c[j * rows + i] = d->data[i * d->rowStride + j * d->colStride];
    BF_VEC_PUSH(*bufs, c);
    BfkBlockRef r = {i0, j0, rows, cols, c};
    BF_VEC_PUSH(*refs, r);
    return 0;
  }
  if (!is_block(t)) return -1;
  const BfMatBlock *B = (const BfMatBlock *)m;
  const BfSize nb = num_blocks(m);
  for (BfSize b = 0; b < nb; ++b) {
    BfSize r, c;
    if (!block_at(m, b, &r, &c)) return -4; /* a block index outside the grid */
    if (!B->block[b]) continue; /* an unset block is zero */
    const int64_t bi = (int64_t)B->rowOffset[r], bj = (int64_t)B->colOffset[c];
    if (B->block[b]->numRows != B->rowOffset[r + 1] - B->rowOffset[r] ||
        B->block[b]->numCols != B->colOffset[c + 1] - B->colOffset[c])
      return -1;
    const int rc = collect(B->block[b], i0 + bi, j0 + bj, refs, bufs);
    if (rc) return rc;
  }
  return 0;
}

/* the factors of `m` in order: a product's factors (nested products
 * flattened), or `m` itself */
static void factors_of(BfMat const *m, MatVec *out) {
  if (bfMatGetType(m) == BF_TYPE_MAT_PRODUCT) {
    const BfMatProduct *p = (const BfMatProduct *)m;
    for (BfSize i = 0; i < p->factorArr.num_elts; ++i) factors_of((BfMat const *)p->factorArr.data[i], out);
  } else {
    BF_VEC_PUSH(*out, m);
  }
}

/* the device product of `m`, or NULL with the error set (NOT_IMPLEMENTED:
 * a leaf that is not dense) */
static BfMat *pack(BfMat const *m) {
  MatVec fs;
  memset(&fs, 0, sizeof(fs));
  factors_of(m, &fs);
  const size_t nf = fs.n;
  RefVec *refs = calloc(nf ? nf : 1, sizeof(RefVec));
  BufVec bufs;
  memset(&bufs, 0, sizeof(bufs));
  int64_t *rows = malloc((nf ? nf : 1) * sizeof(int64_t)), *cols = malloc((nf ? nf : 1) * sizeof(int64_t));
  size_t *nb = malloc((nf ? nf : 1) * sizeof(size_t));
  const BfkBlockRef **bl = malloc((nf ? nf : 1) * sizeof(BfkBlockRef *));
  int rc = nf ? 0 : -1;
  for (size_t f = 0; f < nf && !rc; ++f) {
    rows[f] = (int64_t)bfMatGetNumRows(fs.a[f]);
    cols[f] = (int64_t)bfMatGetNumCols(fs.a[f]);
    if (f > 0 && cols[f - 1] != rows[f]) rc = -3;
    if (!rc) rc = collect(fs.a[f], 0, 0, &refs[f], &bufs);
    nb[f] = refs[f].n;
    bl[f] = refs[f].a;
  }
  BfMat *P = NULL;
  if (!rc) P = bfk_product_from_blocks(nf, rows, cols, nb, bl);
  else
    bfSetError(rc == -2   ? BF_ERROR_MEMORY_ERROR
               : rc == -3 ? BF_ERROR_INCOMPATIBLE_SHAPES
               : rc == -4 ? BF_ERROR_INVALID_ARGUMENTS
                          : BF_ERROR_NOT_IMPLEMENTED);
  for (size_t i = 0; i < bufs.n; ++i) free(bufs.a[i]);
  BF_VEC_FREE(bufs);
  for (size_t f = 0; f < nf; ++f) BF_VEC_FREE(refs[f]);
  free(refs); free(rows); free(cols); free(nb); free(bl);
  BF_VEC_FREE(fs);
  return P;
}

static int packable(BfMat const *m) {
  const BfType t = bfMatGetType(m);
  if (t == BF_TYPE_MAT_DENSE_COMPLEX) return 1;
  if (t == BF_TYPE_MAT_PRODUCT) {
    const BfMatProduct *p = (const BfMatProduct *)m;
    for (BfSize i = 0; i < p->factorArr.num_elts; ++i)
      if (!packable((BfMat const *)p->factorArr.data[i])) return 0;
    return p->factorArr.num_elts > 0;
  }
  if (!is_block(t)) return 0;
  const BfMatBlock *B = (const BfMatBlock *)m;
  for (BfSize b = 0; b < num_blocks(m); ++b)
    if (B->block[b] && !packable(B->block[b])) return 0;
  return 1;
}

BfMat *bfhipMatPack(BfMat const *mat) {
  if (!mat) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  return pack(mat);
}

/* packed once for this apply, applied on the device, released */
static BfMat *packed_mul(BfMat const *A, BfMat const *X) {
  BfMat *P = pack(A);
  if (!P) return NULL;
  BfMat *Y = bfMatMul(P, X);
  bfMatDelete(&P);
  return Y;
}

/* ------------------------------------------------------------- products */

static BfType prod_get_type(BfMat const *m) { (void)m; return BF_TYPE_MAT_PRODUCT; }

static BfSize prod_num_bytes(BfMat const *m) {
  const BfMatProduct *p = (const BfMatProduct *)m;
  BfSize b = sizeof(BfMatProduct);
  for (BfSize i = 0; i < p->factorArr.num_elts; ++i) b += bfMatNumBytes((BfMat const *)p->factorArr.data[i]);
  return b;
}

static void prod_delete(BfMat **m) {
  BfMatProduct *p = (BfMatProduct *)*m;
  bfMatProductDeinitAndDealloc(&p);
  *m = NULL;
}

static BfMat *prod_mul(BfMat const *A, BfMat const *X) {
  const BfMatProduct *p = (const BfMatProduct *)A;
  if (p->factorArr.num_elts == 0) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  if (packable(A)) return packed_mul(A, X);
  /* factor by factor, the last first; factors of dense leaves still run
   * on the device */
  BfMat *cur = NULL;
  for (BfSize k = p->factorArr.num_elts; k-- > 0;) {
    BfMat const *f = (BfMat const *)p->factorArr.data[k];
    BfMat const *in = cur ? cur : X;
    BfMat *next = packable(f) ? packed_mul(f, in) : bfMatMul(f, in);
    if (cur) bfMatDelete(&cur);
    if (!next) return NULL;
    cur = next;
  }
  return cur;
}

static BfMatVtable prodVtable = {prod_get_type, prod_num_bytes, prod_delete, NULL, prod_mul, NULL};

BfMatProduct *bfMatProductNew(void) {
  BfMatProduct *p = calloc(1, sizeof(BfMatProduct));
  if (!p) bfSetError(BF_ERROR_MEMORY_ERROR);
  return p;
}

void bfMatProductInit(BfMatProduct *prod) {
  if (!prod) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  bfMatInitBase(&prod->super, &prodVtable, 0, 0);
  bfInitPtrArrayWithDefaultCapacity(&prod->factorArr);
}

void bfMatProductDeinit(BfMatProduct *prod) {
  if (!prod) return;
  for (BfSize i = 0; i < prod->factorArr.num_elts; ++i) {
    BfMat *f = (BfMat *)prod->factorArr.data[i];
    bfMatDelete(&f);
  }
  bfPtrArrayDeinit(&prod->factorArr);
}

void bfMatProductDealloc(BfMatProduct **prod) {
  if (!prod) return;
  free(*prod);
  *prod = NULL;
}

void bfMatProductDeinitAndDealloc(BfMatProduct **prod) {
  if (!prod || !*prod) return;
  bfMatProductDeinit(*prod);
  bfMatProductDealloc(prod);
}

BfSize bfMatProductNumFactors(BfMatProduct const *prod) { return prod ? prod->factorArr.num_elts : 0; }

BfMat *bfMatProductGetFactor(BfMatProduct *prod, BfSize i) {
  return prod ? (BfMat *)bfPtrArrayGet(&prod->factorArr, i) : NULL;
}

BfMat const *bfMatProductGetFactorConst(BfMatProduct const *prod, BfSize i) {
  return prod ? (BfMat const *)bfPtrArrayGet(&prod->factorArr, i) : NULL;
}

void bfMatProductPostMultiply(BfMatProduct *prod, BfMat *mat) {
  if (!prod || !mat) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  if (prod->factorArr.num_elts > 0 && bfMatGetNumCols(&prod->super) != bfMatGetNumRows(mat)) {
    bfSetError(BF_ERROR_INCOMPATIBLE_SHAPES);
    return;
  }
  bfPtrArrayAppend(&prod->factorArr, mat);
  bfk_mat_refresh_shape(&prod->super);
}

BfMat *bfMatProductPopLastFactor(BfMatProduct *prod) {
  if (!prod) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  BfMat *f = (BfMat *)bfPtrArrayPopLast(&prod->factorArr);
  bfk_mat_refresh_shape(&prod->super);
  return f;
}

BfMat *bfMatProductToMat(BfMatProduct *prod) { return prod ? &prod->super : NULL; }

BfMatProduct *bfMatToMatProduct(BfMat *mat) {
  if (!mat || bfMatGetType(mat) != BF_TYPE_MAT_PRODUCT) { bfSetError(BF_ERROR_TYPE_ERROR); return NULL; }
  return (BfMatProduct *)mat;
}

/* ------------------------------------------------------ block containers */

/* the base arrays of a container with nb blocks on an r x c block grid */
static int block_init(BfMatBlock *B, BfMatVtable *vt, BfSize nb, BfSize r, BfSize c, bool **owns) {
  bfMatInitBase(&B->super, vt, 0, 0);
  B->vtbl = NULL;
  B->block = calloc(nb ? nb : 1, sizeof(BfMat *));
  B->rowOffset = calloc(r + 1, sizeof(BfSize));
  B->colOffset = calloc(c + 1, sizeof(BfSize));
  *owns = calloc(nb ? nb : 1, sizeof(bool));
  if (!B->block || !B->rowOffset || !B->colOffset || !*owns) {
    bfSetError(BF_ERROR_MEMORY_ERROR);
    return -1;
  }
  return 0;
}

/* blocks set by the caller through the arrays are owned, as in the
 * reference (bfMatBlockDeinit deletes every block) */
static void block_deinit(BfMatBlock *B, BfSize nb, bool *owns, int ownAll) {
  for (BfSize b = 0; B->block && b < nb; ++b)
    if (B->block[b] && (ownAll || (owns && owns[b]))) bfMatDelete(&B->block[b]);
  free(B->block);
  free(B->rowOffset);
  free(B->colOffset);
  free(owns);
  B->block = NULL;
  B->rowOffset = B->colOffset = NULL;
}

/* shape from the offsets */
static void block_shape(BfMatBlock *B, BfSize r, BfSize c) {
  B->super.numRows = B->rowOffset[r];
  B->super.numCols = B->colOffset[c];
}

/* offsets from the set blocks' shapes (block-diagonal and dense grids) */
static void offsets_from_blocks(BfMatBlock *B, BfSize nb, BfSize r, BfSize c, BfMat const *owner) {
  BfSize *rs = calloc(r, sizeof(BfSize)), *cs = calloc(c, sizeof(BfSize));
  for (BfSize b = 0; b < nb; ++b) {
    BfSize i, j;
    if (!B->block[b] || !block_at(owner, b, &i, &j)) continue;
    rs[i] = B->block[b]->numRows;
    cs[j] = B->block[b]->numCols;
  }
  B->rowOffset[0] = B->colOffset[0] = 0;
  for (BfSize i = 0; i < r; ++i) B->rowOffset[i + 1] = B->rowOffset[i] + rs[i];
  for (BfSize j = 0; j < c; ++j) B->colOffset[j + 1] = B->colOffset[j] + cs[j];
  free(rs);
  free(cs);
  block_shape(B, r, c);
}

static BfSize blocks_bytes(BfMatBlock const *B, BfSize nb, BfSize base) {
  BfSize s = base;
  for (BfSize b = 0; b < nb; ++b)
    if (B->block[b]) s += bfMatNumBytes(B->block[b]);
  return s;
}

/* block-diagonal */

static BfType diag_get_type(BfMat const *m) { (void)m; return BF_TYPE_MAT_BLOCK_DIAG; }
static BfSize diag_num_bytes(BfMat const *m) {
  return blocks_bytes((BfMatBlock const *)m, num_blocks(m), sizeof(BfMatBlockDiag));
}
static void diag_delete(BfMat **m) {
  BfMatBlockDiag *d = (BfMatBlockDiag *)*m;
  bfMatBlockDiagDeinitAndDealloc(&d);
  *m = NULL;
}
static BfMatVtable diagVtable = {diag_get_type, diag_num_bytes, diag_delete, NULL, packed_mul, NULL};

BfMatBlockDiag *bfMatBlockDiagNew(void) {
  BfMatBlockDiag *d = calloc(1, sizeof(BfMatBlockDiag));
  if (!d) bfSetError(BF_ERROR_MEMORY_ERROR);
  return d;
}

void bfMatBlockDiagInit(BfMatBlockDiag *mat, BfSize numBlockRows, BfSize numBlockCols) {
  if (!mat) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  const BfSize nb = numBlockRows < numBlockCols ? numBlockRows : numBlockCols;
  mat->numBlockRows = numBlockRows;
  mat->numBlockCols = numBlockCols;
  block_init(&mat->super, &diagVtable, nb, numBlockRows, numBlockCols, &mat->ownsBlock);
}

BfMatBlockDiag *bfMatBlockDiagNewFromBlocks(BfPtrArray *blocks, BfPolicy policy) {
  const BfSize nb = bfPtrArraySize(blocks);
  for (BfSize i = 0; i < nb; ++i)
    if (!blocks->data[i]) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  BfMatBlockDiag *d = bfMatBlockDiagNew();
  if (!d) return NULL;
  bfMatBlockDiagInit(d, nb, nb);
  for (BfSize i = 0; i < nb; ++i) {
    d->super.block[i] = bfMatGet((BfMat *)blocks->data[i], policy);
    d->ownsBlock[i] = policy != BF_POLICY_VIEW;
  }
  offsets_from_blocks(&d->super, nb, nb, nb, &d->super.super);
  return d;
}

void bfMatBlockDiagDeinit(BfMatBlockDiag *mat) {
  if (!mat) return;
  block_deinit(&mat->super, num_blocks(&mat->super.super), mat->ownsBlock, 0);
  mat->ownsBlock = NULL;
}

void bfMatBlockDiagDealloc(BfMatBlockDiag **mat) {
  if (!mat) return;
  free(*mat);
  *mat = NULL;
}

void bfMatBlockDiagDeinitAndDealloc(BfMatBlockDiag **mat) {
  if (!mat || !*mat) return;
  bfMatBlockDiagDeinit(*mat);
  bfMatBlockDiagDealloc(mat);
}

BfSize bfMatBlockDiagNumBlocks(BfMatBlockDiag const *mat) { return mat ? num_blocks(&mat->super.super) : 0; }

BfMat *bfMatBlockDiagGetBlock(BfMatBlockDiag *mat, BfSize i) {
  if (!mat || i >= bfMatBlockDiagNumBlocks(mat)) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  return mat->super.block[i];
}

BfMat const *bfMatBlockDiagGetBlockConst(BfMatBlockDiag const *mat, BfSize i) {
  if (!mat || i >= bfMatBlockDiagNumBlocks(mat)) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  return mat->super.block[i];
}

void bfMatBlockDiagSetBlock(BfMatBlockDiag *mat, BfSize i, BfMat *block) {
  if (!mat || i >= bfMatBlockDiagNumBlocks(mat)) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  if (mat->super.block[i] && mat->ownsBlock[i] && mat->super.block[i] != block) bfMatDelete(&mat->super.block[i]);
  mat->super.block[i] = block;
  mat->ownsBlock[i] = true;
  offsets_from_blocks(&mat->super, num_blocks(&mat->super.super), mat->numBlockRows, mat->numBlockCols,
                      &mat->super.super);
}

BfMat *bfMatBlockDiagToMat(BfMatBlockDiag *mat) { return mat ? &mat->super.super : NULL; }

BfMatBlockDiag *bfMatToMatBlockDiag(BfMat *mat) {
  if (!mat || bfMatGetType(mat) != BF_TYPE_MAT_BLOCK_DIAG) { bfSetError(BF_ERROR_TYPE_ERROR); return NULL; }
  return (BfMatBlockDiag *)mat;
}

/* block-COO */

static BfType coo_get_type(BfMat const *m) { (void)m; return BF_TYPE_MAT_BLOCK_COO; }
static BfSize coo_num_bytes(BfMat const *m) {
  return blocks_bytes((BfMatBlock const *)m, num_blocks(m), sizeof(BfMatBlockCoo));
}
static void coo_delete(BfMat **m) {
  BfMatBlockCoo *o = (BfMatBlockCoo *)*m;
  bfMatBlockCooDeinitAndDealloc(&o);
  *m = NULL;
}
/* the caller may have filled the offsets after Init: the shape is read
 * from them at every apply */
static BfMat *coo_mul(BfMat const *A, BfMat const *X) {
  BfMatBlockCoo *o = (BfMatBlockCoo *)A;
  block_shape(&o->super, o->numBlockRows, o->numBlockCols);
  if (X->numRows != A->numCols) { bfSetError(BF_ERROR_INCOMPATIBLE_SHAPES); return NULL; }
  return packed_mul(A, X);
}
static BfMatVtable cooVtable = {coo_get_type, coo_num_bytes, coo_delete, NULL, coo_mul, NULL};

BfMatBlockCoo *bfMatBlockCooNew(void) {
  BfMatBlockCoo *o = calloc(1, sizeof(BfMatBlockCoo));
  if (!o) bfSetError(BF_ERROR_MEMORY_ERROR);
  return o;
}

void bfMatBlockCooInit(BfMatBlockCoo *mat, BfSize numBlockRows, BfSize numBlockCols, BfSize numBlocks) {
  if (!mat) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  mat->numBlocks = numBlocks;
  mat->numBlockRows = numBlockRows;
  mat->numBlockCols = numBlockCols;
  mat->rowInd = calloc(numBlocks ? numBlocks : 1, sizeof(BfSize));
  mat->colInd = calloc(numBlocks ? numBlocks : 1, sizeof(BfSize));
  if (!mat->rowInd || !mat->colInd) { bfSetError(BF_ERROR_MEMORY_ERROR); return; }
  block_init(&mat->super, &cooVtable, numBlocks, numBlockRows, numBlockCols, &mat->ownsBlock);
}

void bfMatBlockCooDeinit(BfMatBlockCoo *mat) {
  if (!mat) return;
  block_deinit(&mat->super, mat->numBlocks, mat->ownsBlock, 1);
  mat->ownsBlock = NULL;
  free(mat->rowInd);
  free(mat->colInd);
  mat->rowInd = mat->colInd = NULL;
}

void bfMatBlockCooDealloc(BfMatBlockCoo **mat) {
  if (!mat) return;
  free(*mat);
  *mat = NULL;
}

void bfMatBlockCooDeinitAndDealloc(BfMatBlockCoo **mat) {
  if (!mat || !*mat) return;
  bfMatBlockCooDeinit(*mat);
  bfMatBlockCooDealloc(mat);
}

BfMat *bfMatBlockCooToMat(BfMatBlockCoo *mat) {
  if (!mat) return NULL;
  block_shape(&mat->super, mat->numBlockRows, mat->numBlockCols);
  return &mat->super.super;
}

BfMatBlockCoo *bfMatToMatBlockCoo(BfMat *mat) {
  if (!mat || bfMatGetType(mat) != BF_TYPE_MAT_BLOCK_COO) { bfSetError(BF_ERROR_TYPE_ERROR); return NULL; }
  return (BfMatBlockCoo *)mat;
}

/* block-dense */

static BfType dense_get_type(BfMat const *m) { (void)m; return BF_TYPE_MAT_BLOCK_DENSE; }
static BfSize dense_num_bytes(BfMat const *m) {
  return blocks_bytes((BfMatBlock const *)m, num_blocks(m), sizeof(BfMatBlockDense));
}
static void dense_delete(BfMat **m) {
  BfMatBlockDense *e = (BfMatBlockDense *)*m;
  bfMatBlockDenseDeinitAndDealloc(&e);
  *m = NULL;
}
static BfMatVtable denseVtable = {dense_get_type, dense_num_bytes, dense_delete, NULL, packed_mul, NULL};

BfMatBlockDense *bfMatBlockDenseNew(void) {
  BfMatBlockDense *e = calloc(1, sizeof(BfMatBlockDense));
  if (!e) bfSetError(BF_ERROR_MEMORY_ERROR);
  return e;
}

void bfMatBlockDenseInit(BfMatBlockDense *mat, BfSize numBlockRows, BfSize numBlockCols) {
  if (!mat) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  mat->numBlockRows = numBlockRows;
  mat->numBlockCols = numBlockCols;
  block_init(&mat->super, &denseVtable, numBlockRows * numBlockCols, numBlockRows, numBlockCols, &mat->ownsBlock);
}

BfMatBlockDense *bfMatBlockDenseNewFromBlocks(BfSize numRowBlocks, BfSize numColBlocks, BfPtrArray *blocks,
                                              BfPolicy policy) {
  const BfSize nb = numRowBlocks * numColBlocks;
  if (bfPtrArraySize(blocks) != nb) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  BfMatBlockDense *e = bfMatBlockDenseNew();
  if (!e) return NULL;
  bfMatBlockDenseInit(e, numRowBlocks, numColBlocks);
  for (BfSize b = 0; b < nb; ++b) {
    BfMat *m = (BfMat *)blocks->data[b];
    e->super.block[b] = m ? bfMatGet(m, policy) : NULL;
    e->ownsBlock[b] = m && policy != BF_POLICY_VIEW;
  }
  offsets_from_blocks(&e->super, nb, numRowBlocks, numColBlocks, &e->super.super);
  return e;
}

void bfMatBlockDenseDeinit(BfMatBlockDense *mat) {
  if (!mat) return;
  block_deinit(&mat->super, mat->numBlockRows * mat->numBlockCols, mat->ownsBlock, 0);
  mat->ownsBlock = NULL;
}

void bfMatBlockDenseDealloc(BfMatBlockDense **mat) {
  if (!mat) return;
  free(*mat);
  *mat = NULL;
}

void bfMatBlockDenseDeinitAndDealloc(BfMatBlockDense **mat) {
  if (!mat || !*mat) return;
  bfMatBlockDenseDeinit(*mat);
  bfMatBlockDenseDealloc(mat);
}

BfMat *bfMatBlockDenseGetBlock(BfMatBlockDense *mat, BfSize i, BfSize j) {
  if (!mat || i >= mat->numBlockRows || j >= mat->numBlockCols) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  return mat->super.block[i * mat->numBlockCols + j];
}

void bfMatBlockDenseSetBlock(BfMatBlockDense *mat, BfSize i, BfSize j, BfMat *block) {
  if (!mat || i >= mat->numBlockRows || j >= mat->numBlockCols) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  const BfSize b = i * mat->numBlockCols + j;
  if (mat->super.block[b] && mat->ownsBlock[b] && mat->super.block[b] != block) bfMatDelete(&mat->super.block[b]);
  mat->super.block[b] = block;
  mat->ownsBlock[b] = true;
  offsets_from_blocks(&mat->super, mat->numBlockRows * mat->numBlockCols, mat->numBlockRows, mat->numBlockCols,
                      &mat->super.super);
}

BfMat *bfMatBlockDenseToMat(BfMatBlockDense *mat) { return mat ? &mat->super.super : NULL; }

BfMatBlockDense *bfMatToMatBlockDense(BfMat *mat) {
  if (!mat || bfMatGetType(mat) != BF_TYPE_MAT_BLOCK_DENSE) { bfSetError(BF_ERROR_TYPE_ERROR); return NULL; }
  return (BfMatBlockDense *)mat;
}

/* A block-COO matrix's shape follows the offsets its caller fills after
 * bfMatBlockCooInit; the dispatchers read it through here. */
void bfk_mat_refresh_shape(BfMat const *m) {
  if (!m) return;
  if (m->vtbl == &cooVtable) {
    BfMatBlockCoo *o = (BfMatBlockCoo *)m;
    if (o->super.rowOffset && o->super.colOffset) block_shape(&o->super, o->numBlockRows, o->numBlockCols);
  } else if (m->vtbl == &prodVtable) { /* rows of the first factor, columns of the last */
    BfMatProduct *p = (BfMatProduct *)m;
    const BfSize n = p->factorArr.num_elts;
    p->super.numRows = n ? bfMatGetNumRows((BfMat const *)p->factorArr.data[0]) : 0;
    p->super.numCols = n ? bfMatGetNumCols((BfMat const *)p->factorArr.data[n - 1]) : 0;
  }
}

/* -------------------------------------------------------- dense views */

BfMatDenseComplex *bfMatDenseComplexNewViewFromPtr(BfSize numRows, BfSize numCols, BfComplex *data) {
  BfMatDenseComplex *m = bfMatDenseComplexNew();
  if (!m) return NULL;
  bfMatDenseComplexInitViewFromPtr(m, numRows, numCols, data);
  return m;
}

/* ------------------------------------------------------ complex vectors */

/* marks a BfVec as a BfVecComplex (the reference dispatches through a
 * vector vtable; only complex vectors exist here) */
static int vecComplexTag;
#define VEC_VIEW 1

BfVecComplex *bfVecComplexNew(void) {
  BfVecComplex *v = calloc(1, sizeof(BfVecComplex));
  if (!v) bfSetError(BF_ERROR_MEMORY_ERROR);
  return v;
}

void bfVecComplexInit(BfVecComplex *vec, BfSize size) {
  if (!vec) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  vec->super.vtbl = &vecComplexTag;
  vec->super.props = 0;
  vec->super.size = size;
  vec->stride = 1;
  vec->data = calloc(size ? size : 1, sizeof(BfComplex));
  if (!vec->data) bfSetError(BF_ERROR_MEMORY_ERROR);
}

void bfVecComplexInitView(BfVecComplex *vec, BfSize size, BfSize stride, BfComplex *data) {
  if (!vec) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  vec->super.vtbl = &vecComplexTag;
  vec->super.props = VEC_VIEW;
  vec->super.size = size;
  vec->stride = stride;
  vec->data = data;
}

void bfVecComplexDeinit(BfVecComplex *vec) {
  if (!vec) return;
  if (!(vec->super.props & VEC_VIEW)) free(vec->data);
  vec->data = NULL;
}

void bfVecComplexDeinitAndDealloc(BfVecComplex **vec) {
  if (!vec || !*vec) return;
  bfVecComplexDeinit(*vec);
  free(*vec);
  *vec = NULL;
}

BfVec *bfVecComplexToVec(BfVecComplex *vec) { return vec ? &vec->super : NULL; }

BfVecComplex *bfVecToVecComplex(BfVec *vec) {
  if (!vec || vec->vtbl != &vecComplexTag) { bfSetError(BF_ERROR_TYPE_ERROR); return NULL; }
  return (BfVecComplex *)vec;
}

void bfVecDelete(BfVec **vec) {
  if (!vec || !*vec) return;
  if ((*vec)->vtbl != &vecComplexTag) { bfSetError(BF_ERROR_TYPE_ERROR); return; }
  BfVecComplex *v = (BfVecComplex *)*vec;
  bfVecComplexDeinitAndDealloc(&v);
  *vec = NULL;
}

BfVec *bfMatMulVec(BfMat const *mat, BfVec const *vec) {
  if (!mat || !vec || vec->vtbl != &vecComplexTag) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return NULL; }
  const BfVecComplex *x = (const BfVecComplex *)vec;
  BfMatDenseComplex *X = bfMatDenseComplexNew();
  if (!X) return NULL;
  bfMatDenseComplexInit(X, x->super.size, 1);
  if (!X->data) { free(X); return NULL; }
  for (BfSize i = 0; i < x->super.size; ++i) X->data[i] = x->data[i * x->stride];
  BfMat *Y = bfMatMul(mat, &X->super);
  bfMatDenseComplexDeinitAndDealloc(&X);
  if (!Y) return NULL;
  if (bfMatGetType(Y) != BF_TYPE_MAT_DENSE_COMPLEX || Y->numCols != 1) {
    bfMatDelete(&Y);
    bfSetError(BF_ERROR_TYPE_ERROR);
    return NULL;
  }
  BfMatDenseComplex *Yd = (BfMatDenseComplex *)Y;
  BfVecComplex *y = bfVecComplexNew();
  if (!y) { bfMatDelete(&Y); return NULL; }
  bfVecComplexInit(y, Y->numRows);
  if (!y->data) { free(y); bfMatDelete(&Y); return NULL; }
  for (BfSize i = 0; i < Y->numRows; ++i) y->data[i] = Yd->data[i * Yd->rowStride];
  bfMatDelete(&Y);
  return &y->super;
}
