/*
 * A C11 caller of the library, compiled against include/bf_amd/bf.h alone,
 * the way the reference's drivers call the reference library
 * (examples/simple/bf_all_blocks.c:96-147, bf_one_block.c:161-280):
 *
 *   1. the kernel of n seeded points in the unit square against themselves
 *      (single layer), factored with bfFacHelm2MakeMultilevel and applied
 *      with bfMatMul to a bfComplexRandn right-hand side in tree order, then
 *      compared with the dense product computed here from glibc's j0 / y0
 *      (the reference's H0, bessel.c:284-296);
 *   2. one separated pair of depth-2 nodes through bfFacHelm2MakeSingleLevel:
 *      the factors read through the containers' struct fields (types, block
 *      grids, rowInd / colInd, dense blocks), then the product applied to a
 *      vector on the source node's points and compared with the dense
 *      kernel block.
 *
 * Usage: bf_caller n  -- prints the two relative errors; exit status 0 when
 * both are below 1e-9 and every structural check holds.
 */
#define _DEFAULT_SOURCE 1
#include <bf_amd/bf.h>

#include <math.h>
#include <stdio.h>
#include <stdlib.h>

static int failures = 0;
#define CHECK(cond)                                                   \
  do {                                                                \
    if (!(cond)) {                                                    \
      fprintf(stderr, "bf_caller: check failed: %s (line %d)\n", #cond, __LINE__); \
      ++failures;                                                     \
    }                                                                 \
  } while (0)

static BfComplex kernel(double k, BfReal const *t, BfReal const *s) {
  const double r = hypot(t[0] - s[0], t[1] - s[1]);
  return r == 0 ? 0 : I * (j0(k * r) + I * y0(k * r)) / 4;
}

int main(int argc, char **argv) {
  const BfSize n = argc > 1 ? (BfSize)atol(argv[1]) : 2048;
  const double k = 20.0;
  bfSeed(0);
  BfPoints2 *points = bfPoints2NewWithCapacity(n);
  const BfReal lo[2] = {0, 0}, hi[2] = {1, 1};
  for (BfSize i = 0; i < n; ++i) {
    BfPoint2 p;
    bfPoint2SampleUniformlyFromBoundingBox(lo, hi, p);
    bfPoints2Append(points, p);
  }
  BfHelm2 helm = {.k = k, .layerPot = BF_LAYER_POTENTIAL_SINGLE};

  BfQuadtree quadtree;
  bfQuadtreeInit(&quadtree, points, NULL);
  CHECK(bfGetError() == BF_ERROR_NONE);
  BfTree *tree = bfQuadtreeToTree(&quadtree);
  BfPerm const *perm = bfTreeGetPermConst(tree);

  /* 1. the whole kernel matrix, in tree order on both sides */
  BfMat *A = bfFacHelm2MakeMultilevel(&helm, &quadtree, &quadtree);
  CHECK(A != NULL && bfMatGetType(A) == BF_TYPE_MAT_HIP_BUTTERFLY);
  BfMatDenseComplex *x = bfMatDenseComplexNewRandn(n, 1);
  BfMat *y = A ? bfMatMul(A, bfMatDenseComplexToMat(x)) : NULL;
  CHECK(y != NULL && bfMatGetNumRows(y) == n && bfMatGetNumCols(y) == 1);
  double num = 0, den = 0;
  if (y) {
    BfMatDenseComplex *Y = bfMatToMatDenseComplex(y);
    for (BfSize i = 0; i < n; i += 7) {
      BfComplex acc = 0;
      BfReal const *t = points->data[perm->index[i]];
      for (BfSize j = 0; j < n; ++j) acc += kernel(k, t, points->data[perm->index[j]]) * x->data[j];
      const BfComplex d = Y->data[i * Y->rowStride] - acc;
      num += creal(d) * creal(d) + cimag(d) * cimag(d);
      den += creal(acc) * creal(acc) + cimag(acc) * cimag(acc);
    }
  }
  const double errAll = sqrt(num / den);
  printf("multilevel: relative error vs dense rows %.3e\n", errAll);
  CHECK(errAll < 1e-9);

  /* 2. one separated pair of depth-2 nodes */
  BfQuadtreeNode *s = NULL, *t = NULL;
  for (BfSize a = quadtree.levelOffset[2]; a < quadtree.levelOffset[3] && !s; ++a)
    for (BfSize b = quadtree.levelOffset[2]; b < quadtree.levelOffset[3]; ++b)
      if (bfQuadtreeNodesAreSeparated(&quadtree.nodes[a], &quadtree.nodes[b])) {
        s = &quadtree.nodes[a];
        t = &quadtree.nodes[b];
        break;
      }
  CHECK(s && t && bfQuadtreeNodeGetQuadtreeConst(s) == &quadtree);
  double errOne = INFINITY;
  if (s && t) {
    BfMatProduct *F = bfMatToMatProduct(bfFacHelm2MakeSingleLevel(&helm, s, t));
    CHECK(F != NULL);
    const BfSize nf = F ? bfMatProductNumFactors(F) : 0;
    CHECK(nf >= 2);
    for (BfSize i = 0; i < nf; ++i) {
      BfMat *f = bfMatProductGetFactor(F, i);
      const BfType type = bfMatGetType(f);
      /* the evaluation and first factors are block diagonal, the others COO */
      CHECK(type == ((i == 0 || i == nf - 1) ? BF_TYPE_MAT_BLOCK_DIAG : BF_TYPE_MAT_BLOCK_COO));
      BfMatBlock *B = (BfMatBlock *)f;
      BfSize nb, nbr, nbc;
      if (type == BF_TYPE_MAT_BLOCK_COO) {
        BfMatBlockCoo *c = (BfMatBlockCoo *)f;
        nb = c->numBlocks, nbr = c->numBlockRows, nbc = c->numBlockCols;
        for (BfSize j = 0; j < nb; ++j) CHECK(c->rowInd[j] < nbr && c->colInd[j] < nbc);
      } else {
        BfMatBlockDiag *d = (BfMatBlockDiag *)f;
        nbr = d->numBlockRows, nbc = d->numBlockCols, nb = nbr < nbc ? nbr : nbc;
      }
      CHECK(B->rowOffset[0] == 0 && B->rowOffset[nbr] == bfMatGetNumRows(f));
      CHECK(B->colOffset[0] == 0 && B->colOffset[nbc] == bfMatGetNumCols(f));
      for (BfSize j = 0; j < nb; ++j) CHECK(bfMatGetType(B->block[j]) == BF_TYPE_MAT_DENSE_COMPLEX);
    }
    const BfSize ms = bfQuadtreeNodeGetNumPoints(t), ns = bfQuadtreeNodeGetNumPoints(s);
    CHECK(F && bfMatGetNumRows(bfMatProductToMat(F)) == ms && bfMatGetNumCols(bfMatProductToMat(F)) == ns);
    BfMatDenseComplex *q = bfMatDenseComplexNewRandn(ns, 1);
    BfMat *v = F ? bfMatMul(bfMatProductToMat(F), bfMatDenseComplexToMat(q)) : NULL;
    CHECK(v != NULL);
    if (v) {
      BfMatDenseComplex *V = bfMatToMatDenseComplex(v);
      num = den = 0;
      for (BfSize i = 0; i < ms; ++i) {
        BfComplex acc = 0;
        BfReal const *tp = points->data[perm->index[t->offset[0] + i]];
        for (BfSize j = 0; j < ns; ++j) acc += kernel(k, tp, points->data[perm->index[s->offset[0] + j]]) * q->data[j];
        const BfComplex d = V->data[i * V->rowStride] - acc;
        num += creal(d) * creal(d) + cimag(d) * cimag(d);
        den += creal(acc) * creal(acc) + cimag(acc) * cimag(acc);
      }
      errOne = sqrt(num / den);
      bfMatDelete(&v);
    }
    printf("single level: %lu factors, relative error vs the dense block %.3e\n", (unsigned long)nf, errOne);
    CHECK(errOne < 1e-9);
    bfMatDenseComplexDeinitAndDealloc(&q);
    BfMat *Fm = bfMatProductToMat(F);
    bfMatDelete(&Fm);
  }

  bfMatDelete(&y);
  bfMatDelete(&A);
  bfMatDenseComplexDeinitAndDealloc(&x);
  bfQuadtreeDeinit(&quadtree);
  bfPoints2DeinitAndDealloc(&points);
  CHECK(bfGetError() == BF_ERROR_NONE);
  printf("bf_caller: %s\n", failures ? "FAILED" : "ok");
  return failures ? 1 : 0;
}
