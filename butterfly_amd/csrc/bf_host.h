/* Internal declarations of the host library (C11). */
#pragma once

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/bf_amd/bf.h"
#include "bf_plan.h"

/* ---- launchers in bf_kernels.hip */
int bfk_init(void);
int bfk_device_count(void);
int bfk_scratch(void **ptr, size_t *bytes);
size_t bfk_lds_cap(void);
void *bfk_stream_handle(void);
/* the quadtree on the GPU (bf_tree.hip) */
int bfk_quadtree(int64_t n, const double *xy, int64_t leafSize,
                 void (*rootBox)(const double mn[2], const double mx[2], double outMin[2], double outMax[2]),
                 BfQuadtreeNode **nodesOut, int64_t *numNodesOut, int64_t *levelCount, int *numLevelsOut,
                 BfSize *perm);
/* rank model and randomized range finder (bf_rank.hip) */
int bfk_mlp_nparam(void);
int bfk_mlp_groups(int B);
int bfk_mlp_step(float *dP, float *dM, float *dV, const float *dX, const float *dY, const int32_t *dIdx, int64_t b0,
                 int B, float lr, float c1, float c2, float *dPart, float *dLossPart, double *dLossAcc);
int bfk_mlp_grad(const float *dP, const float *dX, const float *dY, const int32_t *dIdx, int B, float *dPart,
                 float *dLossPart);
int bfk_mlp_predict(const float *dP, const float *dX, int64_t n, float *dOut);
int bfk_sketch_jobs_bytes(void);
int bfk_sketch_resid_chunks(int64_t maxEntries);
int bfk_sketch_range(int64_t batch, const void *dJobs, uint64_t seed, const double *dA, double *dOm, double *dY,
                     int64_t maxMW);
int bfk_sketch_project(int64_t batch, const void *dJobs, const int64_t *dSOff, const double *dSig, const double *dA,
                       double *dY, double *dBH, double *dPart, int64_t maxMW, int64_t maxNW, int64_t maxMN);
size_t bfk_lstsq_lds_bytes(int p, int n);
void *bfk_malloc(size_t bytes);
void bfk_free(void *p);
void bfk_free_async(void *p);
int bfk_h2d(void *dst, const void *src, size_t bytes);
int bfk_h2d_side(void *dst, const void *src, size_t bytes);
int bfk_d2h(void *dst, const void *src, size_t bytes);
int bfk_memset(void *dst, size_t bytes);
int bfk_sync(void);
void bfk_get_times(double out[8]);
void bfk_get_time_sums(double sums[8], double counts[8]);
void bfk_reset_times(void);
int bfk_set_apply_config(int wavesPerBlock, int unroll, int mfmaTiles);
int bfk_set_compensation(int on);
int bfk_sample(const BfkKernelJob *dJobs, int64_t numJobs, const double *dSrcXY, const double *dTgtXY,
               const double *dSrcN, const double *dTgtN, BfkLayer layer, double *dArena);
int bfk_lstsq_begin(size_t maxOversizeBytes, int64_t numOversize);
int bfk_lstsq_launch(const BfkLstsqJob *dJobs, const int32_t *dOrder, int64_t numJobs, size_t ldsBytes,
                     const double *dSrcXY, const double *dSrcN, BfkLayer layer, double *dArena);
int bfk_lstsq_end(void);
void bfk_set_lstsq_waves(int64_t two, int64_t four, int64_t *old2, int64_t *old4);
int bfk_pair_green(int64_t n, const double *dA, const double *dB, const double *dNA, double k, double *dOut);
int bfk_scale_rows(int64_t n, int nrhs, const double *dW, const void *dX, void *dU, int fp32);
int bfk_bie_post(int64_t n, int nrhs, int ncorr, double alpha, const int32_t *dCol, const double *dC,
                 const void *dX, const void *dU, void *dY, int fp32);
int bfk_dotc_multi(int64_t n, int nv, const double *dV, int64_t ldv, const double *dW, double *hOut);
int bfk_axpy_multi(int64_t n, int nv, const double *dV, int64_t ldv, const double *hH, double s, double *dW);
int bfk_convert(int64_t n, const void *src, void *dst, int to32);
int bfk_svd_jacobi(int64_t batch, const int64_t *dAOff, const int32_t *dM, const int32_t *dN, double *dA,
                   const int64_t *dVOff, double *dV, const int64_t *dSOff, double *dSig, int maxSweeps);
int bfk_svd_jacobi_big(int64_t nb, const int32_t *dBig, const int64_t *dAOff, const int32_t *dM, const int32_t *dN,
                       double *dA, const int64_t *dVOff, double *dV, const int64_t *dSOff, double *dSig,
                       int maxSweeps, int maxN);
int bfk_svd_tall_max_n(void);
int bfk_svd_jacobi_tall(int64_t nb, const int32_t *dIdx, const int64_t *dAOff, const int32_t *dM, const int32_t *dN,
                        double *dA, double *dS, const int64_t *dVOff, double *dV, const int64_t *dSOff, double *dSig,
                        int maxSweeps, int maxN);
size_t bfk_ortho_units(int rows, int n);
int bfk_ortho(const BfkOrthoJob *dJobs, const int32_t *dOrder, int numGroups, const int64_t *groupStart,
              const size_t *groupBytes, const BfkOrthoItem *dItems, double *dArena, const double *dRin,
              double *dRout);
int bfk_apply(const BfkSeg *dSegs, const int64_t *stageSeg, int numStages, const BfkTerm *dTerms,
              const double *dArena, const double *dX, double *dPing, double *dPong, double *dY, int nrhs,
              const int32_t *stageCfg);
int bfk_apply_f32(const BfkSeg *dSegs, const int64_t *stageSeg, int numStages, const BfkTerm *dTerms,
                  const float *dArenaHi, const float *dArenaLo, const float *dX, float *dPing, float *dPong,
                  float *dY, int nrhs, const int32_t *stageCfg);
int bfk_apply_stage_ms(const BfkSeg *dSegs, const int64_t *stageSeg, int s, const BfkTerm *dTerms,
                       const double *dArena, const double *dX, double *dPing, double *dPong, double *dY, int nrhs,
                       int cfg, int reps, double *ms);
int bfk_apply_f32_stage_ms(const BfkSeg *dSegs, const int64_t *stageSeg, int s, const BfkTerm *dTerms,
                           const float *dArenaHi, const float *dArenaLo, const float *dX, float *dPing, float *dPong,
                           float *dY, int nrhs, int cfg, int reps, double *ms);
int bfk_graph_begin(void);
int bfk_graph_end(void **exec);
int bfk_graph_launch(void *exec, int slot);
void bfk_graph_destroy(void *exec);
int bfk_apply_config_stamp(void);
int bfk_to_f32(const double *dArena, int64_t numEntries, float *dHi, float *dLo);
int bfk_h0(int64_t n, const double *x, double *out);
int bfk_copy(const void *src, void *dst, size_t bytes);
int bfk_lstsq_loaded(int64_t batch, int p, int n, const double *A, const double *B, double *X);

/* ---- growable arrays */
#define BF_VEC(T) struct { T *a; size_t n, cap; }
#define BF_VEC_PUSH(v, x)                                                        \
  do {                                                                           \
    if ((v).n == (v).cap) {                                                      \
      size_t nc_ = (v).cap ? 2 * (v).cap : 64;                                   \
      void *na_ = realloc((v).a, nc_ * sizeof(*(v).a));                          \
      if (!na_) abort();                                                         \
      (v).a = na_;                                                               \
      (v).cap = nc_;                                                             \
    }                                                                            \
    (v).a[(v).n++] = (x);                                                        \
  } while (0)
#define BF_VEC_FREE(v) (free((v).a), (v).a = NULL, (v).n = (v).cap = 0)

/* ---- the device factorization */

enum { BF_BLOCK_LSTSQ = 0, BF_BLOCK_KERNEL = 1, BF_BLOCK_IDENTITY = 2 };

typedef struct {
  int64_t dataOff;   /* complex entries into the arena, -1 for identity */
  int32_t rowInd, colInd, m, n;
  int32_t kind;
  int32_t job;       /* index into lstsqJobs or kernelJobs, -1 for identity */
} BfHipBlock;

typedef struct {
  int32_t isDiag;
  int32_t pad;
  int64_t numBlockRows, numBlockCols, numBlocks;
  int64_t firstBlock;          /* into blocks */
  int64_t rowOffStart;         /* into offs: numBlockRows + 1 entries */
  int64_t colOffStart;         /* into offs: numBlockCols + 1 entries */
} BfHipFactor;

typedef struct {
  int32_t isButterfly;
  int32_t numFactors;
  int64_t row0, col0, m, n;
  int64_t firstFactor;         /* into factors */
  int64_t denseOff;            /* arena offset of a dense leaf */
} BfHipLeaf;

/* right-hand-side classes of the tuned launch shapes: 1, 2-15, 16-31,
 * 32-63, 64-127, 128-255, 256 and more */
#define BF_NRHS_CLASSES 7
static inline int bf_nrhs_class(size_t nrhs) {
  if (nrhs <= 1) return 0;
  if (nrhs < 16) return 1;
  int c = 2;
  for (size_t v = 32; c < BF_NRHS_CLASSES - 1 && nrhs >= v; v *= 2) ++c;
  return c;
}

typedef BF_VEC(BfHipLeaf) BfHipLeafVec;
typedef BF_VEC(BfHipFactor) BfHipFactorVec;
typedef BF_VEC(BfHipBlock) BfHipBlockVec;
typedef BF_VEC(int64_t) BfI64Vec;
typedef BF_VEC(int32_t) BfI32Vec;
typedef BF_VEC(BfkLstsqJob) BfLstsqJobVec;
typedef BF_VEC(BfkKernelJob) BfKernelJobVec;
typedef BF_VEC(BfkSeg) BfSegVec;
typedef BF_VEC(BfkTerm) BfTermVec;

typedef struct BfMatHipButterfly {
  BfMat super;
  BfHelm2 helm;
  BfFacHelm2Options opts;
  BfSize numSrc, numTgt;

  /* plan */
  BfHipLeafVec leaves;
  BfHipFactorVec factors;
  BfHipBlockVec blocks;
  BfI64Vec offs;
  BfLstsqJobVec lstsqJobs;
  BfKernelJobVec kernelJobs;
  int64_t arenaEntries;

  /* apply plan */
  BfSegVec segs;
  BfTermVec terms;
  BfI64Vec stageSeg;
  int numStages;
  int64_t bufRows[3];  /* rows of ping, pong (per right-hand side) */
  double applyFactorBytes;

  /* device */
  double *dArena;
  float *dArena32, *dArena32lo;
  double *dSrcXY, *dTgtXY;
  double *dSrcN; /* source unit normals in tree order (double layer, combined field) */
  double *dTgtN; /* target unit normals in tree order (S') */
  BfkSeg *dSegs;
  BfkTerm *dTerms;
  void *dPing, *dPong;
  size_t bufBytes;

  /* launch shape per stage chosen by the tuner, per precision and
   * right-hand-side class (NULL: the global settings) */
  int32_t *tuneCfg[3][BF_NRHS_CLASSES];
  /* the apply chain captured as a graph, and what it was captured for */
  void *graphExec;
  const void *graphX;
  void *graphY, *graphPing, *graphPong;
  size_t graphNrhs;
  int graphPrec, graphStamp;
  const int32_t *graphCfg;

  double stats[16];
  int statsCounted;    /* stats 4-8 and 10 are filled in */
  int orthogonal;      /* the factors have been orthogonalized */
  double orthoSeconds; /* wall time of the orthogonalization */
} BfMatHipButterfly;

/* planner (plan.c) */
int bfPlanBuild(BfMatHipButterfly *F, BfQuadtree const *src, BfQuadtree const *tgt);
/* plans one more problem straight into F as its next diagonal block */
int bfPlanBuildAt(BfMatHipButterfly *F, BfQuadtree const *src, BfQuadtree const *tgt, int64_t srcPts,
                  int64_t tgtPts);
/* one butterfly on given source / target levels (node indices) */
int bfPlanButterfly(BfMatHipButterfly *F, BfQuadtree const *src, BfQuadtree const *tgt, const int32_t *const *ls,
                    const size_t *nls, const int32_t *const *lt, const size_t *nlt, int D);
/* apply plan (fac.c) */
int bfApplyPlanBuild(BfMatHipButterfly *F);
/* apply plan and uploads for a plan + arena already in place (fac.c) */
int bfMatHipButterflyFinish(BfMatHipButterfly *F);

/* matrix vtable (mat.c) */
struct BfMatVtable {
  BfType (*GetType)(BfMat const *);
  BfSize (*NumBytes)(BfMat const *);
  void (*Delete)(BfMat **);
  BfMat *(*Copy)(BfMat const *);
  BfMat *(*Mul)(BfMat const *, BfMat const *);
  void (*PermuteRows)(BfMat *, BfPerm const *);
};
void bfMatInitBase(BfMat *mat, BfMatVtable *vtbl, BfSize numRows, BfSize numCols);
/* one block of a factor handed to bfk_product_from_blocks */
typedef struct {
  int64_t i0, j0, m, n;
  BfComplex *data; /* column-major m x n; NULL = identity (m == n) */
} BfkBlockRef;
BfMat *bfk_product_from_blocks(size_t nf, const int64_t *rows, const int64_t *cols, const size_t *nb,
                               const BfkBlockRef *const *blocks);
void bfk_mat_refresh_shape(BfMat const *m);
BfMat *bfMatHipButterflyMul(BfMat const *A, BfMat const *X);
void bfMatHipButterflyDelete(BfMat **mat);
BfSize bfMatHipButterflyNumBytes(BfMat const *mat);
extern BfMatVtable bfMatHipButterflyVtable;

/* timing */
double bfWallTime(void);
/* madvise(MADV_HUGEPAGE) over [p, p + bytes) for tables of 32 MiB and up */
void bfHugePages(void *p, size_t bytes);
/* prints "what" and the ms since `since` to stderr when BF_AMD_TRACE is set */
void bfTrace(const char *what, double since);
