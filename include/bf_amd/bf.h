/*
 * butterfly_amd -- the C ABI of the MI355X 2-D Helmholtz butterfly
 * factor-and-apply path.
 *
 * Every entry point below keeps the name, argument meaning, ownership rule
 * and error behaviour of the reference function it replaces (cited
 * path:line, relative to the reference repository).  Entry points prefixed
 * `bfhip` have no reference counterpart; each says who it is for.
 *
 * Conventions (as in the reference):
 *   - `New` allocates, `Init` fills caller storage, `Deinit` releases what
 *     `Init` made, `Dealloc` frees the struct, `Delete` does both
 *     (docs/implementation.rst:27-48).
 *   - Errors: functions return NULL / 0 / NaN and record a code that
 *     bfGetError() returns and clears (error.c:7-24).  Unlike the
 *     reference's debug builds, a pending error never traps; HIP failures
 *     map to BF_ERROR_RUNTIME_ERROR or BF_ERROR_MEMORY_ERROR.
 *   - Complex numbers are interleaved FP64 (C99 double _Complex, def.h:27).
 *     Dense matrices are row-major with explicit strides.
 *   - A factorization maps X in source-tree order to Y in target-tree
 *     order; callers permute with the tree permutations (helm2_bie.c:73-81).
 *   - Single-threaded host API: the error code, the RNG state and the
 *     library's HIP stream are process-global.
 */
#pragma once

#include <complex.h>
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>
#include <stdio.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---------------------------------------------------------------- types */

/* def.h:23-41 */
typedef double BfReal;
typedef double _Complex BfComplex;
typedef size_t BfSize;
typedef BfReal BfPoint2[2];
typedef BfReal BfVector2[2];
#define BF_EPS_MACH 2.220446049250313e-16
#define BF_SIZE_BAD_VALUE ((BfSize)-1)

/* error.h:3-17 */
enum BfError {
  BF_ERROR_NONE,
  BF_ERROR_INVALID_ARGUMENTS,
  BF_ERROR_RUNTIME_ERROR,
  BF_ERROR_NOT_IMPLEMENTED,
  BF_ERROR_MEMORY_ERROR,
  BF_ERROR_OUT_OF_RANGE,
  BF_ERROR_FILE_ERROR,
  BF_ERROR_TYPE_ERROR,
  BF_ERROR_INCOMPATIBLE_SHAPES,
};

/* types.h:78-144 with the reference's values (the types this path uses,
 * BF_TYPE_COUNT and BF_TYPE_NONE); the device types come after
 * BF_TYPE_NONE, so no reference value changes meaning */
typedef enum BfTypes {
  BF_TYPE_MAT = 0,
  BF_TYPE_MAT_IDENTITY = 8,
  BF_TYPE_MAT_PRODUCT = 10,
  BF_TYPE_MAT_BLOCK_COO = 15,
  BF_TYPE_MAT_BLOCK_DENSE = 16,
  BF_TYPE_MAT_BLOCK_DIAG = 17,
  BF_TYPE_MAT_DENSE_COMPLEX = 19,
  BF_TYPE_QUADTREE = 34,
  BF_TYPE_QUADTREE_NODE = 39,
  BF_TYPE_COUNT = 42,
  BF_TYPE_NONE = 43,
  BF_TYPE_MAT_HIP_BUTTERFLY = 44, /* a Helmholtz butterfly factorization on the device */
  BF_TYPE_MAT_HIP_PRODUCT = 45,   /* a streamed factorization on the device */
  BF_TYPE_MAT_HIP_BIE = 46,       /* a boundary integral operator on the device */
} BfType;

/* layer_pot.h:28-69 */
typedef enum BfLayerPotentials {
  BF_LAYER_POTENTIAL_UNKNOWN = 0,
  BF_LAYER_POTENTIAL_SINGLE,
  BF_LAYER_POTENTIAL_PV_DOUBLE,
  BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_SINGLE,
  BF_LAYER_POTENTIAL_PV_NORMAL_DERIV_DOUBLE,
  BF_LAYER_POTENTIAL_COMBINED_FIELD,
  BF_LAYER_POTENTIAL_COUNT,
} BfLayerPotential;

/* helm2.h:10-15 */
typedef struct BfHelm2 {
  BfReal k;
  BfLayerPotential layerPot;
  BfComplex alpha;
  BfComplex beta;
} BfHelm2;

/* circle.h:5-8 */
typedef struct BfCircle {
  BfReal r;
  BfPoint2 center;
} BfCircle;

/* points.h:39-44, vectors.h:19-24 */
typedef struct BfPoints2 {
  BfPoint2 *data;
  BfSize size;
  BfSize capacity;
  bool isView;
} BfPoints2;

typedef struct BfVectors2 {
  BfVector2 *data;
  BfSize size;
  BfSize capacity;
  bool isView;
} BfVectors2;

/* perm.h:6-10 */
typedef struct BfPerm {
  BfSize *index;
  BfSize size;
  bool isView;
} BfPerm;

/* A quadtree node.  The reference links nodes by pointers
 * (quadtree_node.h, tree_node.h); here nodes live in one array in
 * breadth-first order, so `child` and `parent` are indices into it
 * (-1 = none).  `offset[q]..offset[q+1]` is child q's range of the
 * tree permutation, as in tree_node.h. */
typedef struct BfQuadtreeNode {
  BfReal bboxMin[2], bboxMax[2];
  BfReal split[2];
  BfSize offset[5];
  int32_t child[4];
  int32_t parent;
  int32_t depth;
  int32_t maxDepthBelow; /* depth of the deepest leaf in this subtree */
  int32_t index;         /* position among the parent's children (0..3) */
} BfQuadtreeNode;

/* tree_node.h: the tree-generic node is the quadtree node here (the only
 * tree this path builds), so BfTreeNode pointers are BfQuadtreeNode
 * pointers */
typedef BfQuadtreeNode BfTreeNode;

/* tree.h:30-39 */
typedef struct BfTree {
  void *vtable;          /* unused; kept so `perm` sits where callers expect */
  BfQuadtreeNode *root;
  BfPerm *perm;
} BfTree;

/* quadtree.h:14-24 */
typedef struct BfQuadtree {
  BfTree super;
  BfPoints2 const *points;
  BfVectors2 const *unitNormals;
  BfQuadtreeNode *nodes; /* breadth-first order, nodes[0] is the root */
  BfSize numNodes;
  BfSize leafSize;       /* a node with <= leafSize points is a leaf */
  BfSize *levelOffset;   /* nodes of depth d are [levelOffset[d], levelOffset[d+1]) */
  BfSize numLevels;
} BfQuadtree;

/* --------------------------------------------------------- matrix types */

typedef struct BfMat BfMat;
typedef struct BfMatVtable BfMatVtable;

/* mat.h:30-39 */
typedef enum BfMatProps {
  BF_MAT_PROPS_NONE = 0,
  BF_MAT_PROPS_VIEW = 1 << 0,
  BF_MAT_PROPS_TRANS = 1 << 1,
  BF_MAT_PROPS_CONJ = 1 << 2,
  BF_MAT_PROPS_ORTHO = 1 << 3,
} BfMatProps;

/* mat.h:187-196: the base every matrix type embeds first */
struct BfMat {
  BfMatVtable *vtbl;
  BfMatProps props;
  BfSize numRows;
  BfSize numCols;
};

/* mat_dense_complex.h:69-80 */
typedef struct BfMatDenseComplex {
  BfMat super;
  BfSize rowStride;
  BfSize colStride;
  BfComplex *data;
  void *pyArray;
} BfMatDenseComplex;

/* ----------------------------------------------------------- functions */

/* error.c:10-24 (no trap on a pending error) */
enum BfError bfGetError(void);
void bfSetError(enum BfError error);

/* rand.c:19-76 */
void bfSeed(BfSize seed);
BfReal bfRealUniform1(void);
void bfRealUniform(BfSize n, BfReal *x);
void bfRealRandn(BfSize n, BfReal *x);
void bfComplexRandn(BfSize n, BfComplex *x);
/* points.c:25-28 */
void bfPoint2SampleUniformlyFromBoundingBox(BfReal const bboxMin[2], BfReal const bboxMax[2], BfPoint2 p);

/* points.c: construction and release of point sets */
BfPoints2 *bfPoints2NewWithCapacity(BfSize capacity);
void bfPoints2Append(BfPoints2 *points, BfPoint2 const p);
void bfPoints2DeinitAndDealloc(BfPoints2 **points);
/* new: a view over caller-owned n x 2 coordinates (no counterpart; for
 * FFI callers that hold coordinates in their own arrays) */
BfPoints2 *bfPoints2NewViewFromPtr(BfReal *xy, BfSize n);

/* perm.c:195-220 */
BfPerm *bfPermGetReversePerm(BfPerm const *perm);
void bfPermDelete(BfPerm **perm);

/* quadtree.c:80-98, tree.c:110; tree.h accessors */
BfQuadtree *bfQuadtreeNew(void);
void bfQuadtreeInit(BfQuadtree *tree, BfPoints2 const *points, BfVectors2 const *unitNormals);
/* new: the same with an explicit leaf threshold (reference: fixed at 1,
 * quadtree_node.c:17) */
void bfQuadtreeInitWithLeafSize(BfQuadtree *tree, BfPoints2 const *points,
                                BfVectors2 const *unitNormals, BfSize leafSize);
/* new: the same tree built on the GPU (replaces the host recursion of
 * quadtree_node.c:123-306 with Morton keys, a radix sort and level
 * tables).  Nodes, boxes, splits, point ranges and the breadth-first
 * numbering equal the host builder's; at the reference's leaf size 1 the
 * permutation equals it too, and with larger leaves the points inside a
 * leaf are in Morton order.  BF_ERROR_OUT_OF_RANGE when the tree would be
 * 31 levels deep (points closer than 2^-30 of the root box's side). */
void bfhipQuadtreeInitWithLeafSize(BfQuadtree *tree, BfPoints2 const *points,
                                   BfVectors2 const *unitNormals, BfSize leafSize);
void bfQuadtreeDeinit(BfQuadtree *tree);
void bfQuadtreeDealloc(BfQuadtree **tree);
void bfQuadtreeDeinitAndDealloc(BfQuadtree **tree);
BfTree *bfQuadtreeToTree(BfQuadtree *tree);
BfPerm const *bfTreeGetPermConst(BfTree const *tree);
/* quadtree_node.h:52: the tree a node belongs to (NULL and
 * BF_ERROR_INVALID_ARGUMENTS for a node of no live tree) */
BfQuadtree const *bfQuadtreeNodeGetQuadtreeConst(BfQuadtreeNode const *node);
/* quadtree_node.h: the node as a tree node (the same pointer here) */
BfTreeNode *bfQuadtreeNodeToTreeNode(BfQuadtreeNode *node);
/* quadtree_node.c:321-330, :393-401 */
BfCircle bfQuadtreeNodeGetBoundingCircle(BfQuadtreeNode const *node);
bool bfQuadtreeNodesAreSeparated(BfQuadtreeNode const *node1, BfQuadtreeNode const *node2);
BfSize bfQuadtreeNodeGetNumPoints(BfQuadtreeNode const *node);

/* helm2.c:14-36 */
BfSize bfHelm2RankEstForTwoCircles(BfHelm2 const *helm, BfCircle const *circ1,
                                   BfCircle const *circ2, BfReal C, BfReal eps);
/* helm2.c:282-319: dense kernel matrix, rows = targets, cols = sources, for
 * the single layer, the double layer (PV_DOUBLE, source normals Nsrc), the
 * combined field alpha S + beta D (helm2.c:38-90) and the normal derivative
 * of the single layer S' (PV_NORMAL_DERIV_SINGLE, target normals Ntgt,
 * helm2.c:127-175).  Sampled on the GPU by the sampling kernel; returns a
 * host matrix.  The hypersingular layer (PV_NORMAL_DERIV_DOUBLE), for which
 * the reference has no case either, returns BF_ERROR_NOT_IMPLEMENTED. */
BfMat *bfHelm2GetKernelMatrix(BfHelm2 const *helm, BfPoints2 const *Xsrc, BfPoints2 const *Xtgt,
                              BfVectors2 const *Nsrc, BfVectors2 const *Ntgt);

/* fac_helm2.c:943-1002: multilevel butterfly of the kernel from the points
 * of `srcTree` to the points of `tgtTree`, with the reference's planning
 * rules (dense below 128*128 entries, rank estimates at 1e-15,
 * bfFacHelm2Prepare's level choice).  Factors live in GPU memory; the
 * returned matrix has type BF_TYPE_MAT_HIP_BUTTERFLY and is applied with
 * bfMatMul.  Layer potentials: single, double (PV_DOUBLE) and combined
 * field, the last two with the unit normals of `srcTree` (its proxies use
 * the same layer with the circles' radial normals, layer_pot.h:63-69,
 * fac_helm2.c:83-118), and S' (PV_NORMAL_DERIV_SINGLE) with the unit normals
 * of `tgtTree`: its re-expansions use the single layer (BF_PROXY_LAYER_POT,
 * fac_helm2.c:58-59, :230-231) and the target normals enter only the kernel
 * blocks of the last factors and the dense leaves (:455-465, :755-771).
 * PV_NORMAL_DERIV_DOUBLE returns BF_ERROR_NOT_IMPLEMENTED. */
BfMat *bfFacHelm2MakeMultilevel(BfHelm2 const *helm, BfQuadtree const *srcTree,
                                BfQuadtree const *tgtTree);

/* tree_traversals.h */
typedef enum BfTreeTraversals {
  BF_TREE_TRAVERSAL_UNKNOWN,
  BF_TREE_TRAVERSAL_LR_LEVEL_ORDER,
  BF_TREE_TRAVERSAL_LR_REVERSE_LEVEL_ORDER,
  BF_TREE_TRAVERSAL_POST_ORDER
} BfTreeTraversal;

/* tree_level_iter.h:7-12: the nodes of the subtree under a node, level by
 * level; `levelNodes` holds the current level's BfQuadtreeNode pointers */
/* ptr_array.h:7-12 (at file scope, so C++ callers see the same type) */
typedef struct BfPtrArray {
  void **data;
  BfSize capacity;
  BfSize num_elts;
  bool isView;
} BfPtrArray;
typedef struct BfTreeLevelIter {
  BfTreeTraversal traversal;
  BfPtrArray nodes;
  BfPtrArray *levelNodes;
  void *aux;
} BfTreeLevelIter;
/* tree_level_iter.c:184-330 (LR_LEVEL_ORDER goes down from the node,
 * LR_REVERSE_LEVEL_ORDER up from the deepest level) */
void bfTreeLevelIterInit(BfTreeLevelIter *iter, BfTreeTraversal traversal, BfTreeNode *node);
void bfTreeLevelIterDeinit(BfTreeLevelIter *iter);
BfSize bfTreeLevelIterCurrentDepth(BfTreeLevelIter const *iter);
bool bfTreeLevelIterIsDone(BfTreeLevelIter const *iter);
void bfTreeLevelIterNext(BfTreeLevelIter *iter);
BfSize bfTreeLevelIterGetNumPoints(BfTreeLevelIter const *iter);
bool bfTreeLevelIterCurrentLevelIsInternal(BfTreeLevelIter const *iter);
BfPtrArray *bfTreeLevelIterGetLevelNodes(BfTreeLevelIter *iter);

typedef struct BfMatProduct BfMatProduct;

/* fac_helm2.c:551-651: picks the source level of a butterfly between
 * srcNode and tgtNode with the reference's rules (the deepest fully
 * internal target level caps the source depth; the source level must hold
 * all of srcNode's points, be internal, and every node's rank estimate at
 * 1e-15 must not exceed its points).  Initializes srcLevelIter (reverse
 * level order, left on the chosen level) and tgtLevelIter (level order, on
 * tgtNode) and returns the number of factors, or 0 when the pair cannot be
 * butterflied; the caller deinitializes the iterators. */
BfSize bfFacHelm2Prepare(BfHelm2 const *helm, BfQuadtreeNode const *srcNode, BfQuadtreeNode const *tgtNode,
                         BfTreeLevelIter *srcLevelIter, BfTreeLevelIter *tgtLevelIter);
/* fac_helm2.c:653-704: builds the numFactors factors from the iterators
 * prepared by bfFacHelm2Prepare (advancing them as the reference does):
 * the sampling and re-expansion blocks are computed by the HIP kernels,
 * then copied into host containers of the reference's shape -- a
 * BfMatProduct whose factors, applied last first, are the evaluation
 * factor (BfMatBlockDiag of target-points x proxies kernel blocks), the
 * re-expansion factors (BfMatBlockCoo) and the first factor
 * (BfMatBlockDiag), every block a BfMatDenseComplex.  The product maps the
 * source node's points (tree order) to the target node's points.  srcNode
 * and tgtNode must be at one depth, as in the reference's multilevel
 * split.  NULL with BF_ERROR_INVALID_ARGUMENTS when numFactors < 2. */
BfMatProduct *bfFacHelm2Make(BfHelm2 const *helm, BfQuadtree const *srcTree, BfQuadtree const *tgtTree,
                             BfTreeLevelIter *srcLevelIter, BfTreeLevelIter *tgtLevelIter, BfSize numFactors);
/* fac_helm2.c:706-729: Prepare + Make for one pair of nodes (same depth);
 * the product as a BfMat (BF_TYPE_MAT_PRODUCT).  NULL with
 * BF_ERROR_RUNTIME_ERROR when the pair cannot be butterflied (the reference
 * has no result there either). */
BfMat *bfFacHelm2MakeSingleLevel(BfHelm2 const *helm, BfQuadtreeNode const *srcNode, BfQuadtreeNode const *tgtNode);

/* new: planner options beside the reference rules (no counterpart: the
 * reference hard-codes them, fac_helm2.c:20, :102, :291-296, :523) */
typedef struct BfFacHelm2Options {
  BfReal tol;            /* rank estimate p = ceil(k r1 r2/d - C log10(tol)); reference 1e-15 */
  BfReal rankConst;      /* C above; reference 1 */
  BfSize maxDenseSize;   /* blocks with m*n below this are dense; reference 128*128 */
  bool passThrough;      /* carry shallow leaves down the butterfly levels
                            instead of bfFacHelm2Prepare's depth caps */
  BfReal pinvRcond;      /* re-expansion pseudo-inverses also drop singular
                            values below pinvRcond * sigma_0; reference 0
                            (only max(m,n) eps sigma_0 + eps,
                            mat_dense_complex.c:1797-1809) */
} BfFacHelm2Options;
void bfFacHelm2OptionsInitDefault(BfFacHelm2Options *opts);
BfMat *bfFacHelm2MakeMultilevelWithOptions(BfHelm2 const *helm, BfQuadtree const *srcTree,
                                           BfQuadtree const *tgtTree, BfFacHelm2Options const *opts);

/* new (no counterpart): run only the host planner -- the quadtree level
 * choices, block shapes and apply stages -- without touching a GPU.  The
 * result supports the plan export below and bfMatDelete, not bfMatMul.
 * For tools that size a factorization before running it. */
BfMat *bfhipFacHelm2Plan(BfHelm2 const *helm, BfQuadtree const *srcTree, BfQuadtree const *tgtTree,
                         BfFacHelm2Options const *opts);

/* new (batched independent problems): factor `count` problems -- problem i
 * is helms[i] on srcTrees[i] -> tgtTrees[i] with opts[i], each planned as by
 * bfFacHelm2MakeMultilevelWithOptions -- as one block-diagonal operator
 * diag(A_0, ..., A_{count-1}).  The compression jobs of all problems share
 * one launch per LDS class and every apply stage is one launch over all
 * problems.  X stacks the problems' inputs in order, each in its own
 * source-tree order; Y likewise in target-tree order.  The result is a
 * BF_TYPE_MAT_HIP_BUTTERFLY (bfMatMul, bfhipMatMulDevice, FP32 forms). */
BfMat *bfhipFacHelm2MakeBatched(BfSize count, BfHelm2 const *helms, BfQuadtree const *const *srcTrees,
                                BfQuadtree const *const *tgtTrees, BfFacHelm2Options const *opts);

/* mat.c dispatchers (mat.h:45-110) */
BfType bfMatGetType(BfMat const *mat);
BfSize bfMatGetNumRows(BfMat const *mat);
BfSize bfMatGetNumCols(BfMat const *mat);
BfSize bfMatNumBytes(BfMat const *mat);
BfMat *bfMatCopy(BfMat const *mat);
void bfMatDelete(BfMat **mat);
/* mat.c:183-185: Y = A X.  For a device factorization X is N_src x nrhs in
 * source-tree order and Y comes back N_tgt x nrhs in target-tree order. */
BfMat *bfMatMul(BfMat const *op1, BfMat const *op2);
/* mat_dense_complex.c:741-774: row i moves to row perm->index[i] */
void bfMatPermuteRows(BfMat *mat, BfPerm const *perm);

/* mat_dense_complex.h:69-137 */
BfMatDenseComplex *bfMatDenseComplexNew(void);
void bfMatDenseComplexInit(BfMatDenseComplex *mat, BfSize numRows, BfSize numCols);
void bfMatDenseComplexInitViewFromPtr(BfMatDenseComplex *mat, BfSize numRows, BfSize numCols,
                                      BfComplex *data);
BfMatDenseComplex *bfMatDenseComplexNewRandn(BfSize numRows, BfSize numCols);
BfMatDenseComplex *bfMatDenseComplexNewViewFromPtr(BfSize numRows, BfSize numCols, BfComplex *data);
void bfMatDenseComplexDeinit(BfMatDenseComplex *mat);
void bfMatDenseComplexDeinitAndDealloc(BfMatDenseComplex **mat);
BfComplex *bfMatDenseComplexGetDataPtr(BfMatDenseComplex *mat);
BfMat *bfMatDenseComplexToMat(BfMatDenseComplex *mat);
BfMatDenseComplex *bfMatToMatDenseComplex(BfMat *mat);

/* new (no counterpart): on a device factorization, the number of factors
 * of butterfly leaf `leaf` (the reference's BfMatProduct per block) */
BfSize bfhipFacLeafNumFactors(BfMat const *fac, BfSize leaf);

/* ------------------------------------------------ device entry points */

/* Precision of the apply: FP64 factors and vectors; FP32 factors and
 * vectors (accumulation set by bfhipSetApplyCompensation); or FP32 hi+lo
 * factors with FP32 vectors, the pair summed in FP64 before the product
 * (FP64 products with compensation on one RHS, the FP64 matrix cores on
 * many), so only the vectors round to FP32. */
typedef enum BfhipPrecision {
  BFHIP_PRECISION_FP64 = 0,
  BFHIP_PRECISION_FP32 = 1,
  BFHIP_PRECISION_FP32X2 = 2,
} BfhipPrecision;

/* new (no counterpart): device memory helpers, so callers without a GPU
 * runtime of their own (ctypes, FFI) can stage inputs in HBM */
void *bfhipMalloc(BfSize numBytes);
void bfhipFree(void *ptr);
int bfhipMemcpyHtoD(void *dst, void const *src, BfSize numBytes);
int bfhipMemcpyDtoH(void *dst, void const *src, BfSize numBytes);
int bfhipSynchronize(void);
int bfhipDeviceCount(void);
/* new (no counterpart): device-to-device copy by a plain streaming kernel
 * (16-byte elements), for callers calibrating HBM byte counters */
int bfhipDeviceCopy(void const *src, void *dst, BfSize numBytes);

/* new (no counterpart): Y = A X with X, Y already in device memory, for
 * callers that keep vectors resident in HBM (benchmarks, solvers).
 * X: N_src x nrhs, Y: N_tgt x nrhs, complex, row-major, tree orders;
 * FP64 interleaved for BFHIP_PRECISION_FP64, FP32 interleaved otherwise.
 * Returns 0 on success. */
int bfhipMatMulDevice(BfMat const *fac, void const *dX, void *dY, BfSize nrhs, BfhipPrecision prec);

/* new (no counterpart in the reference): bfhipMatMulDevice without any
 * device allocation, for callers that cannot allocate (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc).
 * The intermediate vectors of the apply use a fixed 64 MiB device scratch
 * array that the library holds on each device until the process ends (two
 * 32 MiB halves); a factorization whose largest stage needs more than a
 * half for nrhs right-hand sides (rows x nrhs x 16 B in FP64, 8 B in FP32)
 * is refused with BF_ERROR_OUT_OF_RANGE.  Calls share that array: calls
 * that overlap on one device race.  Returns 0 on success. */
int bfhipMatMulDeviceNoAlloc(BfMat const *fac, void const *dX, void *dY, BfSize nrhs, BfhipPrecision prec);

/* new: convert the stored factors for the FP32 / FP32x2 applies (kept
 * beside the FP64 factors until bfMatDelete) */
int bfhipFacPrepareMixed(BfMat *fac, BfhipPrecision prec);
/* new: free the FP64 factors (and apply buffers) of a factorization that
 * has FP32 factors, leaving only the FP32 applies; for callers that keep
 * a mixed factorization resident and need the HBM */
int bfhipFacReleaseFp64(BfMat *fac);
/* new (no counterpart): change the basis of every intermediate vector of
 * the butterfly so that the map from it to the output has orthonormal
 * columns (a backward pass of thin QRs over the column blocks of each
 * factor, in FP64 on the device).  The operator is unchanged up to FP64
 * rounding; reduced-precision applies of the result lose about eps
 * relative per stage instead of eps times the cancellation of the
 * equivalent charges.  Call it before bfhipFacPrepareMixed.  Block exports
 * afterwards return the rotated blocks. */
int bfhipFacOrthogonalize(BfMat *fac);
/* new: wall seconds of the last bfhipFacOrthogonalize, -1 on a bad argument */
double bfhipFacOrthogonalizeSeconds(BfMat const *fac);

/* new (no counterpart): factorization statistics
 * out[0] factor wall seconds, [1] tree+plan seconds, [2] sampling+compression
 * seconds, [3] stored complex entries, [4] LstSq blocks, [5] kernel blocks,
 * [6] identity blocks, [7] dense leaves, [8] butterfly leaves, [9] stages,
 * [10] H0 evaluations */
void bfhipFacGetStats(BfMat const *fac, double out[16]);

/* new: average device time (ms) of the last launches of each kernel,
 * measured with HIP events on the library stream:
 * out[0] sampling, [1] compression, [2] apply (sum of stages of one apply),
 * [3] bytes of factor data read by one FP64 apply, [4] number of apply stages */
void bfhipGetKernelTimes(double out[8]);
/* new: the same, summed since the last reset, with the number of
 * measurements per slot (slot 3 is the FP32 apply) */
void bfhipGetKernelTimeSums(double sums[8], double counts[8]);
void bfhipResetKernelTimes(void);
/* new: launch shapes of the apply, for tuning: the 1-RHS apply's waves per
 * block in {1, 2, 4} (one row segment per wave) and, when `unroll` is 1,
 * the power-of-two lane map instead of the default map whose rows match
 * the segment height; the many-RHS apply's right-hand-side
 * tiles of 16 per wave on the matrix cores in {1, 2, 4, 8}, plus 16 for
 * two 16-row tiles per wave (tiles 1, 2 or 4), 0 for the vector-unit
 * kernel, or -1 (the default: the shapes the tuner picks at N = 2^20, 2
 * tiles x 2 row tiles in FP64 and 2 tiles x 1 row tile in FP32) */
int bfhipSetApplyConfig(int wavesPerBlock, int unroll, int mfmaTiles);
/* new (no counterpart): hipGraph replay of the apply's stage chain, off (0,
 * the default) or on.  When on, bfhipMatMulDevice captures the stage
 * launches of a factorization once and replays the graph while the input
 * and output pointers, right-hand-side count, precision, launch shapes
 * and the library's buffers stay the same (no memset or copy inside the
 * graph); anything else recaptures.  For callers that apply one operator
 * many times (iterative solvers, batched serving). */
int bfhipSetApplyGraphs(int on);
/* new (no counterpart): waves per re-expansion block in the compression.
 * Blocks whose LDS workspace (4-KiB classes) exceeds `twoAbove` bytes are
 * solved by 2 waves, above `fourAbove` by 4 (defaults 16384 and 40960:
 * the p >~ 22 blocks, whose LDS leaves room for only a few per CU); the
 * others by one wave.  Every shape gives the reference's pseudo-inverse
 * solution; they differ only in the order of rounding.  A negative value
 * leaves that threshold unchanged; the previous values go to *oldTwo /
 * *oldFour when not NULL.  Applies to the factorizations made after the
 * call. */
void bfhipSetLstsqWaves(int64_t twoAbove, int64_t fourAbove, int64_t *oldTwo, int64_t *oldFour);
/* new (no counterpart): the apply's autotuner.  For each apply stage of
 * `fac` it times the launch shapes that fit `nrhs` (one right-hand side:
 * 1, 2 or 4 waves per block with either lane map; 16 and more: 1, 2, 4 or
 * 8 tiles of 16 right-hand sides per matrix-core wave, 1, 2 or 4 with two
 * row tiles per wave, or the vector-unit kernel) on that stage alone and keeps the fastest, per bucket of stages
 * (rounded log2 of the mean segment height and of the mean columns per
 * segment), for the precision and the class of nrhs (1, 2-15, 16-31,
 * 32-63, 64-127, 128-255, 256+).  Buckets found in the text file
 * `cachePath` (or $BF_AMD_TUNE_CACHE when NULL; none when both are
 * absent) are not timed again; new ones are appended to it.  Later applies
 * of `fac` in that class use the chosen shapes.  Allocates work vectors
 * while it runs. */
int bfhipFacTuneApply(BfMat *fac, BfSize nrhs, BfhipPrecision prec, char const *cachePath);
/* new: the launch shape of each stage used for nrhs right-hand sides at
 * `prec` (-1 = the global settings); returns the number of stages */
int bfhipFacGetApplyConfig(BfMat const *fac, BfSize nrhs, BfhipPrecision prec, int32_t *stageCfg);

/* new: accumulation in the FP32 applies (factors and vectors stored in
 * FP32).  0: FP32 throughout.  1: FP64 sums -- one RHS: FP32 products
 * summed in FP64; many RHS: the FP32 operands widened to FP64 on load and
 * multiplied on the FP64 matrix cores -- so the rounding of a stage is one
 * FP32 rounding of its output however long its rows are.  2 + m (m = 0, 1,
 * 3 or 7): one RHS as 1; many RHS accumulate FP32 matrix-core tiles and
 * fold them into FP64 sums every m + 1 four-column steps.  The default is
 * 2 (a fold every step): at N = 2^20 and 256 RHS it is 1.13x faster than 1
 * at 8.6e-7 instead of 3.9e-7 relative error. */
int bfhipSetApplyCompensation(int on);

/* new: plan export for parity checks (integers, in the oracle's canonical
 * order): leafTab [L x 6], facTab [F x 4], blkTab [B x 7] */
void bfhipFacPlanSizes(BfMat const *fac, BfSize *numLeaves, BfSize *numFactors, BfSize *numBlocks);
void bfhipFacPlanExport(BfMat const *fac, int64_t *leafTab, int64_t *facTab, int64_t *blkTab);
/* copy block `idx` (export order) as m x n row-major complex FP64 */
int bfhipFacBlockData(BfMat const *fac, BfSize idx, BfComplex *out);
int bfhipFacDenseData(BfMat const *fac, BfSize leafIdx, BfComplex *out);
/* geometry of block `idx`: for a re-expansion block
 * [equiv cx, cy, r, test cx, cy, r, orig cx, cy, r, origOff (-1 = circle), p, n];
 * for a kernel block [src cx, cy, r, tgtOff, 0.., srcOff (-1 = circle), m, n]
 * (offsets index the tree-ordered points) */
int bfhipFacBlockGeometry(BfMat const *fac, BfSize idx, BfReal out[12]);

/* new: device H0^(1) on n arguments (host arrays), out = [J0, Y0] pairs */
int bfhipHankel0(BfSize n, BfReal const *x, BfReal *out);

/* new: batched pseudo-inverse least squares on the device with the
 * reference's truncation rule (mat_dense_complex.c:1768-1849):
 * X_b = pinv(A_b) B_b, A_b p x p, B_b p x n, all row-major complex,
 * host arrays.  Returns 0 on success. */
int bfhipLstSqBatched(BfSize batch, BfSize p, BfSize n, BfComplex const *A, BfComplex const *B,
                      BfComplex *X);

/* ------------------------------------------- streaming factorization */

/* def.h:57-61 */
typedef enum BfPolicies {
  BF_POLICY_VIEW,
  BF_POLICY_COPY,
  BF_POLICY_STEAL,
} BfPolicy;

/* ------------------------------------------- the block matrix algebra */

/* ptr_array.h:7-40 (the struct is declared with BfTreeLevelIter above) */
void bfInitPtrArray(BfPtrArray *arr, BfSize capacity);
void bfInitPtrArrayWithDefaultCapacity(BfPtrArray *arr);
void bfPtrArrayDeinit(BfPtrArray *arr);
BfSize bfPtrArraySize(BfPtrArray const *arr);
bool bfPtrArrayIsEmpty(BfPtrArray const *arr);
void bfPtrArrayAppend(BfPtrArray *arr, void *ptr);
void *bfPtrArrayGet(BfPtrArray const *arr, BfSize pos);
void *bfPtrArrayPopLast(BfPtrArray *arr);

/* mat.h:199: the matrix itself (VIEW: the caller keeps ownership; STEAL:
 * ownership passes) or a copy (COPY) */
BfMat *bfMatGet(BfMat *mat, BfPolicy policy);

/* mat_block.h:39-66.  Block i covers rows rowOffset[r] .. rowOffset[r+1]
 * and columns colOffset[c] .. colOffset[c+1] of its block row r and block
 * column c (block-diagonal: r = c = i; block-dense: block r * numBlockCols
 * + c; block-COO: r = rowInd[i], c = colInd[i]).  Callers may fill the
 * arrays directly, as the reference's drivers do. */
typedef struct BfMatBlock {
  BfMat super;
  void *vtbl;        /* unused (the reference's block vtable) */
  BfMat **block;
  BfSize *rowOffset; /* numBlockRows + 1 entries */
  BfSize *colOffset; /* numBlockCols + 1 entries */
} BfMatBlock;

/* mat_block_diag.h:33-37 */
typedef struct BfMatBlockDiag {
  BfMatBlock super;
  BfSize numBlockRows, numBlockCols; /* the block grid (numBlocks = min of the two) */
  bool *ownsBlock;                   /* blocks this matrix deletes */
} BfMatBlockDiag;

/* mat_block_coo.h:34-46 */
typedef struct BfMatBlockCoo {
  BfMatBlock super;
  BfSize numBlocks;
  BfSize *rowInd;
  BfSize *colInd;
  BfSize numBlockRows, numBlockCols;
  bool *ownsBlock;
} BfMatBlockCoo;

/* mat_block_dense.h:60-62 */
typedef struct BfMatBlockDense {
  BfMatBlock super;
  BfSize numBlockRows, numBlockCols;
  bool *ownsBlock;
} BfMatBlockDense;

/* mat_product.h:28-31: the product F_0 F_1 ... F_{n-1}; applied last
 * factor first */
struct BfMatProduct {
  BfMat super;
  BfPtrArray factorArr;
};

/* The containers are host objects; bfMatMul on one whose leaves are dense
 * complex matrices (or containers of them) packs the leaves into a device
 * product -- one apply stage per factor, the leaves of a factor as that
 * stage's blocks -- and runs the butterfly's stage kernels; X and Y are
 * host matrices.  bfhipMatPack returns that device product for repeated
 * applies (BF_TYPE_MAT_HIP_PRODUCT, bfMatMul / bfMatDelete).  A product
 * with other factors (a device factorization, say) is applied factor by
 * factor, each with its own bfMatMul (mat_product.c:211-245). */
BfMatProduct *bfMatProductNew(void);
void bfMatProductInit(BfMatProduct *prod);
void bfMatProductDeinit(BfMatProduct *prod);
void bfMatProductDealloc(BfMatProduct **prod);
void bfMatProductDeinitAndDealloc(BfMatProduct **prod);
BfSize bfMatProductNumFactors(BfMatProduct const *prod);
BfMat *bfMatProductGetFactor(BfMatProduct *prod, BfSize i);
BfMat const *bfMatProductGetFactorConst(BfMatProduct const *prod, BfSize i);
/* appends `mat` (the product takes ownership): prod <- prod * mat */
void bfMatProductPostMultiply(BfMatProduct *prod, BfMat *mat);
BfMat *bfMatProductPopLastFactor(BfMatProduct *prod);
BfMat *bfMatProductToMat(BfMatProduct *prod);
BfMatProduct *bfMatToMatProduct(BfMat *mat);

BfMatBlockDiag *bfMatBlockDiagNew(void);
BfMatBlockDiag *bfMatBlockDiagNewFromBlocks(BfPtrArray *blocks, BfPolicy policy);
void bfMatBlockDiagInit(BfMatBlockDiag *mat, BfSize numBlockRows, BfSize numBlockCols);
void bfMatBlockDiagDeinit(BfMatBlockDiag *mat);
void bfMatBlockDiagDealloc(BfMatBlockDiag **mat);
void bfMatBlockDiagDeinitAndDealloc(BfMatBlockDiag **mat);
BfSize bfMatBlockDiagNumBlocks(BfMatBlockDiag const *mat);
BfMat *bfMatBlockDiagGetBlock(BfMatBlockDiag *mat, BfSize i);
BfMat const *bfMatBlockDiagGetBlockConst(BfMatBlockDiag const *mat, BfSize i);
/* takes ownership of `block`; the offsets follow the blocks' shapes */
void bfMatBlockDiagSetBlock(BfMatBlockDiag *mat, BfSize i, BfMat *block);
BfMat *bfMatBlockDiagToMat(BfMatBlockDiag *mat);
BfMatBlockDiag *bfMatToMatBlockDiag(BfMat *mat);

BfMatBlockCoo *bfMatBlockCooNew(void);
/* the caller fills rowOffset, colOffset, rowInd, colInd and block */
void bfMatBlockCooInit(BfMatBlockCoo *mat, BfSize numBlockRows, BfSize numBlockCols, BfSize numBlocks);
void bfMatBlockCooDeinit(BfMatBlockCoo *mat);
void bfMatBlockCooDealloc(BfMatBlockCoo **mat);
void bfMatBlockCooDeinitAndDealloc(BfMatBlockCoo **mat);
BfMat *bfMatBlockCooToMat(BfMatBlockCoo *mat);
BfMatBlockCoo *bfMatToMatBlockCoo(BfMat *mat);

BfMatBlockDense *bfMatBlockDenseNew(void);
BfMatBlockDense *bfMatBlockDenseNewFromBlocks(BfSize numRowBlocks, BfSize numColBlocks, BfPtrArray *blocks,
                                              BfPolicy policy);
void bfMatBlockDenseInit(BfMatBlockDense *mat, BfSize numBlockRows, BfSize numBlockCols);
void bfMatBlockDenseDeinit(BfMatBlockDense *mat);
void bfMatBlockDenseDealloc(BfMatBlockDense **mat);
void bfMatBlockDenseDeinitAndDealloc(BfMatBlockDense **mat);
BfMat *bfMatBlockDenseGetBlock(BfMatBlockDense *mat, BfSize i, BfSize j);
/* takes ownership of `block`; the offsets follow the blocks' shapes */
void bfMatBlockDenseSetBlock(BfMatBlockDense *mat, BfSize i, BfSize j, BfMat *block);
BfMat *bfMatBlockDenseToMat(BfMatBlockDense *mat);
BfMatBlockDense *bfMatToMatBlockDense(BfMat *mat);

/* new (no counterpart): the device product of a container of dense
 * leaves, for repeated applies */
BfMat *bfhipMatPack(BfMat const *mat);

/* vec.h:74-78, vec_complex.h:20-34: complex vectors (entry i at
 * data[i * stride]) */
typedef struct BfVec {
  void *vtbl;
  int props;
  BfSize size;
} BfVec;
typedef struct BfVecComplex {
  BfVec super;
  BfSize stride;
  BfComplex *data;
} BfVecComplex;
BfVecComplex *bfVecComplexNew(void);
void bfVecComplexInit(BfVecComplex *vec, BfSize size);
void bfVecComplexInitView(BfVecComplex *vec, BfSize size, BfSize stride, BfComplex *data);
void bfVecComplexDeinit(BfVecComplex *vec);
void bfVecComplexDeinitAndDealloc(BfVecComplex **vec);
BfVec *bfVecComplexToVec(BfVecComplex *vec);
BfVecComplex *bfVecToVecComplex(BfVec *vec);
void bfVecDelete(BfVec **vec);
/* mat.h:87: y = A x for a complex vector, through bfMatMul on an n x 1
 * matrix (so a device factorization or a container applies on the GPU) */
BfVec *bfMatMulVec(BfMat const *mat, BfVec const *vec);

/* fac.h:6-44.  The trees are quadtrees of this library (BfQuadtree's
 * BfTree super); rows of fed blocks follow the row tree's order and their
 * columns the column tree's. */
typedef struct BfFacSpec {
  BfTree *rowTree;
  BfTree *colTree;
  BfSize rowTreeInitDepth;
  BfReal tol;
  BfSize minNumRows;
  BfSize minNumCols;
  bool compareRelativeErrors;
  bool bailIfBottomedOut;
} BfFacSpec;

typedef struct BfFac BfFac;
typedef struct BfFacSpan BfFacSpan;
typedef struct BfFacStreamer BfFacStreamer;

/* fac_streamer.h:9-20.  Feed takes the dense complex block of the current
 * column leaf (all rows); the SVDs of each round run batched on the
 * device.  Where the reference stops on a row leaf that does not compress
 * (unless bailIfBottomedOut), this one passes the leaf's rows through. */
BfFacStreamer *bfFacStreamerNew(void);
void bfFacStreamerInit(BfFacStreamer *facStreamer, BfFacSpec const *facSpec);
void bfFacStreamerDeinit(BfFacStreamer *facStreamer);
void bfFacStreamerDealloc(BfFacStreamer **facStreamer);
void bfFacStreamerDelete(BfFacStreamer **facStreamer);
BfSize bfFacStreamerGetNumRows(BfFacStreamer const *facStreamer);
bool bfFacStreamerFeed(BfFacStreamer *facStreamer, BfMat *Phi);
bool bfFacStreamerIsDone(BfFacStreamer const *facStreamer);
BfFac *bfFacStreamerGetFac(BfFacStreamer const *facStreamer);
BfFacSpan *bfFacStreamerGetFacSpan(BfFacStreamer const *facStreamer);
BfQuadtreeNode const *bfFacStreamerGetCurrentColumnNode(BfFacStreamer const *facStreamer);
BfPerm const *bfFacStreamerGetRowTreeReversePerm(BfFacStreamer const *facStreamer);

/* fac.h:52-56: the product Psi W0 ... W_L as a device matrix
 * (BF_TYPE_MAT_HIP_PRODUCT, applied with bfMatMul; every policy copies) */
void bfFacDeinitAndDealloc(BfFac **fac);
BfMat *bfFacGetMat(BfFac const *fac, BfPolicy policy);
/* fac.c:53-75: the same product as host containers: a BfMatProduct of
 * Psi, W0, ..., W_L, each a BfMatBlockCoo of BfMatDenseComplex blocks on
 * the grid of its blocks' row and column boundaries (a block that spans
 * several grid cells is cut into one block per cell); every policy
 * copies.  bfMatMul on it packs the blocks into a device product. */
BfMatProduct *bfFacGetMatProduct(BfFac const *fac, BfPolicy policy);
BfSize bfFacGetNumBytes(BfFac const *fac);
/* fac_span.h: the partial factorizations side by side; the span does not
 * own them (they stay with the streamer) */
BfMat *bfFacSpanGetMat(BfFacSpan const *facSpan, BfPolicy policy);
void bfFacSpanDelete(BfFacSpan **facSpan);
/* new: introspection for parity checks */
BfSize bfFacSpanGetNumFacs(BfFacSpan const *facSpan);
BfSize bfFacGetNumW(BfFac const *fac);
BfSize bfFacGetRowNodes(BfFac const *fac, int32_t *out, BfSize cap);
int bfFacGetFactorShape(BfFac const *fac, BfSize which, BfSize *rows, BfSize *cols, BfSize *numBlocks);
int bfFacGetFactorDense(BfFac const *fac, BfSize which, BfComplex *out);

/* ------------------------------------ boundary integral equation, GMRES */

/* new (helm2_bie.c:95-121 assembled on the device): A = alpha I +
 * (S + C) W with S the device factorization `fac` of the kernel on the
 * boundary points (square, sources = targets; not owned, keep it alive),
 * W the quadrature weights (original point order), C the Kapur-Rokhlin
 * corrections of order krOrder (0, 2 or 6; quadrature.c:12-200, applied
 * before the weights as helm2_bie.c:98-103 does) along the boundary order
 * of `points`; `perm` is the tree permutation of fac's quadtree.  Vectors
 * are in tree order.  The corrections are values of fac's own layer:
 * G for S, alpha G for the combined field, S' with the row point's normal
 * (taken from fac's target tree) for S' (krComplexKernel, helm2.c:372-378);
 * the double layer takes none (krOrder must be 0). */
BfMat *bfhipHelm2BieNew(BfMat *fac, BfHelm2 const *helm, BfPoints2 const *points, BfReal const *weights,
                        BfPerm const *perm, BfSize krOrder, BfReal alpha);
/* new: y = A x on device vectors (n x nrhs, FP64 or FP32; FP32 needs
 * bfhipFacPrepareMixed on the factorization) */
int bfhipBieMulDevice(BfMat const *op, void const *dX, void *dY, BfSize nrhs, BfhipPrecision prec);
/* linalg.c:47-317: GMRES without restarts, maxNumIter Krylov vectors, one
 * right-hand side at a time, stopping at |r| <= tol |b|; A a device
 * boundary operator, B (and X0) host dense complex; the basis lives in HBM
 * (FP64, classical Gram-Schmidt twice).  M (left preconditioning) must be
 * NULL here. */
BfMat *bfSolveGMRES(BfMat const *A, BfMat const *B, BfMat *X0, BfReal tol, BfSize maxNumIter, BfSize *numIter,
                    BfMat const *M);
/* new: mixed precision GMRES: FP64 residuals and refinement, inner
 * GMRES(restart) with the FP32 apply; host b, x (tree order); returns 0
 * when |b - A x| <= tol |b| (FP64 apply), 1 if maxIter ran out */
int bfhipSolveGmresMixed(BfMat const *A, BfComplex const *b, BfComplex *x, BfReal tol, BfSize restart,
                         BfSize maxIter, BfSize *numIter, BfSize *numRefine, BfReal *relRes);

/* mat.c:63-73: write a device factorization in the bfMatDump framing (the
 * type, then its payload: the leaves as BF_TYPE_MAT_PRODUCT of
 * BF_TYPE_MAT_BLOCK_DIAG / BF_TYPE_MAT_BLOCK_COO factors of dense complex
 * and identity blocks, mat_product.c:123-140, mat_block_diag.c:241-265,
 * mat_block_coo.c:260-290); the layout is in host_io.c.  The reference
 * defines no Dump for dense complex blocks; here they are their size and
 * row-major entries. */
void bfMatDump(BfMat const *mat, FILE *fp);
void bfMatSave(BfMat const *mat, char const *path);
/* new (no counterpart): read a factorization written by bfMatDump /
 * bfMatSave back into HBM (no factorization kernels run); for callers that
 * reuse a factorization across runs */
BfMat *bfhipFacLoad(FILE *fp);
BfMat *bfhipFacLoadFile(char const *path);

/* new (no counterpart; linalg.c:1040-1120 restated): batched truncated
 * SVDs on the device, A_b = U_b W_b with U_b m_b x k_b orthonormal and
 * W_b = S V^H (k_b x n_b), k_b = #{sigma >= tol sigma_0}; A, U, W are
 * packed column-major host arrays, U and W sized for full rank min(m,n). */
int bfhipSvdBatched(BfSize batch, BfSize const *m, BfSize const *n, BfComplex const *A, BfReal tol, BfSize *k,
                    BfComplex *U, BfComplex *W);
/* new (no counterpart): the widest matrix (columns) that bfhipSvdBatched
 * sends through its tall path (QR, then Jacobi on the triangle in LDS) --
 * m >= 2n and at least 2^15 entries -- or 0 when this GPU's LDS cannot
 * hold it (then such matrices take the one-launch-per-step path); -1
 * without a device */
int bfhipSvdTallMaxN(void);

/* new (no counterpart): the same truncated SVDs through a randomized range
 * finder of width width_b (Halko, Martinsson & Tropp 2011, Alg. 4.1 and
 * 5.1: Y = A Omega with a Gaussian Omega, Q = orth(Y), SVD of Q^H A).  A
 * sketch is accepted when ||A - Q Q^H A||_F <= tol sigma_0 and fewer than
 * width_b terms pass the cut; otherwise, or when width_b is 0 or not
 * below min(m_b, n_b), the matrix takes the full SVD (fellBack_b = 1 for a
 * rejected sketch).  The full SVD takes matrices whose smaller side is at
 * most 1024; a sketch needs only its width within that.  `seed` fixes
 * Omega. */
int bfhipSvdBatchedSketch(BfSize batch, BfSize const *m, BfSize const *n, BfComplex const *A, BfReal tol,
                          BfSize const *width, uint64_t seed, BfSize *k, BfComplex *U, BfComplex *W, int *fellBack);

/* new (no counterpart): the rank model, an FP32 MLP (6 -> 64 -> 64 -> 64
 * -> 1, GELU) that predicts the numerical rank of a node-pair block from
 * bfhipRankFeatures; trained (mini-batch Adam on the squared error) and
 * evaluated by HIP kernels.  Features are n x 6 row-major floats. */
typedef struct BfhipRankModel BfhipRankModel;
void bfhipRankFeatures(BfReal const colBox[4], BfReal const rowBox[4], BfReal k, BfSize level, BfReal tol,
                       float out[6]);
BfhipRankModel *bfhipRankModelNew(uint64_t seed);
void bfhipRankModelDelete(BfhipRankModel **model);
int bfhipRankModelTrain(BfhipRankModel *model, BfSize n, float const *features, float const *ranks, BfSize epochs,
                        BfSize batch, float lr, float *lastLoss);
int bfhipRankModelPredict(BfhipRankModel const *model, BfSize n, float const *features, float *ranks);
/* the gradient (standardized units, parameter layout of GetParams) and the
 * mean squared error over all n samples, without a step */
int bfhipRankModelGradient(BfhipRankModel const *model, BfSize n, float const *features, float const *ranks,
                           float *grad, float *loss);
BfSize bfhipRankModelNumParams(void);
int bfhipRankModelGetParams(BfhipRankModel const *model, float *params, float norm[14]);
int bfhipRankModelSetParams(BfhipRankModel *model, float const *params, float const norm[14]);

/* new (no counterpart): with a model, every truncated SVD of the stream
 * takes the randomized path with width = predicted rank + oversample (a
 * rejected sketch falls back to the full SVD); with or without one, the
 * (features, numerical rank) pair of every SVD is recorded for training.
 * `k` is the wavenumber of the kernel behind the stream (0 if none). */
int bfhipFacStreamerSetRankModel(BfFacStreamer *facStreamer, BfhipRankModel const *model, BfReal k,
                                 BfSize oversample);
BfSize bfhipFacStreamerGetRankSamples(BfFacStreamer const *facStreamer, float *features, float *ranks, BfSize cap);
void bfhipFacStreamerGetSketchStats(BfFacStreamer const *facStreamer, BfSize *svds, BfSize *sketched,
                                    BfSize *fellBack);

#ifdef __cplusplus
}
#endif
