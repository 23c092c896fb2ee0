/*
 * Host infrastructure of the path: error state, the seeded RNG, point
 * sets, permutations and the quadtree.  C11.
 */
#define _DEFAULT_SOURCE 1
#include <math.h>
#include <pthread.h>
#include <stdio.h>
#include <stdlib.h>
#include <sys/mman.h>
#include <time.h>

#include "bf_host.h"

double bfWallTime(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return ts.tv_sec + 1e-9 * ts.tv_nsec;
}

/* Large host tables (the plan's job and block tables run to gigabytes at
 * 2^20 points): ask for transparent huge pages, so that their first touch
 * takes a fault per 2 MiB instead of per 4 KiB.  Advice only: the call
 * changes nothing when the kernel declines. */
void bfHugePages(void *p, size_t bytes) {
  if (!p || bytes < ((size_t)32 << 20)) return;
  const uintptr_t a = ((uintptr_t)p + 4095) & ~(uintptr_t)4095;
  const uintptr_t e = ((uintptr_t)p + bytes) & ~(uintptr_t)4095;
  if (e > a) (void)madvise((void *)a, e - a, MADV_HUGEPAGE);
}

/* BF_AMD_TRACE=1 in the environment: the host steps of a factorization
 * print their wall times to stderr */
static int g_trace = -1;

void bfTrace(const char *what, double since) {
  if (g_trace < 0) {
    const char *e = getenv("BF_AMD_TRACE");
    g_trace = e && e[0] && e[0] != '0';
  }
  if (g_trace) fprintf(stderr, "[bf] %-28s %9.2f ms\n", what, 1e3 * (bfWallTime() - since));
}

/* ------------------------------------------------------------ errors */

/* error.c:7-24: one process-global code; reading it clears it */
static enum BfError g_error = BF_ERROR_NONE;

enum BfError bfGetError(void) {
  enum BfError e = g_error;
  g_error = BF_ERROR_NONE;
  return e;
}

void bfSetError(enum BfError error) { g_error = error; }

/* --------------------------------------------------------------- RNG */

/* rand.c:19-76: splitmix64 expands the seed into the four words of a
 * xoshiro256+ state; doubles take the top 53 bits of each draw; normals
 * come in Box-Muller pairs (cos, sin) of consecutive uniform pairs. */
static uint64_t g_sm;
static uint64_t g_xs[4];

static uint64_t sm_next(void) {
  uint64_t z = (g_sm += UINT64_C(0x9e3779b97f4a7c15));
  z = (z ^ (z >> 30)) * UINT64_C(0xbf58476d1ce4e5b9);
  z = (z ^ (z >> 27)) * UINT64_C(0x94d049bb133111eb);
  return z ^ (z >> 31);
}

static uint64_t xs_next(void) {
  const uint64_t sum = g_xs[0] + g_xs[3];
  const uint64_t shifted = g_xs[1] << 17;
  g_xs[2] ^= g_xs[0];
  g_xs[3] ^= g_xs[1];
  g_xs[1] ^= g_xs[2];
  g_xs[0] ^= g_xs[3];
  g_xs[2] ^= shifted;
  g_xs[3] = (g_xs[3] << 45) | (g_xs[3] >> 19);
  return sum;
}

void bfSeed(BfSize seed) {
  g_sm = (uint64_t)seed;
  for (int i = 0; i < 4; ++i) g_xs[i] = sm_next();
}

BfReal bfRealUniform1(void) { return (BfReal)(xs_next() >> 11) * 0x1.0p-53; }

void bfRealUniform(BfSize n, BfReal *x) {
  for (BfSize i = 0; i < n; ++i) x[i] = bfRealUniform1();
}

void bfRealRandn(BfSize n, BfReal *x) {
  const BfReal twoPi = 6.283185307179586;
  BfSize pairs = n / 2;
  bfRealUniform(2 * pairs, x);
  for (BfSize i = 0; i < pairs; ++i) {
    BfReal mag = sqrt(-2 * log(x[2 * i]));
    BfReal ang = twoPi * x[2 * i + 1];
    x[2 * i] = mag * cos(ang);
    x[2 * i + 1] = mag * sin(ang);
  }
  if (n % 2 == 1) {
    BfReal u[2];
    bfRealUniform(2, u);
    x[n - 1] = sqrt(-2 * log(u[0])) * cos(twoPi * u[1]);
  }
}

void bfComplexRandn(BfSize n, BfComplex *x) { bfRealRandn(2 * n, (BfReal *)x); }

void bfPoint2SampleUniformlyFromBoundingBox(BfReal const bboxMin[2], BfReal const bboxMax[2], BfPoint2 p) {
  p[0] = (bboxMax[0] - bboxMin[0]) * bfRealUniform1() + bboxMin[0];
  p[1] = (bboxMax[1] - bboxMin[1]) * bfRealUniform1() + bboxMin[1];
}

/* ------------------------------------------------------------ points */

BfPoints2 *bfPoints2NewWithCapacity(BfSize capacity) {
  BfPoints2 *p = calloc(1, sizeof(BfPoints2));
  if (!p) { bfSetError(BF_ERROR_MEMORY_ERROR); return NULL; }
  p->capacity = capacity ? capacity : 1;
  p->data = malloc(p->capacity * sizeof(BfPoint2));
  if (!p->data) { free(p); bfSetError(BF_ERROR_MEMORY_ERROR); return NULL; }
  return p;
}

void bfPoints2Append(BfPoints2 *points, BfPoint2 const pt) {
  if (points->isView) { bfSetError(BF_ERROR_INVALID_ARGUMENTS); return; }
  if (points->size == points->capacity) {
    BfSize nc = 2 * points->capacity;
    void *nd = realloc(points->data, nc * sizeof(BfPoint2));
    if (!nd) { bfSetError(BF_ERROR_MEMORY_ERROR); return; }
    points->data = nd;
    points->capacity = nc;
  }
  points->data[points->size][0] = pt[0];
  points->data[points->size][1] = pt[1];
  points->size++;
}

BfPoints2 *bfPoints2NewViewFromPtr(BfReal *xy, BfSize n) {
  BfPoints2 *p = calloc(1, sizeof(BfPoints2));
  if (!p) { bfSetError(BF_ERROR_MEMORY_ERROR); return NULL; }
  p->data = (BfPoint2 *)xy;
  p->size = p->capacity = n;
  p->isView = true;
  return p;
}

void bfPoints2DeinitAndDealloc(BfPoints2 **points) {
  if (!points || !*points) return;
  if (!(*points)->isView) free((*points)->data);
  free(*points);
  *points = NULL;
}

/* -------------------------------------------------------------- perms */

BfPerm *bfPermGetReversePerm(BfPerm const *perm) {
  BfPerm *rev = calloc(1, sizeof(BfPerm));
  if (!rev) { bfSetError(BF_ERROR_MEMORY_ERROR); return NULL; }
  rev->size = perm->size;
  rev->index = malloc(perm->size * sizeof(BfSize));
  if (!rev->index) { free(rev); bfSetError(BF_ERROR_MEMORY_ERROR); return NULL; }
  for (BfSize i = 0; i < perm->size; ++i) rev->index[i] = BF_SIZE_BAD_VALUE;
  for (BfSize i = 0; i < perm->size; ++i) {
    BfSize j = perm->index[i];
    if (j >= perm->size || rev->index[j] != BF_SIZE_BAD_VALUE) {
      free(rev->index);
      free(rev);
      bfSetError(BF_ERROR_INVALID_ARGUMENTS);
      return NULL;
    }
    rev->index[j] = i;
  }
  return rev;
}

void bfPermDelete(BfPerm **perm) {
  if (!perm || !*perm) return;
  if (!(*perm)->isView) free((*perm)->index);
  free(*perm);
  *perm = NULL;
}

/* ----------------------------------------------------------- quadtree */

/* Child q of a node split at s holds the points with
 * (x > sx) == (q >> 1) and (y > sy) == (q & 1): q = 0 lower-left,
 * 1 upper-left, 2 lower-right, 3 upper-right; ties go low
 * (quadtree_node.c:148-194). */
static inline int quadrant_of(const BfReal *pt, const BfReal *split) {
  return ((pt[0] > split[0]) << 1) | (pt[1] > split[1]);
}

typedef BF_VEC(BfQuadtreeNode) NodeVec;

/* The root box: the points' bounding box (bmn, bmx) stretched to a square
 * about its centre (quadtree_node.c:284-306, bbox.c:18-30).  Both tree
 * builders call this, so their boxes agree bit for bit. */
static void root_square(const BfReal bmn[2], const BfReal bmx[2], BfReal mn[2], BfReal mx[2]) {
  mn[0] = bmn[0];
  mn[1] = bmn[1];
  mx[0] = bmx[0];
  mx[1] = bmx[1];
  BfReal w = mx[0] - mn[0], h = mx[1] - mn[1];
  if (w > h) {
    BfReal c = (mn[1] + mx[1]) / 2;
    mn[1] = w * (mn[1] - c) / h + c;
    mx[1] = w * (mx[1] - c) / h + c;
  } else {
    BfReal c = (mn[0] + mx[0]) / 2;
    mn[0] = h * (mn[0] - c) / w + c;
    mx[0] = h * (mx[0] - c) / w + c;
  }
}

typedef struct {
  NodeVec nodes;
  BfPoint2 const *pts;
  BfSize *perm;
  uint8_t *qc; /* quadrant of the point at each tree position, per split */
  BfSize leafSize;
  int topLevels;  /* levels split here before the rest goes to threads */
  BfI64Vec defer; /* nodes left for the threads */
} TreeBuilder;

/* Partition perm[lo:hi) by quadrant, visiting the index range in the same
 * order as the reference's in-place sift so that the tree order of points
 * inside every node matches it: for each of the first three quadrants,
 * skip the leading run that already belongs to it, then swap every later
 * member forward onto the boundary.  The quadrants are computed once into
 * qc (one gather of the points per split) and swapped along with perm. */
static void sift(TreeBuilder *b, BfSize lo, BfSize hi, const BfReal *split, BfSize off[5]) {
  off[0] = lo;
  off[4] = hi;
  BfSize *perm = b->perm;
  uint8_t *qc = b->qc;
  for (BfSize j = lo; j < hi; ++j) qc[j] = (uint8_t)quadrant_of(b->pts[perm[j]], split);
  for (int q = 0; q < 3; ++q) {
    BfSize bd = lo;
    while (bd < hi && qc[bd] == q) ++bd;
    for (BfSize j = bd + 1; j < hi; ++j) {
      if (qc[j] == q) {
        const BfSize t = perm[bd];
        perm[bd] = perm[j];
        perm[j] = t;
        qc[j] = qc[bd];
        qc[bd] = (uint8_t)q;
        ++bd;
      }
    }
    off[q + 1] = bd;
    lo = bd;
  }
}

static void build_node(TreeBuilder *b, size_t ni) {
  BfQuadtreeNode nd = b->nodes.a[ni];
  BfSize off[5];
  sift(b, nd.offset[0], nd.offset[4], nd.split, off);
  memcpy(b->nodes.a[ni].offset, off, sizeof(off));
  for (int q = 0; q < 4; ++q) {
    BfSize cnt = off[q + 1] - off[q];
    if (cnt == 0) continue;
    BfQuadtreeNode c;
    memset(&c, 0, sizeof(c));
    const BfQuadtreeNode *p = &b->nodes.a[ni];
    c.bboxMin[0] = (q >> 1) ? p->split[0] : p->bboxMin[0];
    c.bboxMax[0] = (q >> 1) ? p->bboxMax[0] : p->split[0];
    c.bboxMin[1] = (q & 1) ? p->split[1] : p->bboxMin[1];
    c.bboxMax[1] = (q & 1) ? p->bboxMax[1] : p->split[1];
    c.split[0] = (c.bboxMin[0] + c.bboxMax[0]) / 2;
    c.split[1] = (c.bboxMin[1] + c.bboxMax[1]) / 2;
    for (int r = 0; r < 4; ++r) c.offset[r] = off[q + 1];
    c.offset[0] = off[q];
    c.offset[4] = off[q + 1];
    for (int r = 0; r < 4; ++r) c.child[r] = -1;
    c.parent = (int32_t)ni;
    c.depth = p->depth + 1;
    c.index = q;
    BF_VEC_PUSH(b->nodes, c);
    size_t ci = b->nodes.n - 1;
    b->nodes.a[ni].child[q] = (int32_t)ci;
    if (cnt > b->leafSize) {
      if (c.depth < b->topLevels) build_node(b, ci);
      else if (b->topLevels > 0) BF_VEC_PUSH(b->defer, (int64_t)ci);
      else build_node(b, ci);
    }
  }
}

/* The subtrees below the top levels are disjoint index ranges of perm and
 * qc: each is built by its own builder (node 0 = its root) on a thread,
 * then appended to the main node list.  The final breadth-first numbering
 * depends only on the tree's shape, so the result is the serial one. */
typedef struct {
  TreeBuilder b;
} SubtreeTask;

static void *subtree_task(void *arg) {
  SubtreeTask *t = arg;
  build_node(&t->b, 0);
  return NULL;
}

static int build_deferred(TreeBuilder *b) {
  const size_t nd = b->defer.n;
  if (!nd) return 0;
  SubtreeTask *tasks = calloc(nd, sizeof(SubtreeTask));
  pthread_t *th = malloc(nd * sizeof(pthread_t));
  int *ok = calloc(nd, sizeof(int));
  if (!tasks || !th || !ok) {
    free(tasks); free(th); free(ok);
    return -1;
  }
  for (size_t i = 0; i < nd; ++i) {
    TreeBuilder *sb = &tasks[i].b;
    sb->pts = b->pts;
    sb->perm = b->perm;
    sb->qc = b->qc;
    sb->leafSize = b->leafSize;
    BF_VEC_PUSH(sb->nodes, b->nodes.a[b->defer.a[i]]);
  }
  for (size_t i = 0; i < nd; ++i) ok[i] = pthread_create(&th[i], NULL, subtree_task, &tasks[i]) == 0;
  for (size_t i = 0; i < nd; ++i) {
    if (ok[i]) pthread_join(th[i], NULL);
    else subtree_task(&tasks[i]);
  }
  for (size_t i = 0; i < nd; ++i) {
    NodeVec *sn = &tasks[i].b.nodes;
    const int32_t root = (int32_t)b->defer.a[i];
    const int32_t base = (int32_t)b->nodes.n - 1; /* local k >= 1 goes to base + k */
    for (size_t k = 0; k < sn->n; ++k) {
      BfQuadtreeNode nd2 = sn->a[k];
      for (int q = 0; q < 4; ++q)
        if (nd2.child[q] >= 0) nd2.child[q] = nd2.child[q] + base;
      if (k == 0) {
        nd2.parent = b->nodes.a[root].parent;
        b->nodes.a[root] = nd2;
      } else {
        nd2.parent = nd2.parent == 0 ? root : nd2.parent + base;
        BF_VEC_PUSH(b->nodes, nd2);
      }
    }
    BF_VEC_FREE(*sn);
  }
  free(tasks);
  free(th);
  free(ok);
  return 0;
}

/* Live quadtrees, so that a node can name its tree
 * (bfQuadtreeNodeGetQuadtreeConst): nodes live in one array per tree and
 * carry no pointer back to it.  Single-threaded, like the rest of the
 * host API. */
static BF_VEC(BfQuadtree const *) g_trees;

static void tree_register(BfQuadtree const *t) {
  for (size_t i = 0; i < g_trees.n; ++i)
    if (g_trees.a[i] == t) return;
  BF_VEC_PUSH(g_trees, t);
}

static void tree_unregister(BfQuadtree const *t) {
  for (size_t i = 0; i < g_trees.n; ++i)
    if (g_trees.a[i] == t) {
      g_trees.a[i] = g_trees.a[--g_trees.n];
      return;
    }
}

BfQuadtree const *bfQuadtreeNodeGetQuadtreeConst(BfQuadtreeNode const *node) {
  for (size_t i = 0; node && i < g_trees.n; ++i) {
    BfQuadtree const *t = g_trees.a[i];
    if (t->nodes && node >= t->nodes && node < t->nodes + t->numNodes) return t;
  }
  bfSetError(BF_ERROR_INVALID_ARGUMENTS);
  return NULL;
}

BfTreeNode *bfQuadtreeNodeToTreeNode(BfQuadtreeNode *node) { return node; }

BfQuadtree *bfQuadtreeNew(void) {
  BfQuadtree *t = calloc(1, sizeof(BfQuadtree));
  if (!t) bfSetError(BF_ERROR_MEMORY_ERROR);
  return t;
}

void bfQuadtreeInitWithLeafSize(BfQuadtree *tree, BfPoints2 const *points, BfVectors2 const *unitNormals,
                                BfSize leafSize) {
  memset(tree, 0, sizeof(*tree));
  BfSize n = points ? points->size : 0;
  if (n < 2 || leafSize < 1) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return;
  }
  tree->points = points;
  tree->unitNormals = unitNormals;
  tree->leafSize = leafSize;
  TreeBuilder b;
  memset(&b, 0, sizeof(b));
  b.pts = (BfPoint2 const *)points->data;
  b.leafSize = leafSize;
  b.perm = malloc(n * sizeof(BfSize));
  b.qc = malloc(n);
  if (!b.perm || !b.qc) {
    free(b.perm);
    free(b.qc);
    bfSetError(BF_ERROR_MEMORY_ERROR);
    return;
  }
  /* large trees: two levels here, up to 16 subtrees on threads */
  b.topLevels = n >= ((BfSize)1 << 16) ? 2 : 0;
  for (BfSize i = 0; i < n; ++i) b.perm[i] = i;

  BfReal bmn[2] = {INFINITY, INFINITY}, bmx[2] = {-INFINITY, -INFINITY}, mn[2], mx[2];
  for (BfSize i = 0; i < n; ++i)
    for (int d = 0; d < 2; ++d) {
      bmn[d] = fmin(bmn[d], b.pts[i][d]);
      bmx[d] = fmax(bmx[d], b.pts[i][d]);
    }
  root_square(bmn, bmx, mn, mx);
  BfQuadtreeNode root;
  memset(&root, 0, sizeof(root));
  root.bboxMin[0] = mn[0]; root.bboxMin[1] = mn[1];
  root.bboxMax[0] = mx[0]; root.bboxMax[1] = mx[1];
  root.split[0] = (mn[0] + mx[0]) / 2;
  root.split[1] = (mn[1] + mx[1]) / 2;
  root.offset[0] = 0;
  root.offset[4] = n;
  for (int r = 0; r < 4; ++r) root.child[r] = -1;
  root.parent = -1;
  root.index = -1;
  BF_VEC_PUSH(b.nodes, root);
  build_node(&b, 0);
  if (build_deferred(&b)) { bfSetError(BF_ERROR_MEMORY_ERROR); return; }
  BF_VEC_FREE(b.defer);
  free(b.qc);

  /* renumber breadth-first, children in quadrant order (the reference's
   * LR level order, tree_level_iter.c:10-43) */
  size_t nn = b.nodes.n;
  int32_t *order = malloc(nn * sizeof(int32_t)), *newId = malloc(nn * sizeof(int32_t));
  BfQuadtreeNode *nodes = malloc(nn * sizeof(BfQuadtreeNode));
  if (!order || !newId || !nodes) { bfSetError(BF_ERROR_MEMORY_ERROR); return; }
  size_t head = 0, tail = 0;
  order[tail++] = 0;
  while (head < tail) {
    int32_t v = order[head++];
    for (int q = 0; q < 4; ++q)
      if (b.nodes.a[v].child[q] >= 0) order[tail++] = b.nodes.a[v].child[q];
  }
  for (size_t i = 0; i < nn; ++i) newId[order[i]] = (int32_t)i;
  int32_t maxDepth = 0;
  for (size_t i = 0; i < nn; ++i) {
    nodes[i] = b.nodes.a[order[i]];
    if (nodes[i].parent >= 0) nodes[i].parent = newId[nodes[i].parent];
    for (int q = 0; q < 4; ++q)
      if (nodes[i].child[q] >= 0) nodes[i].child[q] = newId[nodes[i].child[q]];
    if (nodes[i].depth > maxDepth) maxDepth = nodes[i].depth;
  }
  for (size_t i = nn; i-- > 0;) {
    int32_t d = nodes[i].depth;
    for (int q = 0; q < 4; ++q)
      if (nodes[i].child[q] >= 0 && nodes[nodes[i].child[q]].maxDepthBelow > d)
        d = nodes[nodes[i].child[q]].maxDepthBelow;
    nodes[i].maxDepthBelow = d;
  }
  tree->levelOffset = calloc((size_t)maxDepth + 2, sizeof(BfSize));
  for (size_t i = 0; i < nn; ++i) tree->levelOffset[nodes[i].depth + 1] = i + 1;
  tree->numLevels = (BfSize)maxDepth + 1;
  free(order);
  free(newId);
  BF_VEC_FREE(b.nodes);
  tree->nodes = nodes;
  tree->numNodes = nn;
  tree->super.root = nodes;
  tree->super.perm = calloc(1, sizeof(BfPerm));
  tree->super.perm->index = b.perm;
  tree->super.perm->size = n;
  tree_register(tree);
}

/* The same tree built on the GPU (bf_tree.hip: Morton keys, a radix sort
 * and the node tables level by level).  The nodes, boxes and point ranges
 * are those of the builder above; the point order inside a leaf is the
 * key order, which is the reference's order when every leaf holds one
 * point (leaf size 1). */
void bfhipQuadtreeInitWithLeafSize(BfQuadtree *tree, BfPoints2 const *points, BfVectors2 const *unitNormals,
                                   BfSize leafSize) {
  memset(tree, 0, sizeof(*tree));
  BfSize n = points ? points->size : 0;
  if (n < 2 || leafSize < 1 || n > (BfSize)INT32_MAX) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return;
  }
  BfSize *perm = malloc(n * sizeof(BfSize));
  if (!perm) {
    bfSetError(BF_ERROR_MEMORY_ERROR);
    return;
  }
  BfQuadtreeNode *nodes = NULL;
  int64_t numNodes = 0, levelCount[64];
  int numLevels = 0;
  const int rc = bfk_quadtree((int64_t)n, (const double *)points->data, (int64_t)leafSize, root_square, &nodes,
                              &numNodes, levelCount, &numLevels, perm);
  if (rc) {
    free(perm);
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : rc == -3 ? BF_ERROR_OUT_OF_RANGE : BF_ERROR_RUNTIME_ERROR);
    return;
  }
  for (int64_t i = numNodes; i-- > 0;) { /* children follow their parents */
    int32_t d = nodes[i].depth;
    for (int q = 0; q < 4; ++q)
      if (nodes[i].child[q] >= 0 && nodes[nodes[i].child[q]].maxDepthBelow > d)
        d = nodes[nodes[i].child[q]].maxDepthBelow;
    nodes[i].maxDepthBelow = d;
  }
  tree->levelOffset = calloc((size_t)numLevels + 1, sizeof(BfSize));
  if (!tree->levelOffset) {
    free(perm);
    free(nodes);
    bfSetError(BF_ERROR_MEMORY_ERROR);
    return;
  }
  for (int d = 0; d < numLevels; ++d) tree->levelOffset[d + 1] = tree->levelOffset[d] + (BfSize)levelCount[d];
  tree->numLevels = (BfSize)numLevels;
  tree->points = points;
  tree->unitNormals = unitNormals;
  tree->leafSize = leafSize;
  tree->nodes = nodes;
  tree->numNodes = (BfSize)numNodes;
  tree->super.root = nodes;
  tree->super.perm = calloc(1, sizeof(BfPerm));
  tree->super.perm->index = perm;
  tree->super.perm->size = n;
  tree_register(tree);
}

void bfQuadtreeInit(BfQuadtree *tree, BfPoints2 const *points, BfVectors2 const *unitNormals) {
  /* LEAF_SIZE_THRESHOLD = 1 (quadtree_node.c:17) */
  bfQuadtreeInitWithLeafSize(tree, points, unitNormals, 1);
}

void bfQuadtreeDeinit(BfQuadtree *tree) {
  if (!tree) return;
  tree_unregister(tree);
  free(tree->nodes);
  free(tree->levelOffset);
  if (tree->super.perm) {
    free(tree->super.perm->index);
    free(tree->super.perm);
  }
  memset(tree, 0, sizeof(*tree));
}

void bfQuadtreeDealloc(BfQuadtree **tree) {
  if (!tree) return;
  free(*tree);
  *tree = NULL;
}

void bfQuadtreeDeinitAndDealloc(BfQuadtree **tree) {
  if (!tree || !*tree) return;
  bfQuadtreeDeinit(*tree);
  bfQuadtreeDealloc(tree);
}

BfTree *bfQuadtreeToTree(BfQuadtree *tree) { return &tree->super; }

BfPerm const *bfTreeGetPermConst(BfTree const *tree) { return tree->perm; }

BfCircle bfQuadtreeNodeGetBoundingCircle(BfQuadtreeNode const *node) {
  BfCircle c;
  c.r = hypot(node->bboxMax[0] - node->bboxMin[0], node->bboxMax[1] - node->bboxMin[1]) / 2;
  c.center[0] = (node->bboxMin[0] + node->bboxMax[0]) / 2;
  c.center[1] = (node->bboxMin[1] + node->bboxMax[1]) / 2;
  return c;
}

bool bfQuadtreeNodesAreSeparated(BfQuadtreeNode const *node1, BfQuadtreeNode const *node2) {
  BfCircle c1 = bfQuadtreeNodeGetBoundingCircle(node1);
  BfCircle c2 = bfQuadtreeNodeGetBoundingCircle(node2);
  BfReal R = hypot(c2.center[0] - c1.center[0], c2.center[1] - c1.center[1]);
  return R > c1.r + c2.r + 1e1 * BF_EPS_MACH;
}

BfSize bfQuadtreeNodeGetNumPoints(BfQuadtreeNode const *node) { return node->offset[4] - node->offset[0]; }

/* helm2.c:14-36 */
BfSize bfHelm2RankEstForTwoCircles(BfHelm2 const *helm, BfCircle const *circ1, BfCircle const *circ2, BfReal C,
                                   BfReal eps) {
  BfReal r1 = circ1->r, r2 = circ2->r;
  BfReal R = hypot(circ2->center[0] - circ1->center[0], circ2->center[1] - circ1->center[1]);
  BfReal d = R - r1 - r2;
  if (!(helm->k > 0) || !(C > 0) || !(eps > 0) || !(d > 0)) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return BF_SIZE_BAD_VALUE;
  }
  BfReal p = helm->k * r1 * r2 / d - C * log10(eps);
  return (BfSize)ceil(p);
}

/* ------------------------------------------------- tree level iterators */

/* tree_level_iter.c:10-330: the nodes of the subtree under a node in LR
 * level order (breadth first, children in quadrant order), one level at a
 * time, from the top (LR_LEVEL_ORDER) or from the deepest level up
 * (LR_REVERSE_LEVEL_ORDER).  `levelNodes` views the current level; Next
 * past the last level leaves the view on it (IsDone turns true going
 * down; going up the view stays on the top level). */
typedef struct {
  BfSize current, numLevels;
  BfSize *offsets;
  BfPtrArray view;
} LevelIterAux;

static void level_view(BfTreeLevelIter *it) {
  LevelIterAux *a = it->aux;
  const BfSize lv = a->current < a->numLevels ? a->current : a->numLevels - 1;
  a->view.data = it->nodes.data + a->offsets[lv];
  a->view.num_elts = a->view.capacity = a->offsets[lv + 1] - a->offsets[lv];
  a->view.isView = true;
  it->levelNodes = &a->view;
}

void bfTreeLevelIterInit(BfTreeLevelIter *it, BfTreeTraversal traversal, BfTreeNode *node) {
  memset(it, 0, sizeof(*it));
  it->traversal = traversal == BF_TREE_TRAVERSAL_UNKNOWN ? BF_TREE_TRAVERSAL_LR_LEVEL_ORDER : traversal;
  BfQuadtree const *t = bfQuadtreeNodeGetQuadtreeConst(node);
  if (!t || (it->traversal != BF_TREE_TRAVERSAL_LR_LEVEL_ORDER &&
             it->traversal != BF_TREE_TRAVERSAL_LR_REVERSE_LEVEL_ORDER)) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return;
  }
  bfInitPtrArrayWithDefaultCapacity(&it->nodes);
  bfPtrArrayAppend(&it->nodes, node);
  for (BfSize i = 0; i < it->nodes.num_elts; ++i) {
    BfQuadtreeNode const *v = it->nodes.data[i];
    for (int q = 0; q < 4; ++q)
      if (v->child[q] >= 0) bfPtrArrayAppend(&it->nodes, (void *)&t->nodes[v->child[q]]);
  }
  LevelIterAux *a = calloc(1, sizeof(LevelIterAux));
  const BfQuadtreeNode *first = it->nodes.data[0], *last = it->nodes.data[it->nodes.num_elts - 1];
  a->numLevels = (BfSize)(last->depth - first->depth + 1);
  a->offsets = calloc(a->numLevels + 1, sizeof(BfSize));
  BfSize lv = 1;
  for (BfSize i = 1; i < it->nodes.num_elts; ++i)
    if (((BfQuadtreeNode const *)it->nodes.data[i])->depth != ((BfQuadtreeNode const *)it->nodes.data[i - 1])->depth)
      a->offsets[lv++] = i;
  a->offsets[a->numLevels] = it->nodes.num_elts;
  a->current = it->traversal == BF_TREE_TRAVERSAL_LR_LEVEL_ORDER ? 0 : a->numLevels - 1;
  it->aux = a;
  level_view(it);
}

void bfTreeLevelIterDeinit(BfTreeLevelIter *it) {
  if (!it) return;
  bfPtrArrayDeinit(&it->nodes);
  LevelIterAux *a = it->aux;
  if (a) {
    free(a->offsets);
    free(a);
  }
  it->aux = NULL;
  it->levelNodes = NULL;
}

BfSize bfTreeLevelIterCurrentDepth(BfTreeLevelIter const *it) {
  if (!it || !it->levelNodes || it->levelNodes->num_elts == 0) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return BF_SIZE_BAD_VALUE;
  }
  return (BfSize)((BfQuadtreeNode const *)it->levelNodes->data[0])->depth;
}

bool bfTreeLevelIterIsDone(BfTreeLevelIter const *it) {
  LevelIterAux const *a = it->aux;
  return a->current == a->numLevels;
}

void bfTreeLevelIterNext(BfTreeLevelIter *it) {
  LevelIterAux *a = it->aux;
  if (it->traversal == BF_TREE_TRAVERSAL_LR_LEVEL_ORDER) {
    if (a->current < a->numLevels) ++a->current;
  } else if (a->current > 0) {
    --a->current;
  }
  if (a->current < a->numLevels) level_view(it);
}

BfSize bfTreeLevelIterGetNumPoints(BfTreeLevelIter const *it) {
  BfSize n = 0;
  for (BfSize i = 0; i < it->levelNodes->num_elts; ++i)
    n += bfQuadtreeNodeGetNumPoints(it->levelNodes->data[i]);
  return n;
}

bool bfTreeLevelIterCurrentLevelIsInternal(BfTreeLevelIter const *it) {
  for (BfSize i = 0; i < it->levelNodes->num_elts; ++i) {
    BfQuadtreeNode const *v = it->levelNodes->data[i];
    if (v->child[0] < 0 && v->child[1] < 0 && v->child[2] < 0 && v->child[3] < 0) return false;
  }
  return true;
}

BfPtrArray *bfTreeLevelIterGetLevelNodes(BfTreeLevelIter *it) { return it ? it->levelNodes : NULL; }
