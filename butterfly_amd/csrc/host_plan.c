/*
 * Planner of the multilevel butterfly (fac_helm2.c:42-1002), producing
 * flat tables: top-level leaves (dense blocks or butterflies), factors,
 * blocks, and the sampling / re-expansion jobs the GPU runs.  The plan is
 * integer-exact with the reference's decisions in its default mode; the
 * pass-through mode adds the config-driven planner.
 */
#define _DEFAULT_SOURCE 1
#include <math.h>
#include <pthread.h>
#include <stdatomic.h>
#include <unistd.h>

#include "bf_host.h"

enum { MODE_PROXY = 0, MODE_CHARGES = 1, MODE_VALUES = 2 };

typedef struct {
  BfQuadtree const *t;
  double *cx, *cy, *cr; /* bounding circles per node */
} TreeView;

/* one top-level block of the plan: a dense leaf or a separated pair */
typedef struct {
  int32_t s, t, separated;
} PlanItem;
typedef BF_VEC(PlanItem) PlanItemVec;

typedef struct {
  BfMatHipButterfly *F;
  TreeView src, tgt;
  double k, tol, C;
  PlanItemVec *items; /* when set, the recursion only lists the top-level blocks */
} Planner;

typedef BF_VEC(int32_t) Level;

static void tree_view(TreeView *v, BfQuadtree const *t) {
  v->t = t;
  v->cx = malloc(t->numNodes * sizeof(double));
  v->cy = malloc(t->numNodes * sizeof(double));
  v->cr = malloc(t->numNodes * sizeof(double));
  for (size_t i = 0; i < t->numNodes; ++i) {
    BfCircle c = bfQuadtreeNodeGetBoundingCircle(&t->nodes[i]);
    v->cx[i] = c.center[0];
    v->cy[i] = c.center[1];
    v->cr[i] = c.r;
  }
}

static void tree_view_free(TreeView *v) {
  free(v->cx);
  free(v->cy);
  free(v->cr);
}

static inline const BfQuadtreeNode *node(const TreeView *v, int32_t i) { return &v->t->nodes[i]; }
static inline int64_t npts(const TreeView *v, int32_t i) {
  return (int64_t)(v->t->nodes[i].offset[4] - v->t->nodes[i].offset[0]);
}
static inline int is_leaf(const TreeView *v, int32_t i) {
  const BfQuadtreeNode *n = node(v, i);
  return n->child[0] < 0 && n->child[1] < 0 && n->child[2] < 0 && n->child[3] < 0;
}

/* p = ceil(k r1 r2 / d - C log10(tol)) with (r1, circle 1) first, in the
 * argument order of each reference call site (helm2.c:14-36) */
static int64_t rank_est(const Planner *P, const TreeView *a, int32_t ia, const TreeView *b, int32_t ib) {
  double r1 = a->cr[ia], r2 = b->cr[ib];
  double R = hypot(b->cx[ib] - a->cx[ia], b->cy[ib] - a->cy[ia]);
  double d = R - r1 - r2;
  double p = P->k * r1 * r2 / d - P->C * log10(P->tol);
  return (int64_t)ceil(p);
}

static int separated(const TreeView *a, int32_t ia, const TreeView *b, int32_t ib) {
  double R = hypot(b->cx[ib] - a->cx[ia], b->cy[ib] - a->cy[ia]);
  return R > a->cr[ia] + b->cr[ib] + 1e1 * BF_EPS_MACH;
}

static int children(const TreeView *v, int32_t i, int32_t out[4], int selfIfLeaf) {
  int n = 0;
  for (int q = 0; q < 4; ++q)
    if (node(v, i)->child[q] >= 0) out[n++] = node(v, i)->child[q];
  if (n == 0 && selfIfLeaf) out[n++] = i;
  return n;
}

static int64_t alloc_arena(BfMatHipButterfly *F, int64_t entries) {
  int64_t off = F->arenaEntries;
  F->arenaEntries += entries;
  return off;
}

static void push_kernel_job(Planner *P, BfHipBlock *b, const TreeView *sv, int32_t s, int srcIsPoints,
                            int64_t ncols, int32_t t) {
  BfMatHipButterfly *F = P->F;
  BfkKernelJob J;
  memset(&J, 0, sizeof(J));
  J.scx = sv->cx[s];
  J.scy = sv->cy[s];
  J.sr = sv->cr[s];
  J.k = P->k;
  J.srcOff = srcIsPoints ? (int64_t)node(sv, s)->offset[0] : -1;
  J.tgtOff = (int64_t)node(&P->tgt, t)->offset[0];
  J.m = b->m;
  J.n = (int32_t)ncols;
  b->job = (int32_t)F->kernelJobs.n;
  b->dataOff = alloc_arena(F, (int64_t)b->m * b->n);
  J.out = b->dataOff;
  BF_VEC_PUSH(F->kernelJobs, J);
}

/* re-expansion from `orig` (points of source node o when origIsPoints,
 * else n proxies on its circle) to m proxies on the circle of source node
 * e, tested on m proxies on the circle of target node tt */
static void push_lstsq_job(Planner *P, BfHipBlock *b, int32_t e, int32_t o, int origIsPoints, int32_t tt) {
  BfMatHipButterfly *F = P->F;
  BfkLstsqJob J;
  memset(&J, 0, sizeof(J));
  J.ecx = P->src.cx[e]; J.ecy = P->src.cy[e]; J.er = P->src.cr[e];
  J.tcx = P->tgt.cx[tt]; J.tcy = P->tgt.cy[tt]; J.tr = P->tgt.cr[tt];
  J.ocx = P->src.cx[o]; J.ocy = P->src.cy[o]; J.orr = P->src.cr[o];
  J.k = P->k;
  J.rcond = F->opts.pinvRcond;
  J.origOff = origIsPoints ? (int64_t)node(&P->src, o)->offset[0] : -1;
  J.p = b->m;
  J.n = b->n;
  b->job = (int32_t)F->lstsqJobs.n;
  b->dataOff = alloc_arena(F, (int64_t)b->m * b->n);
  J.out = b->dataOff;
  BF_VEC_PUSH(F->lstsqJobs, J);
}

static int64_t push_offsets(BfMatHipButterfly *F, const int64_t *sizes, int64_t n) {
  int64_t start = (int64_t)F->offs.n;
  int64_t acc = 0;
  BF_VEC_PUSH(F->offs, 0);
  for (int64_t i = 0; i < n; ++i) {
    acc += sizes[i];
    BF_VEC_PUSH(F->offs, acc);
  }
  return start;
}

/* Butterfly between S and T from level lists Ls[0..D] (Ls[0] = {S}) and
 * Lt[0..D] (Lt[0] = {T}): the first factor is block diagonal over Ls[D]
 * (makeFirstFactor, fac_helm2.c:42-160), factor i = 1..D is block-sparse
 * with one block per (target child, source child) (makeFactor, :222-401),
 * the last is block diagonal over Lt[D] (makeLastFactor, :403-509).
 * With `modes` a pair may carry the source charges themselves or the
 * target values themselves instead of proxies, whichever the estimate
 * says is smaller, so leaves carried below their depth cost nothing. */
static void make_butterfly(Planner *P, BfHipLeaf *leaf, Level *Ls, Level *Lt, int D, int modes) {
  BfMatHipButterfly *F = P->F;
  const TreeView *sv = &P->src, *tv = &P->tgt;
  int32_t S = Ls[0].a[0], T = Lt[0].a[0];
  leaf->numFactors = D + 2;
  leaf->firstFactor = (int64_t)F->factors.n;

  /* first factor */
  int64_t nb = (int64_t)Ls[D].n;
  int8_t *mode = malloc(nb);
  int64_t *msize = malloc(nb * sizeof(int64_t));
  int64_t *cs = malloc(nb * sizeof(int64_t));
  BfHipFactor f0;
  memset(&f0, 0, sizeof(f0));
  f0.isDiag = 1;
  f0.numBlockRows = f0.numBlockCols = f0.numBlocks = nb;
  f0.firstBlock = (int64_t)F->blocks.n;
  for (int64_t i = 0; i < nb; ++i) {
    int32_t s = Ls[D].a[i];
    int64_t p = rank_est(P, sv, s, tv, T);
    int64_t ns = npts(sv, s);
    BfHipBlock b;
    memset(&b, 0, sizeof(b));
    b.rowInd = b.colInd = (int32_t)i;
    b.n = (int32_t)ns;
    if (modes && is_leaf(sv, s) && p >= ns) {
      mode[i] = MODE_CHARGES;
      b.m = (int32_t)ns;
      b.kind = BF_BLOCK_IDENTITY;
      b.dataOff = -1;
      b.job = -1;
    } else {
      mode[i] = MODE_PROXY;
      b.m = (int32_t)p;
      b.kind = BF_BLOCK_LSTSQ;
      push_lstsq_job(P, &b, s, s, 1, T);
    }
    msize[i] = b.m;
    cs[i] = b.n;
    BF_VEC_PUSH(F->blocks, b);
  }
  f0.rowOffStart = push_offsets(F, msize, nb);
  f0.colOffStart = push_offsets(F, cs, nb);
  BF_VEC_PUSH(F->factors, f0);
  free(cs);

  /* middle factors */
  for (int st = 1; st <= D; ++st) {
    Level *sPar = &Ls[D - st], *sCh = &Ls[D - st + 1];
    Level *tPar = &Lt[st - 1], *tCh = &Lt[st];
    int64_t nRows = (int64_t)tCh->n * (int64_t)sPar->n;
    int64_t nCols = (int64_t)tPar->n * (int64_t)sCh->n;
    BfHipFactor f;
    memset(&f, 0, sizeof(f));
    f.isDiag = 0;
    f.numBlockRows = nRows;
    f.numBlockCols = nCols;
    f.numBlocks = (int64_t)tCh->n * (int64_t)sCh->n;
    f.firstBlock = (int64_t)F->blocks.n;
    int64_t *rowSize = calloc(nRows, sizeof(int64_t));
    int8_t *newMode = malloc(nRows);
    int64_t tcIdx = 0;
    for (int64_t tp = 0; tp < (int64_t)tPar->n; ++tp) {
      int32_t tkids[4];
      int ntk = children(tv, tPar->a[tp], tkids, 1);
      for (int a = 0; a < ntk; ++a, ++tcIdx) {
        int32_t tc = tkids[a];
        int64_t scIdx = 0;
        for (int64_t sp = 0; sp < (int64_t)sPar->n; ++sp) {
          int32_t spn = sPar->a[sp];
          int32_t skids[4];
          int nsk = children(sv, spn, skids, 1);
          int64_t row = tcIdx * (int64_t)sPar->n + sp;
          int64_t est = 0;
          int anyValues = 0, allCharges = 1;
          for (int c = 0; c < nsk; ++c) {
            int64_t r1 = rank_est(P, sv, skids[c], tv, tPar->a[tp]);
            int64_t r2 = rank_est(P, sv, spn, tv, tc);
            int64_t r = r1 > r2 ? r1 : r2;
            if (r > est) est = r;
            int64_t col = tp * (int64_t)sCh->n + scIdx + c;
            if (mode[col] == MODE_VALUES) anyValues = 1;
            if (mode[col] != MODE_CHARGES) allCharges = 0;
          }
          int rm = MODE_PROXY;
          int64_t rows = est;
          if (modes) {
            if (anyValues || (is_leaf(tv, tc) && est >= npts(tv, tc))) {
              rm = MODE_VALUES;
              rows = npts(tv, tc);
            } else if (is_leaf(sv, spn) && allCharges && est >= npts(sv, spn)) {
              rm = MODE_CHARGES;
              rows = npts(sv, spn);
            }
          }
          rowSize[row] = rows;
          newMode[row] = (int8_t)rm;
          for (int c = 0; c < nsk; ++c) {
            int64_t col = tp * (int64_t)sCh->n + scIdx + c;
            int cm = mode[col];
            BfHipBlock b;
            memset(&b, 0, sizeof(b));
            b.rowInd = (int32_t)row;
            b.colInd = (int32_t)col;
            b.m = (int32_t)rows;
            b.n = (int32_t)msize[col];
            if ((cm == MODE_CHARGES && rm == MODE_CHARGES) || (cm == MODE_VALUES && rm == MODE_VALUES)) {
              b.kind = BF_BLOCK_IDENTITY;
              b.dataOff = -1;
              b.job = -1;
            } else if (rm == MODE_VALUES) {
              b.kind = BF_BLOCK_KERNEL;
              push_kernel_job(P, &b, sv, skids[c], cm == MODE_CHARGES, b.n, tc);
            } else {
              b.kind = BF_BLOCK_LSTSQ;
              push_lstsq_job(P, &b, spn, skids[c], cm == MODE_CHARGES, tc);
            }
            BF_VEC_PUSH(F->blocks, b);
          }
          scIdx += nsk;
        }
      }
    }
    f.rowOffStart = push_offsets(F, rowSize, nRows);
    f.colOffStart = push_offsets(F, msize, nCols);
    BF_VEC_PUSH(F->factors, f);
    free(mode);
    free(msize);
    mode = newMode;
    msize = rowSize;
  }

  /* last factor */
  nb = (int64_t)Lt[D].n;
  BfHipFactor fl;
  memset(&fl, 0, sizeof(fl));
  fl.isDiag = 1;
  fl.numBlockRows = fl.numBlockCols = fl.numBlocks = nb;
  fl.firstBlock = (int64_t)F->blocks.n;
  int64_t *rs = malloc(nb * sizeof(int64_t));
  for (int64_t i = 0; i < nb; ++i) {
    int32_t t = Lt[D].a[i];
    BfHipBlock b;
    memset(&b, 0, sizeof(b));
    b.rowInd = b.colInd = (int32_t)i;
    b.m = (int32_t)npts(tv, t);
    b.n = (int32_t)msize[i];
    if (mode[i] == MODE_VALUES) {
      b.kind = BF_BLOCK_IDENTITY;
      b.dataOff = -1;
      b.job = -1;
    } else {
      b.kind = BF_BLOCK_KERNEL;
      push_kernel_job(P, &b, sv, S, mode[i] == MODE_CHARGES, b.n, t);
    }
    rs[i] = b.m;
    BF_VEC_PUSH(F->blocks, b);
  }
  fl.rowOffStart = push_offsets(F, rs, nb);
  fl.colOffStart = push_offsets(F, msize, nb);
  BF_VEC_PUSH(F->factors, fl);
  free(rs);
  free(mode);
  free(msize);
}

static BfHipLeaf *new_leaf(BfMatHipButterfly *F, const Planner *P, int32_t s, int32_t t, int isButterfly) {
  BfHipLeaf l;
  memset(&l, 0, sizeof(l));
  l.isButterfly = isButterfly;
  l.row0 = (int64_t)node(&P->tgt, t)->offset[0];
  l.col0 = (int64_t)node(&P->src, s)->offset[0];
  l.m = npts(&P->tgt, t);
  l.n = npts(&P->src, s);
  l.denseOff = -1;
  BF_VEC_PUSH(F->leaves, l);
  return &F->leaves.a[F->leaves.n - 1];
}

static void make_dense(Planner *P, int32_t s, int32_t t) {
  BfHipLeaf *l = new_leaf(P->F, P, s, t, 0);
  BfHipBlock b;
  memset(&b, 0, sizeof(b));
  b.m = (int32_t)l->m;
  b.n = (int32_t)l->n;
  push_kernel_job(P, &b, &P->src, s, 1, l->n, t);
  P->F->leaves.a[P->F->leaves.n - 1].denseOff = b.dataOff;
}

/* breadth-first levels of the subtree under `root` */
static int bfs_levels(const TreeView *v, int32_t root, Level **out) {
  int cap = 8, n = 0;
  Level *lv = calloc(cap, sizeof(Level));
  BF_VEC_PUSH(lv[0], root);
  n = 1;
  for (;;) {
    Level nxt;
    memset(&nxt, 0, sizeof(nxt));
    Level *cur = &lv[n - 1];
    for (size_t i = 0; i < cur->n; ++i) {
      int32_t kids[4];
      int nk = children(v, cur->a[i], kids, 0);
      for (int c = 0; c < nk; ++c) BF_VEC_PUSH(nxt, kids[c]);
    }
    if (nxt.n == 0) break;
    if (n == cap) {
      lv = realloc(lv, 2 * cap * sizeof(Level));
      memset(lv + cap, 0, cap * sizeof(Level));
      cap *= 2;
    }
    lv[n++] = nxt;
  }
  *out = lv;
  return n;
}

static int64_t level_points(const TreeView *v, const Level *l) {
  int64_t s = 0;
  for (size_t i = 0; i < l->n; ++i) s += npts(v, l->a[i]);
  return s;
}

static int level_internal(const TreeView *v, const Level *l) {
  for (size_t i = 0; i < l->n; ++i)
    if (is_leaf(v, l->a[i])) return 0;
  return 1;
}

/* allRankEstimatesAreOK (fac_helm2.c:511-530) */
static int all_ranks_ok(const Planner *P, int32_t T, const Level *l) {
  for (size_t i = 0; i < l->n; ++i)
    if (rank_est(P, &P->tgt, T, &P->src, l->a[i]) > npts(&P->src, l->a[i])) return 0;
  return 1;
}

static void free_levels(Level *lv, int n) {
  for (int i = 0; i < n; ++i) BF_VEC_FREE(lv[i]);
  free(lv);
}

static void make_separated(Planner *P, int32_t s, int32_t t) {
  const TreeView *sv = &P->src, *tv = &P->tgt;
  if (P->F->opts.passThrough) {
    int ds = node(sv, s)->maxDepthBelow - node(sv, s)->depth;
    int dt = node(tv, t)->maxDepthBelow - node(tv, t)->depth;
    int D = ds < dt ? ds : dt;
    Level *Ls = calloc(D + 1, sizeof(Level)), *Lt = calloc(D + 1, sizeof(Level));
    BF_VEC_PUSH(Ls[0], s);
    BF_VEC_PUSH(Lt[0], t);
    for (int j = 0; j < D; ++j) {
      int32_t kids[4];
      for (size_t i = 0; i < Ls[j].n; ++i) {
        int nk = children(sv, Ls[j].a[i], kids, 1);
        for (int c = 0; c < nk; ++c) BF_VEC_PUSH(Ls[j + 1], kids[c]);
      }
      for (size_t i = 0; i < Lt[j].n; ++i) {
        int nk = children(tv, Lt[j].a[i], kids, 1);
        for (int c = 0; c < nk; ++c) BF_VEC_PUSH(Lt[j + 1], kids[c]);
      }
    }
    BfMatHipButterfly *F = P->F;
    const size_t nl = F->leaves.n, nfac = F->factors.n, nb = F->blocks.n, no = F->offs.n;
    const size_t njl = F->lstsqJobs.n, njk = F->kernelJobs.n;
    const int64_t na = F->arenaEntries;
    BfHipLeaf *l = new_leaf(F, P, s, t, 1);
    make_butterfly(P, l, Ls, Lt, D, 1);
    free_levels(Ls, D + 1);
    free_levels(Lt, D + 1);
    /* a butterfly that stores more entries than the block itself (a small
     * target facing a large source across a narrow gap, where the proxy
     * count k r1 r2 / d exceeds the points) is replaced by the dense block */
    if (F->arenaEntries - na > npts(sv, s) * npts(tv, t)) {
      F->leaves.n = nl;
      F->factors.n = nfac;
      F->blocks.n = nb;
      F->offs.n = no;
      F->lstsqJobs.n = njl;
      F->kernelJobs.n = njk;
      F->arenaEntries = na;
      make_dense(P, s, t);
    }
    return;
  }

  /* bfFacHelm2Prepare (fac_helm2.c:551-651) */
  Level *sl, *tl;
  int nsl = bfs_levels(sv, s, &sl), ntl = bfs_levels(tv, t, &tl);
  int tgtDepth = node(tv, t)->depth;
  int maxAllow = tgtDepth;
  for (int j = 1; j < ntl && level_internal(tv, &tl[j]); ++j) ++maxAllow;
  int li = nsl - 1;
  int srcDepth = node(sv, s)->depth + li;
  while (srcDepth > maxAllow && li > 0) { --li; --srcDepth; }
  while (level_points(sv, &sl[li]) != npts(sv, s) && li > 0) { --li; --srcDepth; }
  while (!level_internal(sv, &sl[li]) && li > 0) { --li; --srcDepth; }
  while (srcDepth > tgtDepth && !all_ranks_ok(P, t, &sl[li]) && li > 0) { --li; --srcDepth; }
  int nf = all_ranks_ok(P, t, &sl[li]) ? srcDepth - tgtDepth + 2 : 0;
  if (nf == 0) {
    make_dense(P, s, t);
  } else {
    BfHipLeaf *l = new_leaf(P->F, P, s, t, 1);
    make_butterfly(P, l, sl, tl, nf - 2, 0);
  }
  free_levels(sl, nsl);
  free_levels(tl, ntl);
}

static void emit(Planner *P, int32_t s, int32_t t, int sep) {
  if (P->items) {
    PlanItem it = {s, t, sep};
    BF_VEC_PUSH(*P->items, it);
  } else if (sep) {
    make_separated(P, s, t);
  } else {
    make_dense(P, s, t);
  }
}

/* facHelm2MakeMultilevel_rec (fac_helm2.c:860-941) */
static void multilevel_rec(Planner *P, const int32_t *sNodes, int ns, const int32_t *tNodes, int nt) {
  const TreeView *sv = &P->src, *tv = &P->tgt;
  for (int i = 0; i < nt; ++i) {
    for (int j = 0; j < ns; ++j) {
      int32_t s = sNodes[j], t = tNodes[i];
      if ((BfSize)(npts(tv, t) * npts(sv, s)) < P->F->opts.maxDenseSize) {
        emit(P, s, t, 0);
      } else if (separated(sv, s, tv, t)) {
        emit(P, s, t, 1);
      } else {
        /* a leaf stands in for its own child list; two leaves that are
         * still not separated become one dense block */
        int32_t sc[4], tc[4];
        int nsc = children(sv, s, sc, 1), ntc = children(tv, t, tc, 1);
        if (nsc == 1 && ntc == 1 && sc[0] == s && tc[0] == t) {
          emit(P, s, t, 0);
          continue;
        }
        multilevel_rec(P, sc, nsc, tc, ntc);
      }
    }
  }
}

/* ---- parallel build: the top-level blocks are independent, so they are
 * built by a pool of threads into per-chunk tables (chunks of consecutive
 * blocks, handed out in order), which are then concatenated in block order
 * with their indices and arena offsets shifted.  The result is the same
 * plan, entry for entry, as building the blocks one after another. */

typedef struct {
  BfMatHipButterfly T; /* local tables: leaves, factors, blocks, offs, jobs, arena size */
  size_t i0, i1;       /* items */
  /* bases in the merged tables */
  int64_t leafBase, factorBase, blockBase, offsBase, lBase, kBase, arenaBase;
} PlanChunk;

typedef struct {
  const Planner *P;
  const PlanItemVec *items;
  PlanChunk *chunks;
  size_t numChunks;
  BfMatHipButterfly *F;
  int64_t srcPts, tgtPts; /* point offsets of this problem in F's stacked points */
  atomic_size_t next;
  int phase; /* 0 build, 1 merge */
} PlanPool;

static void build_chunk(PlanPool *pool, PlanChunk *c) {
  Planner L = *pool->P;
  L.items = NULL;
  L.F = &c->T;
  c->T.helm = pool->F->helm;
  c->T.opts = pool->F->opts;
  for (size_t i = c->i0; i < c->i1; ++i) emit(&L, pool->items->a[i].s, pool->items->a[i].t, pool->items->a[i].separated);
}

#define COPY_TABLE(dst, src, base) memcpy((dst).a + (base), (src).a, (src).n * sizeof(*(src).a))

/* copies T's tables into F at the given bases (F's tables already sized),
 * shifting indices and arena offsets, and point offsets by the source /
 * target point bases; then frees T's tables */
static void copy_tables(BfMatHipButterfly *F, BfMatHipButterfly *T, const PlanChunk *c, int64_t srcPts,
                        int64_t tgtPts) {
  /* one pass per table: each entry is copied and shifted on the way (the
   * job and block tables run to gigabytes at 2^20 points) */
  BfHipLeaf *L = F->leaves.a + c->leafBase;
  for (size_t i = 0; i < T->leaves.n; ++i) {
    BfHipLeaf l = T->leaves.a[i];
    l.firstFactor += c->factorBase;
    l.row0 += tgtPts;
    l.col0 += srcPts;
    if (l.denseOff >= 0) l.denseOff += c->arenaBase;
    L[i] = l;
  }
  BfHipFactor *Fa = F->factors.a + c->factorBase;
  for (size_t i = 0; i < T->factors.n; ++i) {
    BfHipFactor f = T->factors.a[i];
    f.firstBlock += c->blockBase;
    f.rowOffStart += c->offsBase;
    f.colOffStart += c->offsBase;
    Fa[i] = f;
  }
  BfHipBlock *B = F->blocks.a + c->blockBase;
  for (size_t i = 0; i < T->blocks.n; ++i) {
    BfHipBlock b = T->blocks.a[i];
    if (b.dataOff >= 0) b.dataOff += c->arenaBase;
    if (b.job >= 0) b.job += (int32_t)(b.kind == BF_BLOCK_LSTSQ ? c->lBase : c->kBase);
    B[i] = b;
  }
  COPY_TABLE(F->offs, T->offs, c->offsBase);
  BfkLstsqJob *JL = F->lstsqJobs.a + c->lBase;
  for (size_t i = 0; i < T->lstsqJobs.n; ++i) {
    BfkLstsqJob J = T->lstsqJobs.a[i];
    J.out += c->arenaBase;
    if (J.origOff >= 0) J.origOff += srcPts;
    JL[i] = J;
  }
  BfkKernelJob *JK = F->kernelJobs.a + c->kBase;
  for (size_t i = 0; i < T->kernelJobs.n; ++i) {
    BfkKernelJob J = T->kernelJobs.a[i];
    J.out += c->arenaBase;
    if (J.srcOff >= 0) J.srcOff += srcPts;
    J.tgtOff += tgtPts;
    JK[i] = J;
  }
  BF_VEC_FREE(T->leaves);
  BF_VEC_FREE(T->factors);
  BF_VEC_FREE(T->blocks);
  BF_VEC_FREE(T->offs);
  BF_VEC_FREE(T->lstsqJobs);
  BF_VEC_FREE(T->kernelJobs);
}

static void merge_chunk(PlanPool *pool, PlanChunk *c) { copy_tables(pool->F, &c->T, c, pool->srcPts, pool->tgtPts); }

static void *plan_worker(void *arg) {
  PlanPool *pool = arg;
  for (;;) {
    size_t c = atomic_fetch_add(&pool->next, 1);
    if (c >= pool->numChunks) break;
    if (pool->phase == 0)
      build_chunk(pool, &pool->chunks[c]);
    else
      merge_chunk(pool, &pool->chunks[c]);
  }
  return NULL;
}

static void run_pool(PlanPool *pool, int phase, int nthreads) {
  pool->phase = phase;
  atomic_store(&pool->next, 0);
  pthread_t th[64];
  int started = 0;
  for (int i = 1; i < nthreads; ++i)
    if (pthread_create(&th[started], NULL, plan_worker, pool) == 0) ++started;
  plan_worker(pool);
  for (int i = 0; i < started; ++i) pthread_join(th[i], NULL);
}

/* size a table to `total` entries, growing its capacity geometrically so
 * that repeated appends copy each entry a bounded number of times; a first
 * allocation of a large table gets huge-page advice (not a regrown one:
 * appended problems of a batch would then fault into compaction stalls) */
#define RESERVE(v, total)                                              \
  do {                                                                 \
    if ((v).cap < (size_t)(total)) {                                   \
      size_t nc_ = 2 * (v).cap > (size_t)(total) ? 2 * (v).cap : (size_t)(total); \
      void *na_ = realloc((v).a, (nc_ ? nc_ : 1) * sizeof(*(v).a));    \
      if (!na_) abort();                                               \
      if (!(v).cap) bfHugePages(na_, nc_ * sizeof(*(v).a));            \
      (v).a = na_;                                                     \
      (v).cap = nc_;                                                   \
    }                                                                  \
    (v).n = (size_t)(total);                                           \
  } while (0)

static int plan_threads(void) {
  long n = sysconf(_SC_NPROCESSORS_ONLN);
  if (n < 1) n = 1;
  if (n > 16) n = 16;
  return (int)n;
}

int bfPlanBuild(BfMatHipButterfly *F, BfQuadtree const *src, BfQuadtree const *tgt) {
  return bfPlanBuildAt(F, src, tgt, 0, 0);
}

/* Plans one problem (F->helm, F->opts) and appends it to F's tables as the
 * next diagonal block: leaf rows / columns and point offsets shifted by
 * tgtPts / srcPts, indices and arena offsets after what F already holds. */
int bfPlanBuildAt(BfMatHipButterfly *F, BfQuadtree const *src, BfQuadtree const *tgt, int64_t srcPts,
                  int64_t tgtPts) {
  Planner P;
  memset(&P, 0, sizeof(P));
  P.F = F;
  P.k = F->helm.k;
  P.tol = F->opts.tol;
  P.C = F->opts.rankConst;
  tree_view(&P.src, src);
  tree_view(&P.tgt, tgt);
  if (src->numLevels < 3 || tgt->numLevels < 3) {
    tree_view_free(&P.src);
    tree_view_free(&P.tgt);
    return -1;
  }
  /* start from the second level of both trees (fac_helm2.c:954-982) */
  int32_t s2[16], t2[16];
  int n2s = 0, n2t = 0;
  for (BfSize i = src->levelOffset[2]; i < src->levelOffset[3]; ++i) s2[n2s++] = (int32_t)i;
  for (BfSize i = tgt->levelOffset[2]; i < tgt->levelOffset[3]; ++i) t2[n2t++] = (int32_t)i;
  /* list the top-level blocks in plan order, then build them */
  PlanItemVec items;
  memset(&items, 0, sizeof(items));
  P.items = &items;
  multilevel_rec(&P, s2, n2s, t2, n2t);
  P.items = NULL;
  /* chunks of about equal point count, so one huge block does not sit
   * in a chunk with many others */
  const int nthreads = plan_threads();
  double totalW = 0;
  for (size_t i = 0; i < items.n; ++i)
    totalW += (double)(npts(&P.src, items.a[i].s) + npts(&P.tgt, items.a[i].t));
  const double target = totalW / (double)(64 * nthreads) + 1;
  size_t nc = 0;
  PlanChunk *chunks = calloc(items.n ? items.n : 1, sizeof(PlanChunk));
  for (size_t i = 0; i < items.n;) {
    double w = 0;
    size_t j = i;
    while (j < items.n && (j == i || w < target)) {
      w += (double)(npts(&P.src, items.a[j].s) + npts(&P.tgt, items.a[j].t));
      ++j;
    }
    chunks[nc].i0 = i;
    chunks[nc].i1 = j;
    ++nc;
    i = j;
  }
  PlanPool pool;
  memset(&pool, 0, sizeof(pool));
  pool.P = &P;
  pool.items = &items;
  pool.chunks = chunks;
  pool.numChunks = nc;
  pool.F = F;
  pool.srcPts = srcPts;
  pool.tgtPts = tgtPts;
  const double tp = bfWallTime();
  run_pool(&pool, 0, nthreads);
  bfTrace("plan: chunks planned", tp);
  int64_t nl = (int64_t)F->leaves.n, nf = (int64_t)F->factors.n, nb = (int64_t)F->blocks.n;
  int64_t no = (int64_t)F->offs.n, njl = (int64_t)F->lstsqJobs.n, njk = (int64_t)F->kernelJobs.n;
  int64_t na = F->arenaEntries;
  for (size_t c = 0; c < nc; ++c) {
    BfMatHipButterfly *T = &chunks[c].T;
    chunks[c].leafBase = nl;
    chunks[c].factorBase = nf;
    chunks[c].blockBase = nb;
    chunks[c].offsBase = no;
    chunks[c].lBase = njl;
    chunks[c].kBase = njk;
    chunks[c].arenaBase = na;
    nl += (int64_t)T->leaves.n;
    nf += (int64_t)T->factors.n;
    nb += (int64_t)T->blocks.n;
    no += (int64_t)T->offs.n;
    njl += (int64_t)T->lstsqJobs.n;
    njk += (int64_t)T->kernelJobs.n;
    na += T->arenaEntries;
  }
  RESERVE(F->leaves, nl);
  RESERVE(F->factors, nf);
  RESERVE(F->blocks, nb);
  RESERVE(F->offs, no);
  RESERVE(F->lstsqJobs, njl);
  RESERVE(F->kernelJobs, njk);
  F->arenaEntries = na;
  run_pool(&pool, 1, nthreads);
  bfTrace("plan: chunks merged", tp);
  free(chunks);
  BF_VEC_FREE(items);
  tree_view_free(&P.src);
  tree_view_free(&P.tgt);
  return 0;
}

/* One butterfly between source node S = ls[0][0] and target node
 * T = lt[0][0] on the given levels (bfFacHelm2Make, fac_helm2.c:653-704):
 * ls[j] are the source nodes j levels below S, lt[j] the target nodes j
 * levels below T, j = 0..D, all as node indices of their trees.  The plan
 * has this one leaf, at the node's rows and columns of the whole trees,
 * and uses F's wavenumber and options. */
int bfPlanButterfly(BfMatHipButterfly *F, BfQuadtree const *src, BfQuadtree const *tgt, const int32_t *const *ls,
                    const size_t *nls, const int32_t *const *lt, const size_t *nlt, int D) {
  if (D < 0 || nls[0] != 1 || nlt[0] != 1) return -1;
  Planner P;
  memset(&P, 0, sizeof(P));
  P.F = F;
  P.k = F->helm.k;
  P.tol = F->opts.tol;
  P.C = F->opts.rankConst;
  tree_view(&P.src, src);
  tree_view(&P.tgt, tgt);
  Level *Ls = calloc((size_t)D + 1, sizeof(Level)), *Lt = calloc((size_t)D + 1, sizeof(Level));
  for (int j = 0; j <= D; ++j) {
    for (size_t i = 0; i < nls[j]; ++i) BF_VEC_PUSH(Ls[j], ls[j][i]);
    for (size_t i = 0; i < nlt[j]; ++i) BF_VEC_PUSH(Lt[j], lt[j][i]);
  }
  BfHipLeaf *l = new_leaf(F, &P, Ls[0].a[0], Lt[0].a[0], 1);
  make_butterfly(&P, l, Ls, Lt, D, 0);
  free_levels(Ls, D + 1);
  free_levels(Lt, D + 1);
  tree_view_free(&P.src);
  tree_view_free(&P.tgt);
  return 0;
}
