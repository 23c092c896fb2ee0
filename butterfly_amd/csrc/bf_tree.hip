/*
 * The quadtree on the GPU (gfx950): Morton keys, one radix sort, and the
 * node tables level by level.
 *
 * The host builder (host_base.c) reproduces the reference's recursive
 * in-place sift (quadtree_node.c:123-306): a node holding more than
 * leafSize points is split at the centre of its box into the four
 * quadrants q = 2 (x > sx) + (y > sy), ties going low, boxes halved from
 * the root square.  Which points a node holds is therefore fixed by the
 * geometry alone: the node is the set of points whose first d quadrant
 * choices spell its path.  Here every point computes those choices down
 * to TREE_MAX_DEPTH levels with the same nested-box arithmetic, packed two
 * bits per level into a 62-bit key; a stable radix sort of (key, index)
 * lays the points out in the tree's order, in which every node is one
 * contiguous run of a key prefix; and a loop over depths emits, at depth
 * d, one node per distinct 2d-bit prefix whose parent run holds more than
 * leafSize points.  Nodes of a depth are numbered in key order, which is
 * the breadth-first order of the host builder (children in quadrant
 * order).  With leaf size 1 every leaf holds one point and the
 * permutation is the reference's; with larger leaves the nodes are the
 * same and the points inside a leaf are in key order (index order for
 * equal keys) instead of the sift's order.
 *
 *   k_tree_bbox   per-workgroup min/max of the points, then one workgroup
 *                 over the partials
 *   k_tree_keys   the 62-bit path of every point
 *   k_tree_heads  depth d: 1 where a point starts a prefix run whose parent
 *                 run is split (a node of depth d)
 *   k_tree_emit   depth d: the node record of every such run (box, split,
 *                 child boundaries, parent link; the parent's child slot)
 *
 * Unit of work: 16 B of points read per point per key pass and two 8-B
 * keys per point per depth; the sort is hipCUB's (rocPRIM) radix sort.
 */
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "../../include/bf_amd/bf.h"

#define TR_CHECK(call)                                                                      \
  do {                                                                                      \
    hipError_t e_ = (call);                                                                 \
    if (e_ != hipSuccess) {                                                                 \
      fprintf(stderr, "bf_tree: %s failed: %s\n", #call, hipGetErrorString(e_));            \
      rc = (e_ == hipErrorOutOfMemory || e_ == hipErrorMemoryAllocation) ? -2 : -1;         \
      goto done;                                                                            \
    }                                                                                       \
  } while (0)

extern "C" void *bfk_stream_handle(void);
extern "C" int bfk_init(void);

/* levels below the root a key can hold (2 bits each in 62 bits) */
#define TREE_MAX_DEPTH 31
#define TREE_T 256

static __device__ __forceinline__ uint64_t prefix_of(uint64_t key, int d) {
  return d == 0 ? 0 : key >> (2 * (TREE_MAX_DEPTH - d));
}

/* first index in [lo, hi) whose key is >= v */
static __device__ __forceinline__ int64_t lower_bound(const uint64_t *k, int64_t lo, int64_t hi, uint64_t v) {
  while (lo < hi) {
    const int64_t mid = lo + ((hi - lo) >> 1);
    if (k[mid] < v) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

/* the keys in [lo, hi) whose depth-d prefix is p: [first, last) */
static __device__ __forceinline__ void prefix_run(const uint64_t *k, int64_t n, uint64_t p, int d, int64_t *a,
                                                  int64_t *b) {
  if (d == 0) {
    *a = 0;
    *b = n;
    return;
  }
  const int sh = 2 * (TREE_MAX_DEPTH - d);
  *a = lower_bound(k, 0, n, p << sh);
  *b = (p + 1) << sh == 0 ? n : lower_bound(k, *a, n, (p + 1) << sh);
}

__global__ __launch_bounds__(TREE_T) void k_tree_bbox(int64_t n, const double2 *__restrict__ xy,
                                                      double4 *__restrict__ part) {
  __shared__ double s[4][TREE_T];
  double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
  for (int64_t i = blockIdx.x * (int64_t)TREE_T + threadIdx.x; i < n; i += (int64_t)gridDim.x * TREE_T) {
    const double2 p = xy[i];
    mnx = fmin(mnx, p.x);
    mny = fmin(mny, p.y);
    mxx = fmax(mxx, p.x);
    mxy = fmax(mxy, p.y);
  }
  s[0][threadIdx.x] = mnx;
  s[1][threadIdx.x] = mny;
  s[2][threadIdx.x] = mxx;
  s[3][threadIdx.x] = mxy;
  __syncthreads();
  for (int o = TREE_T / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + o]);
      s[1][threadIdx.x] = fmin(s[1][threadIdx.x], s[1][threadIdx.x + o]);
      s[2][threadIdx.x] = fmax(s[2][threadIdx.x], s[2][threadIdx.x + o]);
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[blockIdx.x] = make_double4(s[0][0], s[1][0], s[2][0], s[3][0]);
}

__global__ __launch_bounds__(TREE_T) void k_tree_bbox_final(int nb, double4 *__restrict__ part) {
  __shared__ double s[4][TREE_T];
  double mnx = INFINITY, mny = INFINITY, mxx = -INFINITY, mxy = -INFINITY;
  for (int i = threadIdx.x; i < nb; i += TREE_T) {
    const double4 p = part[i];
    mnx = fmin(mnx, p.x);
    mny = fmin(mny, p.y);
    mxx = fmax(mxx, p.z);
    mxy = fmax(mxy, p.w);
  }
  s[0][threadIdx.x] = mnx;
  s[1][threadIdx.x] = mny;
  s[2][threadIdx.x] = mxx;
  s[3][threadIdx.x] = mxy;
  __syncthreads();
  for (int o = TREE_T / 2; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) {
      s[0][threadIdx.x] = fmin(s[0][threadIdx.x], s[0][threadIdx.x + o]);
      s[1][threadIdx.x] = fmin(s[1][threadIdx.x], s[1][threadIdx.x + o]);
      s[2][threadIdx.x] = fmax(s[2][threadIdx.x], s[2][threadIdx.x + o]);
      s[3][threadIdx.x] = fmax(s[3][threadIdx.x], s[3][threadIdx.x + o]);
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) part[0] = make_double4(s[0][0], s[1][0], s[2][0], s[3][0]);
}

/* the box of the child in quadrant q of the box (mn, mx), split at its
 * centre: the same operations as the host builder, so the boxes agree
 * bit for bit */
static __device__ __forceinline__ void child_box(double mn[2], double mx[2], int q) {
  const double sx = (mn[0] + mx[0]) / 2, sy = (mn[1] + mx[1]) / 2;
  if (q >> 1) mn[0] = sx;
  else mx[0] = sx;
  if (q & 1) mn[1] = sy;
  else mx[1] = sy;
}

__global__ __launch_bounds__(TREE_T) void k_tree_keys(int64_t n, const double2 *__restrict__ xy, double rx0,
                                                      double ry0, double rx1, double ry1,
                                                      uint64_t *__restrict__ keys, uint64_t *__restrict__ idx) {
  const int64_t i = blockIdx.x * (int64_t)TREE_T + threadIdx.x;
  if (i >= n) return;
  const double2 p = xy[i];
  double mn[2] = {rx0, ry0}, mx[2] = {rx1, ry1};
  uint64_t key = 0;
  for (int d = 0; d < TREE_MAX_DEPTH; ++d) {
    const double sx = (mn[0] + mx[0]) / 2, sy = (mn[1] + mx[1]) / 2;
    const int q = ((p.x > sx) << 1) | (p.y > sy);
    key = key << 2 | (uint64_t)q;
    child_box(mn, mx, q);
  }
  keys[i] = key;
  idx[i] = (uint64_t)i;
}

/* depth d >= 1: flag[i] = 1 if point i starts a run of its depth-d prefix
 * and the run of its depth-(d-1) prefix (the parent) holds more than
 * leafSize points */
__global__ __launch_bounds__(TREE_T) void k_tree_heads(int64_t n, const uint64_t *__restrict__ keys, int d,
                                                       int64_t leafSize, int32_t *__restrict__ flag) {
  const int64_t i = blockIdx.x * (int64_t)TREE_T + threadIdx.x;
  if (i >= n) return;
  const uint64_t p = prefix_of(keys[i], d);
  int f = i == 0 || prefix_of(keys[i - 1], d) != p;
  if (f) {
    int64_t a, b;
    prefix_run(keys, n, p >> 2, d - 1, &a, &b);
    f = b - a > leafSize;
  }
  flag[i] = f;
}

/* depth d: the node of every flagged run.  Level d's nodes get ids
 * base + pos[i] in key order; their run starts go to start[] (sorted), so
 * a child finds its parent among the previous level's starts. */
__global__ __launch_bounds__(TREE_T) void k_tree_emit(int64_t n, const uint64_t *__restrict__ keys, int d,
                                                      int64_t leafSize, const int32_t *__restrict__ flag,
                                                      const int32_t *__restrict__ pos, int64_t base,
                                                      int64_t *__restrict__ start, BfQuadtreeNode *__restrict__ lv,
                                                      const int64_t *__restrict__ parentStart, int64_t parentCount,
                                                      int64_t parentBase, BfQuadtreeNode *__restrict__ parentLv,
                                                      double rx0, double ry0, double rx1, double ry1) {
  const int64_t i = blockIdx.x * (int64_t)TREE_T + threadIdx.x;
  if (i >= n || !flag[i]) return;
  const int64_t local = pos[i];
  const uint64_t p = prefix_of(keys[i], d);
  int64_t a, b;
  prefix_run(keys, n, p, d, &a, &b);
  BfQuadtreeNode nd;
  double mn[2] = {rx0, ry0}, mx[2] = {rx1, ry1};
  for (int l = d - 1; l >= 0; --l) child_box(mn, mx, (int)((p >> (2 * l)) & 3));
  nd.bboxMin[0] = mn[0];
  nd.bboxMin[1] = mn[1];
  nd.bboxMax[0] = mx[0];
  nd.bboxMax[1] = mx[1];
  nd.split[0] = (mn[0] + mx[0]) / 2;
  nd.split[1] = (mn[1] + mx[1]) / 2;
  nd.offset[0] = (BfSize)a;
  nd.offset[4] = (BfSize)b;
  if (b - a > leafSize && d < TREE_MAX_DEPTH) {
    const int sh = 2 * (TREE_MAX_DEPTH - d - 1);
    for (int q = 1; q < 4; ++q) nd.offset[q] = (BfSize)lower_bound(keys, a, b, ((p << 2) | (uint64_t)q) << sh);
  } else {
    for (int q = 1; q < 4; ++q) nd.offset[q] = (BfSize)b;
  }
  for (int q = 0; q < 4; ++q) nd.child[q] = -1;
  nd.depth = d;
  nd.maxDepthBelow = d;
  if (d == 0) {
    nd.parent = -1;
    nd.index = -1;
  } else {
    /* the parent's run starts at the lower bound of the parent prefix */
    int64_t pa, pb;
    prefix_run(keys, n, p >> 2, d - 1, &pa, &pb);
    int64_t lo = 0, hi = parentCount;
    while (lo < hi) {
      const int64_t mid = lo + ((hi - lo) >> 1);
      if (parentStart[mid] < pa) lo = mid + 1;
      else hi = mid;
    }
    nd.parent = (int32_t)(parentBase + lo);
    nd.index = (int32_t)(p & 3);
    parentLv[lo].child[p & 3] = (int32_t)(base + local);
  }
  start[local] = a;
  lv[local] = nd;
}

extern "C" {

/* Builds the quadtree of the n points `xy` (host, x/y interleaved) on the
 * GPU.  rootBox turns the points' bounding box into the root square (the
 * host builder's own arithmetic).  On success *nodesOut (numNodes records,
 * breadth-first) and perm (n entries, tree order) are filled; nodes are
 * malloc'ed.  Returns -3 when the tree would reach depth TREE_MAX_DEPTH
 * (more than leafSize points within 2^-30 of the root box's side). */
int bfk_quadtree(int64_t n, const double *xy, int64_t leafSize,
                 void (*rootBox)(const double mn[2], const double mx[2], double outMin[2], double outMax[2]),
                 BfQuadtreeNode **nodesOut, int64_t *numNodesOut, int64_t *levelCount, int *numLevelsOut,
                 BfSize *perm) {
  if (bfk_init()) return -1;
  hipStream_t st = (hipStream_t)bfk_stream_handle();
  int rc = 0;
  const unsigned grid = (unsigned)((n + TREE_T - 1) / TREE_T);
  const int nbb = 1024;
  double2 *dXY = nullptr;
  double4 *dPart = nullptr;
  uint64_t *kIn = nullptr, *kOut = nullptr, *iIn = nullptr, *iOut = nullptr;
  int32_t *flag = nullptr, *pos = nullptr;
  void *tmp = nullptr;
  size_t tmpBytes = 0, scanBytes = 0, sortBytes = 0;
  int64_t *start[TREE_MAX_DEPTH + 1] = {nullptr};
  BfQuadtreeNode *lv[TREE_MAX_DEPTH + 1] = {nullptr};
  int64_t count[TREE_MAX_DEPTH + 1] = {0}, base[TREE_MAX_DEPTH + 2] = {0};
  int levels = 0;
  double4 bb;
  double mn[2], mx[2], r0[2], r1[2];
  *nodesOut = nullptr;
  *numNodesOut = 0;
  TR_CHECK(hipMalloc((void **)&dXY, (size_t)n * 16));
  TR_CHECK(hipMalloc((void **)&dPart, (size_t)nbb * sizeof(double4)));
  TR_CHECK(hipMalloc((void **)&kIn, (size_t)n * 8));
  TR_CHECK(hipMalloc((void **)&kOut, (size_t)n * 8));
  TR_CHECK(hipMalloc((void **)&iIn, (size_t)n * 8));
  TR_CHECK(hipMalloc((void **)&iOut, (size_t)n * 8));
  TR_CHECK(hipMalloc((void **)&flag, (size_t)n * 4));
  TR_CHECK(hipMalloc((void **)&pos, (size_t)n * 4));
  TR_CHECK(hipMemcpyAsync(dXY, xy, (size_t)n * 16, hipMemcpyHostToDevice, st));
  hipLaunchKernelGGL(k_tree_bbox, dim3(nbb), dim3(TREE_T), 0, st, n, (const double2 *)dXY, dPart);
  TR_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_tree_bbox_final, dim3(1), dim3(TREE_T), 0, st, nbb, dPart);
  TR_CHECK(hipGetLastError());
  TR_CHECK(hipMemcpyAsync(&bb, dPart, sizeof(bb), hipMemcpyDeviceToHost, st));
  TR_CHECK(hipStreamSynchronize(st));
  mn[0] = bb.x;
  mn[1] = bb.y;
  mx[0] = bb.z;
  mx[1] = bb.w;
  rootBox(mn, mx, r0, r1);
  hipLaunchKernelGGL(k_tree_keys, dim3(grid), dim3(TREE_T), 0, st, n, (const double2 *)dXY, r0[0], r0[1], r1[0],
                     r1[1], kIn, iIn);
  TR_CHECK(hipGetLastError());
  TR_CHECK(hipcub::DeviceRadixSort::SortPairs(nullptr, sortBytes, kIn, kOut, iIn, iOut, (int)n, 0,
                                              2 * TREE_MAX_DEPTH, st));
  TR_CHECK(hipcub::DeviceScan::ExclusiveSum(nullptr, scanBytes, flag, pos, (int)n, st));
  tmpBytes = sortBytes > scanBytes ? sortBytes : scanBytes;
  TR_CHECK(hipMalloc(&tmp, tmpBytes));
  TR_CHECK(hipcub::DeviceRadixSort::SortPairs(tmp, sortBytes, kIn, kOut, iIn, iOut, (int)n, 0, 2 * TREE_MAX_DEPTH,
                                              st));
  /* depth 0: the root, the run of the empty prefix */
  for (int d = 0; d <= TREE_MAX_DEPTH; ++d) {
    int64_t m = 1;
    if (d == 0) {
      TR_CHECK(hipMemsetAsync(flag, 0, (size_t)n * 4, st));
      static const int32_t one = 1;
      TR_CHECK(hipMemcpyAsync(flag, &one, 4, hipMemcpyHostToDevice, st));
      TR_CHECK(hipMemsetAsync(pos, 0, (size_t)n * 4, st));
    } else {
      hipLaunchKernelGGL(k_tree_heads, dim3(grid), dim3(TREE_T), 0, st, n, (const uint64_t *)kOut, d, leafSize,
                         flag);
      TR_CHECK(hipGetLastError());
      TR_CHECK(hipcub::DeviceScan::ExclusiveSum(tmp, scanBytes, flag, pos, (int)n, st));
      int32_t lastPos = 0, lastFlag = 0;
      TR_CHECK(hipMemcpyAsync(&lastPos, pos + n - 1, 4, hipMemcpyDeviceToHost, st));
      TR_CHECK(hipMemcpyAsync(&lastFlag, flag + n - 1, 4, hipMemcpyDeviceToHost, st));
      TR_CHECK(hipStreamSynchronize(st));
      m = (int64_t)lastPos + lastFlag;
    }
    if (m == 0) break;
    if (d == TREE_MAX_DEPTH) { /* a run at the last key level would still be split */
      rc = -3;
      goto done;
    }
    count[d] = m;
    base[d + 1] = base[d] + m;
    TR_CHECK(hipMalloc((void **)&start[d], (size_t)m * 8));
    TR_CHECK(hipMalloc((void **)&lv[d], (size_t)m * sizeof(BfQuadtreeNode)));
    hipLaunchKernelGGL(k_tree_emit, dim3(grid), dim3(TREE_T), 0, st, n, (const uint64_t *)kOut, d, leafSize,
                       (const int32_t *)flag, (const int32_t *)pos, base[d], start[d], lv[d],
                       d ? (const int64_t *)start[d - 1] : nullptr, d ? count[d - 1] : 0, d ? base[d - 1] : 0,
                       d ? lv[d - 1] : nullptr, r0[0], r0[1], r1[0], r1[1]);
    TR_CHECK(hipGetLastError());
    levels = d + 1;
  }
  {
    BfQuadtreeNode *nodes = (BfQuadtreeNode *)malloc((size_t)base[levels] * sizeof(BfQuadtreeNode));
    if (!nodes) {
      rc = -2;
      goto done;
    }
    for (int d = 0; d < levels; ++d) {
      const hipError_t e = hipMemcpyAsync(nodes + base[d], lv[d], (size_t)count[d] * sizeof(BfQuadtreeNode),
                                          hipMemcpyDeviceToHost, st);
      if (e != hipSuccess) {
        free(nodes);
        TR_CHECK(e);
      }
      levelCount[d] = count[d];
    }
    const hipError_t e1 = hipMemcpyAsync(perm, iOut, (size_t)n * 8, hipMemcpyDeviceToHost, st);
    const hipError_t e2 = e1 == hipSuccess ? hipStreamSynchronize(st) : e1;
    if (e2 != hipSuccess) {
      free(nodes);
      TR_CHECK(e2);
    }
    *nodesOut = nodes;
    *numNodesOut = base[levels];
    *numLevelsOut = levels;
  }
done:
  (void)hipStreamSynchronize(st);
  for (int d = 0; d <= TREE_MAX_DEPTH; ++d) {
    (void)hipFree(start[d]);
    (void)hipFree(lv[d]);
  }
  (void)hipFree(dXY);
  (void)hipFree(dPart);
  (void)hipFree(kIn);
  (void)hipFree(kOut);
  (void)hipFree(iIn);
  (void)hipFree(iOut);
  (void)hipFree(flag);
  (void)hipFree(pos);
  (void)hipFree(tmp);
  return rc;
}

} /* extern "C" */
