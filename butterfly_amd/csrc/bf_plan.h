/*
 * Plan records shared by the host planner (C) and the HIP kernels.
 * Plain C types only; every record is a fixed-size POD so whole tables
 * are copied to HBM once and indexed by blockIdx.
 */
#pragma once

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* One re-expansion block (helm2.c:321-365):
 *   X (p x n) = pinv(K(test <- equiv)) K(test <- orig)
 * test: p points on circle (tcx, tcy, tr); equiv: p points on circle
 * (ecx, ecy, er); orig: n points on circle (ocx, ocy, orr) when
 * origOff < 0, else the source points [origOff, origOff + n) in
 * source-tree order.  X is written column-major at `out`.  Each job
 * carries its wavenumber k and pseudo-inverse cut rcond, so the jobs of
 * independent problems can share a launch. */
/* 16-byte units of one re-expansion job's LDS (or scratch) workspace:
 * M p x (p + n + 1), w p, v p, points 2p + n, reductions 8, pivots p/4 */
#define BF_LSTSQ_UNITS(p, n) \
  ((size_t)(p) * ((p) + (n) + 1) + (size_t)(p) + (size_t)(p) + (size_t)(2 * (p) + (n)) + 8 + (size_t)(((p) + 3) / 4))

typedef struct BfkLstsqJob {
  double ecx, ecy, er;
  double tcx, tcy, tr;
  double ocx, ocy, orr;
  double k, rcond;
  int64_t origOff;
  int64_t out;
  int32_t p, n;
} BfkLstsqJob;

/* One kernel block (helm2.c:93-125): K(tgt <- src), m x n, rows are the
 * target points [tgtOff, tgtOff + m) in target-tree order, columns are
 * n points on circle (scx, scy, sr) when srcOff < 0, else the source
 * points [srcOff, srcOff + n).  Written column-major at `out`. */
typedef struct BfkKernelJob {
  double scx, scy, sr;
  double k;
  int64_t srcOff;
  int64_t tgtOff;
  int64_t out;
  int32_t m, n;
} BfkKernelJob;

/* The layer potential of a factorization (uniform over its jobs): 0 the
 * single layer S, 1 the double layer D (source normals), 2 the combined
 * field alpha S + beta D; alpha = ar + i ai, beta = br + i bi; 3 the
 * normal derivative S' of the single layer (target normals; sampling only:
 * its re-expansions use the single layer, layer_pot.h:63-69). */
typedef struct BfkLayer {
  int32_t layer, pad;
  double ar, ai, br, bi;
} BfkLayer;

/* Vector buffers an apply stage reads or writes. */
enum { BFK_BUF_X = 0, BFK_BUF_PING = 1, BFK_BUF_PONG = 2, BFK_BUF_Y = 3 };

/* One output row segment of an apply stage: rows [yOff, yOff + m) of
 * buffer ybuf are the sum of terms [t0, t1). */
typedef struct BfkSeg {
  int64_t yOff;
  int32_t m;
  int32_t ybuf;
  int32_t t0, t1;
} BfkSeg;

/* One term: y += A[rowOff:rowOff+m, 0:n] x[xOff:xOff+n] with A stored
 * column-major at aOff with leading dimension lda; aOff < 0 means an
 * identity block (y += x[xOff:xOff+m]). */
typedef struct BfkTerm {
  int64_t aOff;
  int64_t xOff;
  int32_t n;
  int32_t lda;
  int32_t rowOff;
  int32_t xbuf;
} BfkTerm;

/* Orthogonalization of a butterfly (a change of basis of every
 * intermediate vector; the operator is unchanged).  One job is one column
 * block of one factor: its blocks (items [i0, i1)) are first multiplied on
 * the left by the triangular factor R_in of their row (if any), then, when
 * rOut >= 0, the stacked column (rows x n) is replaced by the Q of its
 * thin QR and R (n x n, column-major) is written at rOut for the rows of
 * the previous factor. */
typedef struct BfkOrthoJob {
  int64_t rOut;
  int32_t n, rows;
  int32_t i0, i1;
} BfkOrthoJob;

/* One block of an orthogonalization job: m x n column-major at aOff;
 * rIn >= 0 is the offset of its row's m x m R_in in the input R buffer. */
typedef struct BfkOrthoItem {
  int64_t aOff;
  int64_t rIn;
  int32_t m, pad;
} BfkOrthoItem;

#ifdef __cplusplus
}
#endif
