/*
 * HIP kernels of the 2-D Helmholtz butterfly for gfx950, and their C-ABI
 * launchers (called by the host library in C).
 *
 *   k_sample   kernel-block sampling: K(tgt <- src) blocks written to HBM
 *              (last factors, evaluation blocks, dense leaves;
 *              helm2.c:93-125, fac_helm2.c:403-509, :741-775)
 *   k_lstsq    re-expansion blocks: sample K(test <- equiv) and
 *              K(test <- orig) into LDS, one-sided Jacobi on the rows of
 *              the augmented [Z_equiv | Z_orig], then
 *              X = Z_equiv^+ Z_orig with the reference's truncation
 *              (helm2.c:321-365, mat_dense_complex.c:1768-1849)
 *   k_apply1   one stage of the apply for one right-hand side: every
 *              output row segment sums its column-major blocks times the
 *              input (mat_product.c:211-245, mat_block_coo.c:382-425,
 *              mat_block_diag.c:372-407, mat_block_dense.c:512-572)
 *   k_applyN   the same for many right-hand sides (one lane per column)
 *
 * One wave64 per job / segment; all jobs of a level or stage go into one
 * launch.  Everything here is FP64 unless the name says f32.
 */
#include <hip/hip_runtime.h>
#include <type_traits>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>

#include "bf_h0.h"
#include "bf_plan.h"

#define BFK_CHECK(call)                                                        \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "bf_kernels: %s failed: %s\n", #call, hipGetErrorString(e_)); \
      return (e_ == hipErrorOutOfMemory || e_ == hipErrorMemoryAllocation) ? -2 : -1; \
    }                                                                          \
  } while (0)

static hipStream_t g_stream = nullptr;
/* side streams for the compression launches (default priority, like
 * g_stream); they fork from and join back into g_stream */
#define BFK_SIDE_STREAMS 3
static hipStream_t g_side[BFK_SIDE_STREAMS] = {nullptr, nullptr, nullptr};
static hipEvent_t g_fork = nullptr, g_join[BFK_SIDE_STREAMS] = {nullptr, nullptr, nullptr};
static int g_nextStream = 0;
/* uploads that must not queue behind the kernels of g_stream */
static hipStream_t g_copy = nullptr;
static hipEvent_t g_ev0 = nullptr, g_ev1 = nullptr;
/* the compression launches return before they finish: their events are
 * read at the next bfk_sync / bfk_get_times */
static hipEvent_t g_evL0 = nullptr, g_evL1 = nullptr;
static int g_lPending = 0;
static size_t g_ldsCap = 0;
/* compression: LDS classes above g_lstsq2 bytes per job run 2 waves per
 * job, above g_lstsq4 4 waves (BF_AMD_LSTSQ_WAVES="a,b" overrides both;
 * measured at 2^20 on the curve: 1.34 s with one wave everywhere, 1.14 s
 * with these; lower or higher thresholds and 8 waves were slower) */
static size_t g_lstsq2 = 16384, g_lstsq4 = 40960;
static bool g_tallOk = false; /* k_tall_jacobi has its LDS */
static double g_times[8];   /* last measured ms per category */
static double g_sum[8];     /* accumulated ms per category since the last reset */
static double g_cnt[8];     /* number of measurements since the last reset */

static constexpr double TWO_PI = 6.283185307179586;

/* Fixed device workspace of the allocation-free apply (bfk_scratch): the
 * intermediate vectors of one apply live here instead of in buffers the
 * library allocates.  Each device holds its own copy for the life of the
 * process.  Host code never uses the array's name as an address: the
 * device address comes from hipGetSymbolAddress, looked up once per device
 * with that device current (bfk_init for the library's device, or the
 * first bfk_scratch call on another). */
#define BFK_SCRATCH_BYTES ((size_t)64 << 20)
#define BFK_MAX_DEVICES 16
static __device__ __align__(256) unsigned char scratch[BFK_SCRATCH_BYTES];
static void *g_scratchPtr[BFK_MAX_DEVICES];

/* ----------------------------------------------------------- helpers */

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 c) {
  return make_double2(fma(a.x, b.x, fma(-a.y, b.y, c.x)), fma(a.x, b.y, fma(a.y, b.x, c.y)));
}
__device__ __forceinline__ double2 shfl_xor2(double2 v, int m) {
  return make_double2(__shfl_xor(v.x, m, 64), __shfl_xor(v.y, m, 64));
}

/* proxy point i of n on a circle (circle.c:12-34) */
__device__ __forceinline__ double2 circle_pt(double cx, double cy, double r, int n, int i) {
  double scale = TWO_PI / (double)n;
  double th = scale * (double)i;
  double s, c;
  sincos(th, &s, &c);
  return make_double2(r * c + cx, r * s + cy);
}

/* G(t <- s) with r = |s - t| (points.c:484-505 order of subtraction) */
__device__ __forceinline__ double2 green(double k, double2 t, double2 s) {
  double r = hypot(s.x - t.x, s.y - t.y);
  double gr, gi;
  bf_helm2_single(k, r, &gr, &gi);
  return make_double2(gr, gi);
}

/* the layer potential of a factorization: 0 single, 1 double (with the
 * source normal ns), 2 combined field alpha S + beta D (helm2.c:61-90) */
__device__ __forceinline__ double2 green_layer(const BfkLayer &L, double k, double2 t, double2 s, double2 ns) {
  if (L.layer == 0) return green(k, t, s);
  const double dx = t.x - s.x, dy = t.y - s.y;
  double dr, di;
  bf_helm2_double(k, hypot(dx, dy), ns.x * dx + ns.y * dy, &dr, &di);
  if (L.layer == 1) return make_double2(dr, di);
  const double2 g = green(k, t, s);
  return make_double2(L.ar * g.x - L.ai * g.y + L.br * dr - L.bi * di, L.ar * g.y + L.ai * g.x + L.br * di + L.bi * dr);
}

/* S'(t <- s) = (i/4) k H1(k r) nt.(t - s)/r with the target normal nt, 0 for
 * r < 1e-15 (helm2.c:45-53, :127-175): the double-layer expression with the
 * target's normal in place of the source's */
__device__ __forceinline__ double2 green_tnormal(double k, double2 t, double2 s, double2 nt) {
  const double dx = t.x - s.x, dy = t.y - s.y;
  double dr, di;
  bf_helm2_double(k, hypot(dx, dy), nt.x * dx + nt.y * dy, &dr, &di);
  return make_double2(dr, di);
}

/* unit normal of a point on a proxy circle (circle.c: the radial direction) */
__device__ __forceinline__ double2 circle_normal(double cx, double cy, double r, double2 pt) {
  return make_double2((pt.x - cx) / r, (pt.y - cy) / r);
}

/* the same, as a call: keeps the Hankel coefficients from being hoisted
 * into registers for the whole of a kernel that samples only in its
 * prologue */
__device__ __attribute__((noinline)) double2 green_call(BfkLayer L, double k, double2 t, double2 s, double2 ns) {
  if (L.layer == 3) return green_tnormal(k, t, s, ns); /* ns is the target's normal here */
  return green_layer(L, k, t, s, ns);
}

/* ---------------------------------------------------------- sampling */

#define SAMPLE_THREADS 256
#define SAMPLE_CHUNK 1024

__global__ __launch_bounds__(SAMPLE_THREADS) void k_sample(const BfkKernelJob *__restrict__ jobs,
                                                           const double2 *__restrict__ srcXY,
                                                           const double2 *__restrict__ tgtXY,
                                                           const double2 *__restrict__ srcN,
                                                           const double2 *__restrict__ tgtN, BfkLayer L,
                                                           double2 *__restrict__ arena) {
  __shared__ double2 spts[SAMPLE_CHUNK];
  __shared__ double2 snrm[SAMPLE_CHUNK];
  const BfkKernelJob J = jobs[blockIdx.x];
  const int m = J.m;
  for (int c0 = 0; c0 < J.n; c0 += SAMPLE_CHUNK) {
    int cn = min(SAMPLE_CHUNK, J.n - c0);
    __syncthreads();
    for (int j = threadIdx.x; j < cn; j += SAMPLE_THREADS) {
      const double2 pt = J.srcOff >= 0 ? srcXY[J.srcOff + c0 + j] : circle_pt(J.scx, J.scy, J.sr, J.n, c0 + j);
      spts[j] = pt;
      if (L.layer == 1 || L.layer == 2)
        snrm[j] = J.srcOff >= 0 ? srcN[J.srcOff + c0 + j] : circle_normal(J.scx, J.scy, J.sr, pt);
    }
    __syncthreads();
    const int64_t cnt = (int64_t)m * cn;
    double2 *out = arena + J.out + (int64_t)c0 * m;
    for (int64_t e = threadIdx.x; e < cnt; e += SAMPLE_THREADS) {
      int j = (int)(e / m);
      int i = (int)(e - (int64_t)j * m);
      const double2 t = tgtXY[J.tgtOff + i];
      /* the entry through a call: inlined, the Hankel coefficients were
       * hoisted out of this loop into 428 registers (one wave per SIMD) */
      out[e] = L.layer == 3 ? green_call(L, J.k, t, spts[j], tgtN[J.tgtOff + i])
                            : green_call(L, J.k, t, spts[j], L.layer == 0 ? make_double2(0, 0) : snrm[j]);
    }
  }
}

/* -------------------------------------------------------- compression */

/* Per-job LDS (or scratch) layout, in 16-byte units:
 *   M    p x ld   row-major [Z_equiv | Z_orig], ld = p + n + 1 (odd pad)
 *   w    p        x = sigma^2, y = 1/sigma^2 or 0 (truncated)
 *   v    p        Householder vector
 *   pts  2p + n   test, equiv, orig points (sampling only)
 *   red  8        block reductions
 *   piv  p int32  column pivots (p/4 units rounded up) */
__host__ __device__ __forceinline__ size_t lstsq_units(int p, int n) { return BF_LSTSQ_UNITS(p, n); }

/* Round-robin pairing: position 0 is fixed, the others rotate by `step`. */
__device__ __forceinline__ int rr_row(int pos, int step, int P) {
  return pos == 0 ? 0 : 1 + (pos - 1 + step) % (P - 1);
}

/* ---- the re-expansion solve X = Z_equiv^+ Z_orig for M = [Z_equiv | Z_orig]
 * (p x (p+n)), on the workgroup's LDS slice (or, for the few blocks above
 * the LDS capacity, a slice of global scratch):
 *  1. Householder QR of Z_equiv with column pivoting, Z_equiv P = Q R,
 *     applied to all of M, giving [R | Q^H Z_orig];
 *  1b. X = P R^-1 Q^H Z_orig when the QR factor proves that no singular
 *     value falls below the cut (the pseudo-inverse is then the inverse);
 *  2. otherwise one-sided Jacobi on the rows of R (rotations applied to
 *     whole rows) until they are orthogonal: rows become sigma_i v_i^H,
 *     the right part U^H Q^H Z_orig (the QR step makes the graded rows
 *     converge in a few sweeps);
 *  3. keep sigma_i >= max(p eps sigma_0 + eps, rcond sigma_0)
 *     (mat_dense_complex.c:1797-1809) and form X = P sum_i R'_i^H B'_i /
 *     sigma_i^2, written column-major.
 * Every loop issues its loads in batches before it waits, lane groups
 * reduce with DPP lane swaps, the round-robin needs no integer modulo, and
 * each rotation costs two FP64 divides. */

/* x from the lane CTRL selects inside the lane's row of 16 (DPP).  Every
 * control used here reads a lane of the same row, and the callers run with
 * all lanes active, so no lane keeps an "old" value: the plain DPP move
 * needs no destination initialized beforehand. */
template <int CTRL>
__device__ __forceinline__ double dpp_d(double x) {
  const int lo = __builtin_amdgcn_mov_dpp(__double2loint(x), CTRL, 0xF, 0xF, true);
  const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(x), CTRL, 0xF, 0xF, true);
  return __hiloint2double(hi, lo);
}
template <int CTRL>
__device__ __forceinline__ int dpp_i(int x) {
  return __builtin_amdgcn_mov_dpp(x, CTRL, 0xF, 0xF, true);
}
/* partner lanes for groups of 2, 4, 8 and 16: quad_perm [1,0,3,2],
 * quad_perm [2,3,0,1], row_half_mirror, row_mirror.  After the k-th swap
 * every lane of an aligned group of 2^k lanes holds the group's sum. */
#define DPP_X1 0xB1
#define DPP_X2 0x4E
#define DPP_HM 0x141
#define DPP_M 0x140

/* sum over aligned groups of Lp lanes (Lp a power of two <= 64); every lane
 * of the wave must execute it */
__device__ __forceinline__ double group_sum(double v, int Lp) {
  if (Lp >= 2) v += dpp_d<DPP_X1>(v);
  if (Lp >= 4) v += dpp_d<DPP_X2>(v);
  if (Lp >= 8) v += dpp_d<DPP_HM>(v);
  if (Lp >= 16) v += dpp_d<DPP_M>(v);
  if (Lp >= 32) v += __shfl_xor(v, 16, 64);
  if (Lp >= 64) v += __shfl_xor(v, 32, 64);
  return v;
}

__device__ __forceinline__ void argmax_step(double &v, int &idx, double ov, int oi) {
  if (ov > v || (ov == v && oi < idx)) { v = ov; idx = oi; }
}
/* max over the wave with the smallest index among maximal entries */
__device__ __forceinline__ void wave_argmax(double &v, int &idx) {
  argmax_step(v, idx, dpp_d<DPP_X1>(v), dpp_i<DPP_X1>(idx));
  argmax_step(v, idx, dpp_d<DPP_X2>(v), dpp_i<DPP_X2>(idx));
  argmax_step(v, idx, dpp_d<DPP_HM>(v), dpp_i<DPP_HM>(idx));
  argmax_step(v, idx, dpp_d<DPP_M>(v), dpp_i<DPP_M>(idx));
  argmax_step(v, idx, __shfl_xor(v, 16, 64), __shfl_xor(idx, 16, 64));
  argmax_step(v, idx, __shfl_xor(v, 32, 64), __shfl_xor(idx, 32, 64));
}

/* Reductions over a workgroup of G threads (G/64 waves, G a multiple of
 * 64): inside the wave by lane swaps, then, for G > 64, over the waves
 * through `red` (2 G/64 doubles of LDS) between barriers; every thread gets
 * the same result, summed in the same order.  Every thread must call them. */
template <int G>
__device__ __forceinline__ double blk_sum(double v, double *red) {
  v = group_sum(v, 64);
  if (G == 64) return v;
  __syncthreads(); /* the previous reduction's reads of red are done */
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  v = red[0];
  for (int q = 1; q < G / 64; ++q) v += red[q];
  return v;
}
template <int G>
__device__ __forceinline__ void blk_argmax(double &v, int &idx, double *red) {
  wave_argmax(v, idx);
  if (G == 64) return;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) {
    red[2 * (threadIdx.x >> 6)] = v;
    red[2 * (threadIdx.x >> 6) + 1] = (double)idx;
  }
  __syncthreads();
  v = red[0];
  idx = (int)red[1];
  for (int q = 1; q < G / 64; ++q) argmax_step(v, idx, red[2 * q], (int)red[2 * q + 1]);
}
/* (through `red`, not __syncthreads_or, whose work-group reduction takes
 * static LDS and would leave the kernel short of the full 160 KiB) */
template <int G>
__device__ __forceinline__ bool blk_any(int v, double *red) {
  const bool w = __any(v);
  if (G == 64) return w;
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = w ? 1.0 : 0.0;
  __syncthreads();
  double s = red[0];
  for (int q = 1; q < G / 64; ++q) s += red[q];
  return s > 0;
}

/* round-robin (circle method) partner slot of position pos at step `step`
 * for P (even) rows: position 0 stays, the others rotate */
__device__ __forceinline__ int rr_row_fast(int pos, int step, int P) {
  if (pos == 0) return 0;
  int x = pos - 1 + step;
  if (x >= P - 1) x -= P - 1;
  return 1 + x;
}

/* |M[r][c]|^2 summed over rows [r0, p), loads issued four at a time */
__device__ __forceinline__ double col_norm2(const double2 *__restrict__ M, int ld, int c, int r0, int p) {
  double s0 = 0, s1 = 0;
  int r = r0;
  for (; r + 3 < p; r += 4) {
    const double2 a0 = M[r * ld + c], a1 = M[(r + 1) * ld + c], a2 = M[(r + 2) * ld + c], a3 = M[(r + 3) * ld + c];
    s0 = fma(a0.x, a0.x, fma(a0.y, a0.y, s0));
    s1 = fma(a1.x, a1.x, fma(a1.y, a1.y, s1));
    s0 = fma(a2.x, a2.x, fma(a2.y, a2.y, s0));
    s1 = fma(a3.x, a3.x, fma(a3.y, a3.y, s1));
  }
  for (; r < p; ++r) {
    const double2 a = M[r * ld + c];
    s0 = fma(a.x, a.x, fma(a.y, a.y, s0));
  }
  return s0 + s1;
}

/* |x[c]|^2 summed over c in [0, n) */
__device__ __forceinline__ double row_norm2(const double2 *__restrict__ x, int n) {
  double s0 = 0, s1 = 0;
  int c = 0;
  for (; c + 3 < n; c += 4) {
    const double2 a0 = x[c], a1 = x[c + 1], a2 = x[c + 2], a3 = x[c + 3];
    s0 = fma(a0.x, a0.x, fma(a0.y, a0.y, s0));
    s1 = fma(a1.x, a1.x, fma(a1.y, a1.y, s1));
    s0 = fma(a2.x, a2.x, fma(a2.y, a2.y, s0));
    s1 = fma(a3.x, a3.x, fma(a3.y, a3.y, s1));
  }
  for (; c < n; ++c) s0 = fma(x[c].x, x[c].x, fma(x[c].y, x[c].y, s0));
  return s0 + s1;
}

/* 1/sqrt(x) for normal x > 0: the hardware estimate and two Newton steps
 * (y <- y (3/2 - x y^2 / 2)), accurate to an ulp or two */
__device__ __forceinline__ double rsqrt_nr(double x) {
  double y = __builtin_amdgcn_rsq(x);
  const double h = 0.5 * x;
  y = y * fma(-h * y, y, 1.5);
  y = y * fma(-h * y, y, 1.5);
  return y;
}

/* One-sided Jacobi on the rows of R (p rows, W columns with the right-hand
 * part), round-robin pairs, LP lanes per pair (a power of two fixed at
 * compile time so the strides fold into the LDS offsets), on a workgroup
 * of G threads (one wave, or 2 or 4 waves for the larger blocks; `red` as
 * for blk_sum).  A rotation of
 * rows i < j with g = r_i . conj(r_j), a = |r_i|^2, b = |r_j|^2:
 *   e = g/|g|, zeta = (b - a)/(2|g|), t = sign(zeta)/(|zeta| + sqrt(1 + zeta^2)),
 *   c = 1/sqrt(1 + t^2), s = c t;  r_i <- c r_i - s e r_j,  r_j <- s r_i + c e r_j. */
template <int LP, int G>
__device__ __forceinline__ void jacobi_rows(double2 *__restrict__ M, int p, int W, int ld, double *red) {
  const int lane = threadIdx.x;
  const double eps = 2.220446049250313e-16;
  const int P = p + (p & 1);
  const int npairs = P / 2;
  constexpr int pairsPerPass = G / LP;
  const int slot = lane / LP;
  const int sub = lane - slot * LP;
  const double thr = 4.0 * eps * sqrt((double)p);
  const double thr2 = thr * thr;
  for (int sweep = 0; sweep < 40; ++sweep) {
    int rotated = 0;
    int big = 0; /* a rotation of this sweep had |g|^2 >= 1e-12 a b */
    for (int step = 0; step < P - 1; ++step) {
      for (int pb = 0; pb < npairs; pb += pairsPerPass) {
        const int pr = pb + slot;
        int i = 0, j = 0;
        bool valid = false;
        if (pr < npairs) {
          const int a = rr_row_fast(pr, step, P), b = rr_row_fast(P - 1 - pr, step, P);
          i = min(a, b);
          j = max(a, b);
          valid = j < p;
        }
        double al = 0, be = 0, gx = 0, gy = 0;
        if (valid) {
          const double2 *ri = M + i * ld + sub, *rj = M + j * ld + sub;
          const int cnt = (p - sub + LP - 1) / LP; /* columns sub, sub + LP, ... below p */
          int q = 0;
          for (; q + 1 < cnt; q += 2) {
            const double2 x0 = ri[q * LP], y0 = rj[q * LP], x1 = ri[(q + 1) * LP], y1 = rj[(q + 1) * LP];
            al = fma(x0.x, x0.x, fma(x0.y, x0.y, al));
            be = fma(y0.x, y0.x, fma(y0.y, y0.y, be));
            gx = fma(x0.x, y0.x, fma(x0.y, y0.y, gx)); /* x conj(y) */
            gy = fma(x0.y, y0.x, fma(-x0.x, y0.y, gy));
            al = fma(x1.x, x1.x, fma(x1.y, x1.y, al));
            be = fma(y1.x, y1.x, fma(y1.y, y1.y, be));
            gx = fma(x1.x, y1.x, fma(x1.y, y1.y, gx));
            gy = fma(x1.y, y1.x, fma(-x1.x, y1.y, gy));
          }
          if (q < cnt) {
            const double2 x0 = ri[q * LP], y0 = rj[q * LP];
            al = fma(x0.x, x0.x, fma(x0.y, x0.y, al));
            be = fma(y0.x, y0.x, fma(y0.y, y0.y, be));
            gx = fma(x0.x, y0.x, fma(x0.y, y0.y, gx));
            gy = fma(x0.y, y0.x, fma(-x0.x, y0.y, gy));
          }
        }
        al = group_sum(al, LP);
        be = group_sum(be, LP);
        gx = group_sum(gx, LP);
        gy = group_sum(gy, LP);
        const double ag2 = gx * gx + gy * gy;
        /* rows below 1e-145 in norm lie far under any cut; they are not rotated */
        if (valid && ag2 > 1e-290 && ag2 > thr2 * al * be) {
          rotated = 1;
          big |= ag2 >= 1e-12 * (al * be);
          const double rg = rsqrt_nr(ag2);
          const double ex = gx * rg, ey = gy * rg;
          const double zeta = (be - al) * (0.5 * rg);
          const double az = fabs(zeta);
          double t;
          if (az > 1e150) {
            t = 0.5 / zeta;
          } else {
            /* t only chooses the angle: c = 1/sqrt(1 + t^2) and s = c t,
             * both to full precision, make the rotation unitary whatever
             * t's last bits, so t = sign(zeta) / (|zeta| + sqrt(1 + zeta^2))
             * comes from the hardware estimates and Newton steps (one for
             * the square root, two for the reciprocal of d >= 1) */
            const double q1 = fma(zeta, zeta, 1.0);
            double y = __builtin_amdgcn_rsq(q1);
            y = y * fma(-0.5 * q1 * y, y, 1.5);
            const double d = fma(q1, y, az);
            double r = __builtin_amdgcn_rcp(d);
            r = fma(r, fma(-d, r, 1.0), r);
            r = fma(r, fma(-d, r, 1.0), r);
            t = zeta >= 0 ? r : -r;
          }
          const double cs = rsqrt_nr(fma(t, t, 1.0)), sn = cs * t;
          const double sx = sn * ex, sy = sn * ey, cx = cs * ex, cy = cs * ey;
          double2 *ri = M + i * ld + sub, *rj = M + j * ld + sub;
          const int cnt = (W - sub + LP - 1) / LP;
          int q = 0;
          for (; q + 3 < cnt; q += 4) {
            double2 x[4], y[4];
#pragma unroll
            for (int u = 0; u < 4; ++u) { x[u] = ri[(q + u) * LP]; y[u] = rj[(q + u) * LP]; }
#pragma unroll
            for (int u = 0; u < 4; ++u) {
              ri[(q + u) * LP] = make_double2(fma(cs, x[u].x, -fma(sx, y[u].x, -sy * y[u].y)),
                                              fma(cs, x[u].y, -fma(sx, y[u].y, sy * y[u].x)));
              rj[(q + u) * LP] = make_double2(fma(sn, x[u].x, fma(cx, y[u].x, -cy * y[u].y)),
                                              fma(sn, x[u].y, fma(cx, y[u].y, cy * y[u].x)));
            }
          }
          for (; q < cnt; ++q) {
            const double2 x0 = ri[q * LP], y0 = rj[q * LP];
            ri[q * LP] = make_double2(fma(cs, x0.x, -fma(sx, y0.x, -sy * y0.y)),
                                      fma(cs, x0.y, -fma(sx, y0.y, sy * y0.x)));
            rj[q * LP] = make_double2(fma(sn, x0.x, fma(cx, y0.x, -cy * y0.y)),
                                      fma(sn, x0.y, fma(cx, y0.y, cy * y0.x)));
          }
        }
        __syncthreads();
      }
    }
    if (!blk_any<G>(rotated, red)) break;
    /* cyclic Jacobi converges quadratically: once every rotation of a sweep
     * was below 1e-6 in |g| / sqrt(a b), the rows it leaves are orthogonal
     * to about 1e-12 relative, an order under the 1e-11 singular-value cut,
     * so the next sweep, which would move singular values and vectors by
     * that much, is skipped */
    if (!blk_any<G>(big, red)) break;
  }
}

/* Pivoted QR helpers: S lanes per column (S = 1, 2 or 4, an aligned group
 * inside a quad), each lane over every S-th row from the step's first row,
 * the S partial sums combined by lane swaps.  S is the widest split whose
 * S x (columns) lanes fit in the workgroup, so the late steps of a small
 * block, and the whole QR of a p ~ 16 block, keep most lanes busy. */
template <int G>
__device__ __forceinline__ int qr_split(int ncols) {
  return 4 * ncols <= G ? 4 : 2 * ncols <= G ? 2 : 1;
}

/* the largest |column c|^2 over rows [j, p) among columns c in [j, p), and
 * its column (the smallest index among maximal norms) */
template <int S, int G>
__device__ __forceinline__ void qr_pivot_norms(const double2 *__restrict__ M, int ld, int j, int p, double &best,
                                               int &bidx) {
  const int lane = threadIdx.x;
  for (int t = lane; t < (p - j) * S; t += G) {
    const int c = j + t / S;
    double s0 = 0, s1 = 0;
    int r = j + (t & (S - 1));
    for (; r + S < p; r += 2 * S) {
      const double2 a0 = M[r * ld + c], a1 = M[(r + S) * ld + c];
      s0 = fma(a0.x, a0.x, fma(a0.y, a0.y, s0));
      s1 = fma(a1.x, a1.x, fma(a1.y, a1.y, s1));
    }
    if (r < p) {
      const double2 a0 = M[r * ld + c];
      s0 = fma(a0.x, a0.x, fma(a0.y, a0.y, s0));
    }
    const double sn = group_sum(s0 + s1, S);
    if (sn > best) { best = sn; bidx = c; }
  }
}

/* columns j+1 .. W-1 of M minus v (v^H y) inv over rows [j, p) */
template <int S, int G>
__device__ __forceinline__ void qr_reflect(double2 *__restrict__ M, const double2 *__restrict__ v, int ld, int j,
                                           int p, int W, double inv) {
  const int lane = threadIdx.x;
  for (int t = lane; t < (W - j - 1) * S; t += G) {
    const int c = j + 1 + t / S, r0 = j + (t & (S - 1));
    double sr = 0, si = 0, tr = 0, ti = 0; /* v^H y in two chains */
    int r = r0;
    for (; r + S < p; r += 2 * S) {
      const double2 v0 = v[r], y0 = M[r * ld + c], v1 = v[r + S], y1 = M[(r + S) * ld + c];
      sr = fma(v0.x, y0.x, fma(v0.y, y0.y, sr));
      si = fma(v0.x, y0.y, fma(-v0.y, y0.x, si));
      tr = fma(v1.x, y1.x, fma(v1.y, y1.y, tr));
      ti = fma(v1.x, y1.y, fma(-v1.y, y1.x, ti));
    }
    if (r < p) {
      const double2 v0 = v[r], y0 = M[r * ld + c];
      sr = fma(v0.x, y0.x, fma(v0.y, y0.y, sr));
      si = fma(v0.x, y0.y, fma(-v0.y, y0.x, si));
    }
    sr = group_sum(sr + tr, S) * inv;
    si = group_sum(si + ti, S) * inv;
    r = r0;
    for (; r + 3 * S < p; r += 4 * S) {
      double2 y[4], vv[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) { vv[u] = v[r + u * S]; y[u] = M[(r + u * S) * ld + c]; }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        y[u].x -= sr * vv[u].x - si * vv[u].y;
        y[u].y -= sr * vv[u].y + si * vv[u].x;
        M[(r + u * S) * ld + c] = y[u];
      }
    }
    for (; r < p; r += S) {
      const double2 vr = v[r];
      double2 y = M[r * ld + c];
      y.x -= sr * vr.x - si * vr.y;
      y.y -= sr * vr.y + si * vr.x;
      M[r * ld + c] = y;
    }
  }
}

/* The solve of one re-expansion block by a workgroup of G threads: one
 * wave for the small blocks, 2 or 4 waves for the larger ones, whose LDS
 * (20-60 KiB) leaves room for only a few blocks per CU: more waves per
 * block keep the CU busy where one wave per block would leave it waiting
 * on LDS and dependent arithmetic.  Every branch below is uniform over the
 * workgroup (its conditions come from block reductions). */
template <int G>
__device__ __forceinline__ void qr_jacobi_group(double2 *__restrict__ M, int p, int n, int ld, double2 *__restrict__ w,
                                                double2 *__restrict__ v, int *__restrict__ piv, double *__restrict__ red,
                                                double rcond, double2 *__restrict__ out) {
  const int lane = threadIdx.x;
  const int W = p + n;
  const double eps = 2.220446049250313e-16;

  for (int i = lane; i < p; i += G) piv[i] = i;
  __syncthreads();

  /* ---- 1. pivoted Householder QR (S lanes per column, qr_split) */
  for (int j = 0; j < p; ++j) {
    double best = -1.0;
    int bidx = p;
    switch (qr_split<G>(p - j)) {
      case 4: qr_pivot_norms<4, G>(M, ld, j, p, best, bidx); break;
      case 2: qr_pivot_norms<2, G>(M, ld, j, p, best, bidx); break;
      default: qr_pivot_norms<1, G>(M, ld, j, p, best, bidx); break;
    }
    blk_argmax<G>(best, bidx, red);
    const int q = __builtin_amdgcn_readfirstlane(bidx);
    if (q != j) {
      for (int r = lane; r < p; r += G) {
        const double2 t = M[r * ld + j];
        M[r * ld + j] = M[r * ld + q];
        M[r * ld + q] = t;
      }
      if (lane == 0) { const int t = piv[j]; piv[j] = piv[q]; piv[q] = t; }
    }
    __syncthreads();
    const double sig = sqrt(best);
    if (sig > 0) {
      const double2 x0 = M[j * ld + j];
      const double ax0 = hypot(x0.x, x0.y);
      const double phx = ax0 > 0 ? x0.x / ax0 : 1.0, phy = ax0 > 0 ? x0.y / ax0 : 0.0;
      for (int r = j + lane; r < p; r += G) {
        double2 a = M[r * ld + j];
        if (r == j) a = make_double2(x0.x + phx * sig, x0.y + phy * sig);
        v[r] = a;
      }
      __syncthreads();
      const double inv = 1.0 / (sig * (sig + ax0));
      switch (qr_split<G>(W - j - 1)) {
        case 4: qr_reflect<4, G>(M, v, ld, j, p, W, inv); break;
        case 2: qr_reflect<2, G>(M, v, ld, j, p, W, inv); break;
        default: qr_reflect<1, G>(M, v, ld, j, p, W, inv); break;
      }
      __syncthreads();
      for (int r = j + lane; r < p; r += G)
        M[r * ld + j] = (r == j) ? make_double2(-phx * sig, -phy * sig) : make_double2(0, 0);
    }
    __syncthreads();
  }

  /* ---- 1b. when no singular value can fall below the cut, the
   * pseudo-inverse is the inverse and X = R^-1 Q^H B.  sigma_max <= |R|_F
   * bounds the cut from above and sigma_min >= 1 / |R^-1|_F bounds the
   * smallest singular value from below; Z = R^-1 goes into the (zero)
   * strict lower triangle of M, transposed, with its diagonal in w. */
  {
    double rf = 0;
    for (int c = lane; c < p; c += G) rf += col_norm2(M, ld, c, 0, c + 1);
    rf = blk_sum<G>(rf, red);
    for (int i = lane; i < p; i += G) {
      const double2 d = M[i * ld + i];
      const double m2 = d.x * d.x + d.y * d.y;
      v[i] = make_double2(d.x / m2, -d.y / m2);
    }
    __syncthreads();
    double zf = 0;
    for (int j = lane; j < p; j += G) {
      const double2 zjj = v[j];
      w[j] = zjj;
      zf = zjj.x * zjj.x + zjj.y * zjj.y;
      double2 *zj = M + j * ld; /* z_kj at zj[k] for k < j */
      for (int i = j - 1; i >= 0; --i) {
        const double2 *ri = M + i * ld;
        const double2 rij = ri[j];
        double sx = rij.x * zjj.x - rij.y * zjj.y, sy = rij.x * zjj.y + rij.y * zjj.x;
        double tx = 0, ty = 0;
        int k = i + 1;
        for (; k + 1 < j; k += 2) {
          const double2 a0 = ri[k], b0 = zj[k], a1 = ri[k + 1], b1 = zj[k + 1];
          sx = fma(a0.x, b0.x, fma(-a0.y, b0.y, sx));
          sy = fma(a0.x, b0.y, fma(a0.y, b0.x, sy));
          tx = fma(a1.x, b1.x, fma(-a1.y, b1.y, tx));
          ty = fma(a1.x, b1.y, fma(a1.y, b1.x, ty));
        }
        if (k < j) {
          const double2 a0 = ri[k], b0 = zj[k];
          sx = fma(a0.x, b0.x, fma(-a0.y, b0.y, sx));
          sy = fma(a0.x, b0.y, fma(a0.y, b0.x, sy));
        }
        sx += tx;
        sy += ty;
        const double2 di = v[i];
        const double2 z = make_double2(-(sx * di.x - sy * di.y), -(sx * di.y + sy * di.x));
        zj[i] = z;
        zf = fma(z.x, z.x, fma(z.y, z.y, zf));
      }
    }
    zf = blk_sum<G>(zf, red);
    const double sigMaxUp = sqrt(rf);
    const double cutUp = fmax((double)p * eps * sigMaxUp + eps, rcond * sigMaxUp);
    const bool finiteZ = zf > 0 && zf <= 1.7976931348623157e308;
    bool direct = finiteZ && 1.0 / sqrt(zf) >= cutUp;
    if (!direct && finiteZ) {
      /* a tighter upper bound on sigma_max before giving up: ||R||_F
       * exceeds it by up to sqrt(p) when several singular values are
       * comparable, while (sum sigma^4)^(1/4) = ||R^H R||_F^(1/2) never
       * falls below it and sits much closer.  B = R^H R entry by entry
       * from R's upper triangle (B_ij = sum_{k <= min(i,j)} conj(R_ki) R_kj),
       * summing |B_ij|^2 over i <= j, doubled off the diagonal. */
      double b4 = 0;
      int i = lane / p, j = lane - (lane / p) * p;
      const int di = G / p, dj = G - di * p;
      for (int e = lane; e < p * p; e += G) {
        if (i <= j) {
          double bx = 0, by = 0;
          for (int k = 0; k <= i; ++k) {
            const double2 a = M[k * ld + i], b = M[k * ld + j];
            bx = fma(a.x, b.x, fma(a.y, b.y, bx)); /* conj(a) b */
            by = fma(a.x, b.y, fma(-a.y, b.x, by));
          }
          b4 += (i == j ? 1.0 : 2.0) * fma(bx, bx, by * by);
        }
        j += dj;
        i += di;
        if (j >= p) { j -= p; ++i; }
      }
      b4 = blk_sum<G>(b4, red);
      const double s4 = sqrt(sqrt(b4));
      direct = 1.0 / sqrt(zf) >= fmax((double)p * eps * s4 + eps, rcond * s4);
    }
    __syncthreads();
    if (direct) {
      const int cnt = p * n;
      int c = lane / p, a = lane - (lane / p) * p;
      const int dc = G / p, da = G - dc * p;
      for (int e = lane; e < cnt; e += G) {
        const double2 za = w[a], b0 = M[a * ld + p + c];
        double ax = za.x * b0.x - za.y * b0.y, ay = za.x * b0.y + za.y * b0.x;
        double bx = 0, by = 0;
        int i = a + 1;
        for (; i + 1 < p; i += 2) {
          const double2 z0 = M[i * ld + a], y0 = M[i * ld + p + c], z1 = M[(i + 1) * ld + a],
                        y1 = M[(i + 1) * ld + p + c];
          ax = fma(z0.x, y0.x, fma(-z0.y, y0.y, ax));
          ay = fma(z0.x, y0.y, fma(z0.y, y0.x, ay));
          bx = fma(z1.x, y1.x, fma(-z1.y, y1.y, bx));
          by = fma(z1.x, y1.y, fma(z1.y, y1.x, by));
        }
        if (i < p) {
          const double2 z0 = M[i * ld + a], y0 = M[i * ld + p + c];
          ax = fma(z0.x, y0.x, fma(-z0.y, y0.y, ax));
          ay = fma(z0.x, y0.y, fma(z0.y, y0.x, ay));
        }
        out[(size_t)c * p + piv[a]] = make_double2(ax + bx, ay + by);
        c += dc;
        a += da;
        if (a >= p) { a -= p; ++c; }
      }
      return;
    }
    /* back to R: clear the strict lower triangle again */
    for (int i = lane; i < p; i += G)
      for (int k = 0; k < i; ++k) M[i * ld + k] = make_double2(0, 0);
    __syncthreads();
  }

  /* ---- 2. one-sided Jacobi on the rows of R */
  {
    const int npairs = (p + (p & 1)) / 2;
    int Lp = 64;
    while (Lp > 1 && Lp * npairs > G) Lp >>= 1;
    switch (Lp) {
      case 8: jacobi_rows<8, G>(M, p, W, ld, red); break;
      case 4: jacobi_rows<4, G>(M, p, W, ld, red); break;
      case 2: jacobi_rows<2, G>(M, p, W, ld, red); break;
      case 1: jacobi_rows<1, G>(M, p, W, ld, red); break;
      case 16: jacobi_rows<16, G>(M, p, W, ld, red); break;
      case 32: jacobi_rows<32, G>(M, p, W, ld, red); break;
      default: jacobi_rows<64, G>(M, p, W, ld, red); break;
    }
  }

  /* ---- 3. truncation and X */
  double smax = 0;
  for (int i = lane; i < p; i += G) {
    const double s = row_norm2(M + i * ld, p);
    w[i].x = s;
    smax = fmax(smax, s);
  }
  int dummy = 0;
  blk_argmax<G>(smax, dummy, red);
  __syncthreads();
  const double tol = fmax((double)p * eps * sqrt(smax) + eps, rcond * sqrt(smax));
  for (int i = lane; i < p; i += G) {
    const double s = sqrt(w[i].x);
    w[i].y = (s >= tol) ? 1.0 / w[i].x : 0.0;
  }
  __syncthreads();
  const int cnt = p * n;
  int c = lane / p, a = lane - (lane / p) * p;
  const int dc = G / p, da = G - dc * p;
  for (int e = lane; e < cnt; e += G) {
    double ax = 0, ay = 0, bx = 0, by = 0;
    int i = 0;
    for (; i + 1 < p; i += 2) {
      const double w0 = w[i].y, w1 = w[i + 1].y;
      const double2 x0 = M[i * ld + a], y0 = M[i * ld + p + c], x1 = M[(i + 1) * ld + a], y1 = M[(i + 1) * ld + p + c];
      ax = fma(w0, x0.x * y0.x + x0.y * y0.y, ax); /* conj(x) y */
      ay = fma(w0, x0.x * y0.y - x0.y * y0.x, ay);
      bx = fma(w1, x1.x * y1.x + x1.y * y1.y, bx);
      by = fma(w1, x1.x * y1.y - x1.y * y1.x, by);
    }
    if (i < p) {
      const double w0 = w[i].y;
      const double2 x0 = M[i * ld + a], y0 = M[i * ld + p + c];
      ax = fma(w0, x0.x * y0.x + x0.y * y0.y, ax);
      ay = fma(w0, x0.x * y0.y - x0.y * y0.x, ay);
    }
    out[(size_t)c * p + piv[a]] = make_double2(ax + bx, ay + by);
    /* next e = e + G: advance (c, a) without dividing */
    c += dc;
    a += da;
    if (a >= p) { a -= p; ++c; }
  }
}

struct LstsqView {
  double2 *M, *w, *v, *pts;
  double *red;
  int *piv;
};

__device__ __forceinline__ LstsqView lstsq_view(double2 *base, int p, int n) {
  LstsqView L;
  const int ld = p + n + 1;
  L.M = base;
  L.w = L.M + (size_t)p * ld;
  L.v = L.w + p;
  L.pts = L.v + p;
  L.red = (double *)(L.pts + 2 * p + n);
  L.piv = (int *)(L.pts + 2 * p + n + 8);
  return L;
}

/* One job per block of T threads.  The working set lives in LDS, or for
 * jobs too large for LDS (about one in 10^4 at N = 2^20) in a per-block
 * slice of a global scratch buffer. */
/* five waves per SIMD: the solve fits in 96 VGPRs without spills, and the
 * LDS of the common 8-12 KiB jobs allows about that many per CU anyway */
template <int T, bool GLOBAL>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(5))) void k_lstsq(const BfkLstsqJob *__restrict__ jobs, const int32_t *__restrict__ order,
                                             const double2 *__restrict__ srcXY,
                                             const double2 *__restrict__ srcN, BfkLayer Lay,
                                             double2 *__restrict__ arena, double2 *__restrict__ gwork,
                                             int64_t gstride) {
  extern __shared__ double2 lds[];
  const BfkLstsqJob J = jobs[order ? order[blockIdx.x] : blockIdx.x];
  const int p = J.p, n = J.n, ld = p + n + 1;
  LstsqView L = lstsq_view(GLOBAL ? gwork + (int64_t)blockIdx.x * gstride : lds, p, n);
  double2 *tp = L.pts, *ep = tp + p, *op = ep + p;
  const int lane = threadIdx.x;
  for (int i = lane; i < p; i += T) {
    tp[i] = circle_pt(J.tcx, J.tcy, J.tr, p, i);
    ep[i] = circle_pt(J.ecx, J.ecy, J.er, p, i);
  }
  for (int c = lane; c < n; c += T)
    op[c] = J.origOff >= 0 ? srcXY[J.origOff + c] : circle_pt(J.ocx, J.ocy, J.orr, n, c);
  __syncthreads();
  const int W = p + n;
  /* entry e = i W + c; (i, c) advance by T entries without dividing */
  int i = lane / W, c = lane - (lane / W) * W;
  const int di = T / W, dc = T - di * W;
  for (int e = lane; e < p * W; e += T, i += di, c += dc) {
    if (c >= W) { c -= W; ++i; }
    double2 s = c < p ? ep[c] : op[c - p];
    double2 ns = make_double2(0, 0);
    if (Lay.layer != 0)
      ns = c < p ? circle_normal(J.ecx, J.ecy, J.er, s)
                 : (J.origOff >= 0 ? srcN[J.origOff + c - p] : circle_normal(J.ocx, J.ocy, J.orr, s));
    L.M[i * ld + c] = green_call(Lay, J.k, tp[i], s, ns);
  }
  __syncthreads();
  qr_jacobi_group<T>(L.M, p, n, ld, L.w, L.v, L.piv, L.red, J.rcond, arena + J.out);
}

/* the same solve on matrices given in global memory (row-major A p x p,
 * B p x n per batch entry); used to check the solver on its own */
__global__ __launch_bounds__(64) void k_lstsq_loaded(const double2 *__restrict__ A, const double2 *__restrict__ B,
                                                     int p, int n, double2 *__restrict__ X) {
  extern __shared__ double2 lds[];
  const int ld = p + n + 1;
  LstsqView L = lstsq_view(lds, p, n);
  const double2 *a = A + (size_t)blockIdx.x * p * p;
  const double2 *b = B + (size_t)blockIdx.x * p * n;
  const int W = p + n;
  for (int e = threadIdx.x; e < p * W; e += 64) {
    int i = e / W;
    int c = e - i * W;
    L.M[i * ld + c] = c < p ? a[i * p + c] : b[i * n + (c - p)];
  }
  __syncthreads();
  qr_jacobi_group<64>(L.M, p, n, ld, L.w, L.v, L.piv, L.red, 0.0, X + (size_t)blockIdx.x * p * n);
}

/* ------------------------------------ boundary integral operator, GMRES */

/* G(a_i <- b_i) for a list of point pairs (the quadrature corrections), or
 * S'(a_i <- b_i) with the normal na_i at a_i when na is given */
__global__ void k_pair_green(int64_t n, const double2 *__restrict__ a, const double2 *__restrict__ b,
                             const double2 *__restrict__ na, double k, double2 *__restrict__ out) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = na ? green_tnormal(k, a[i], b[i], na[i]) : green(k, a[i], b[i]);
}

/* u = w .* x, row i scaled by w[i], nrhs columns (row-major) */
template <typename V>
__global__ void k_scale_rows(int64_t n, int nrhs, const double *__restrict__ w, const V *__restrict__ x,
                             V *__restrict__ u) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nrhs) return;
  const double s = w[e / nrhs];
  V v = x[e];
  v.x *= s;
  v.y *= s;
  u[e] = v;
}

/* y = alpha x + y + sum_q c[i][q] u[col[i][q]] (the identity term and the
 * near-singular corrections of one row, after the butterfly wrote y) */
template <typename V>
__global__ void k_bie_post(int64_t n, int nrhs, int ncorr, double alpha, const int32_t *__restrict__ col,
                           const double2 *__restrict__ c, const V *__restrict__ x, const V *__restrict__ u,
                           V *__restrict__ y) {
  const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= n * nrhs) return;
  const int64_t i = e / nrhs;
  const int r = (int)(e - i * nrhs);
  double ax = alpha * (double)x[e].x + (double)y[e].x, ay = alpha * (double)x[e].y + (double)y[e].y;
  for (int q = 0; q < ncorr; ++q) {
    const double2 cq = c[i * ncorr + q];
    const V uq = u[(int64_t)col[i * ncorr + q] * nrhs + r];
    ax = fma(cq.x, (double)uq.x, fma(-cq.y, (double)uq.y, ax));
    ay = fma(cq.x, (double)uq.y, fma(cq.y, (double)uq.x, ay));
  }
  V out;
  out.x = ax;
  out.y = ay;
  y[e] = out;
}

/* Partial sums of h_j = V_j^H w for j < nv: block b writes its share of
 * every h_j to part[b * nv + j] (summed on the host, in block order). */
__global__ __launch_bounds__(256) void k_dotc_multi(int64_t n, int nv, const double2 *__restrict__ V, int64_t ldv,
                                                    const double2 *__restrict__ w, double2 *__restrict__ part) {
  __shared__ double2 red[256];
  const int64_t chunk = (n + gridDim.x - 1) / gridDim.x;
  const int64_t a = (int64_t)blockIdx.x * chunk, e = a + chunk < n ? a + chunk : n;
  for (int j = 0; j < nv; ++j) {
    double sx = 0, sy = 0;
    for (int64_t i = a + threadIdx.x; i < e; i += 256) {
      const double2 v = V[j * ldv + i], x = w[i];
      sx = fma(v.x, x.x, fma(v.y, x.y, sx)); /* conj(v) x */
      sy = fma(v.x, x.y, fma(-v.y, x.x, sy));
    }
    red[threadIdx.x] = make_double2(sx, sy);
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
      if ((int)threadIdx.x < o) {
        red[threadIdx.x].x += red[threadIdx.x + o].x;
        red[threadIdx.x].y += red[threadIdx.x + o].y;
      }
      __syncthreads();
    }
    if (threadIdx.x == 0) part[(int64_t)blockIdx.x * nv + j] = red[0];
    __syncthreads();
  }
}

/* w = s w - V h (nv vectors of V, leading dimension ldv) */
__global__ void k_axpy_multi(int64_t n, int nv, const double2 *__restrict__ V, int64_t ldv,
                             const double2 *__restrict__ h, double s, double2 *__restrict__ w) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  double2 acc = w[i];
  acc.x *= s;
  acc.y *= s;
  for (int j = 0; j < nv; ++j) {
    const double2 v = V[j * ldv + i], c = h[j];
    acc.x -= c.x * v.x - c.y * v.y;
    acc.y -= c.x * v.y + c.y * v.x;
  }
  w[i] = acc;
}

__global__ void k_f64_to_f32(int64_t n, const double2 *__restrict__ a, float2 *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = make_float2((float)a[i].x, (float)a[i].y);
}

__global__ void k_f32_to_f64(int64_t n, const float2 *__restrict__ a, double2 *__restrict__ b) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) b[i] = make_double2((double)a[i].x, (double)a[i].y);
}

/* ------------------------------------------------ batched one-sided SVD */

/* Singular value decomposition of a batch of m x n complex matrices
 * (column-major, m >= 1, n >= 1; the host keeps n <= 1024, as the work grows
 * with n^2), one 256-thread block each, in
 * place: one-sided (Hestenes) Jacobi rotates pairs of columns until they
 * are orthogonal, A V = W with orthogonal columns w_j = sigma_j u_j.  On
 * return A holds W, V (n x n, column-major) the accumulated rotations and
 * sig[j] = |w_j|.  Pairs follow the round-robin order, one group of g
 * threads per pair (g a power of two <= 64, so a pair's reduction stays
 * in one wave). */
__global__ __launch_bounds__(256) void k_svd_jacobi(const int64_t *__restrict__ aOff, const int32_t *__restrict__ mm,
                                                    const int32_t *__restrict__ nn, double2 *__restrict__ A,
                                                    const int64_t *__restrict__ vOff, double2 *__restrict__ V,
                                                    const int64_t *__restrict__ sOff, double *__restrict__ sig,
                                                    int maxSweeps) {
  const int b = blockIdx.x;
  const int m = mm[b], n = nn[b];
  double2 *a = A + aOff[b];
  double2 *v = V + vOff[b];
  const int tid = threadIdx.x;
  const double eps = 2.220446049250313e-16;
  for (int e = tid; e < n * n; e += 256) v[e] = make_double2((e % (n + 1)) == 0 ? 1.0 : 0.0, 0.0);
  __syncthreads();
  const int P = n + (n & 1);
  const int npairs = P / 2;
  int g = 64;
  while (g > 1 && g * npairs > 256) g >>= 1;
  const int slots = 256 / g;
  const int slot = tid / g, sub = tid - slot * g;
  __shared__ int anyRot;
  for (int sweep = 0; sweep < maxSweeps && n > 1; ++sweep) {
    if (tid == 0) anyRot = 0;
    __syncthreads();
    for (int step = 0; step < P - 1; ++step) {
      for (int pb = 0; pb < npairs; pb += slots) {
        const int pr = pb + slot;
        int i = -1, j = -1;
        if (pr < npairs) {
          const int p0 = rr_row(pr, step, P), p1 = rr_row(P - 1 - pr, step, P);
          i = min(p0, p1);
          j = max(p0, p1);
        }
        const bool valid = pr < npairs && j < n;
        double al = 0, be = 0, gx = 0, gy = 0;
        if (valid) {
          const double2 *ci = a + (int64_t)i * m, *cj = a + (int64_t)j * m;
          for (int r = sub; r < m; r += g) {
            double2 x = ci[r], y = cj[r];
            al = fma(x.x, x.x, fma(x.y, x.y, al));
            be = fma(y.x, y.x, fma(y.y, y.y, be));
            gx = fma(x.x, y.x, fma(x.y, y.y, gx)); /* conj(x) y */
            gy = fma(x.x, y.y, fma(-x.y, y.x, gy));
          }
        }
        for (int o = 1; o < g; o <<= 1) {
          al += __shfl_xor(al, o, 64);
          be += __shfl_xor(be, o, 64);
          gx += __shfl_xor(gx, o, 64);
          gy += __shfl_xor(gy, o, 64);
        }
        const double ag = sqrt(gx * gx + gy * gy);
        if (valid && ag > 0 && ag > eps * sqrt(al * be)) {
          if (sub == 0) anyRot = 1;
          /* rotate (x, y) <- (c x - s e y, s x + c e y), e = conj(g)/|g| */
          const double ex = gx / ag, ey = -gy / ag;
          const double zeta = (be - al) / (2 * ag);
          const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
          const double cs = 1 / sqrt(1 + t * t), sn = cs * t;
          double2 *ci = a + (int64_t)i * m, *cj = a + (int64_t)j * m;
          for (int r = sub; r < m; r += g) {
            double2 x = ci[r], y = cj[r];
            double ux = ex * y.x - ey * y.y, uy = ex * y.y + ey * y.x;
            ci[r] = make_double2(cs * x.x - sn * ux, cs * x.y - sn * uy);
            cj[r] = make_double2(sn * x.x + cs * ux, sn * x.y + cs * uy);
          }
          double2 *vi = v + (int64_t)i * n, *vj = v + (int64_t)j * n;
          for (int r = sub; r < n; r += g) {
            double2 x = vi[r], y = vj[r];
            double ux = ex * y.x - ey * y.y, uy = ex * y.y + ey * y.x;
            vi[r] = make_double2(cs * x.x - sn * ux, cs * x.y - sn * uy);
            vj[r] = make_double2(sn * x.x + cs * ux, sn * x.y + cs * uy);
          }
        }
        __syncthreads();
      }
    }
    if (!anyRot) break;
    __syncthreads();
  }
  double *s = sig + sOff[b];
  for (int j = slot; j < n; j += slots) {
    const double2 *cj = a + (int64_t)j * m;
    double q = 0;
    for (int r = sub; r < m; r += g) q = fma(cj[r].x, cj[r].x, fma(cj[r].y, cj[r].y, q));
    for (int o = 1; o < g; o <<= 1) q += __shfl_xor(q, o, 64);
    if (sub == 0) s[j] = sqrt(q);
  }
}

/* The same Jacobi for large matrices, spread over the chip: one launch per
 * round-robin step, workgroup (p, b) rotates pair p of big matrix b with
 * 256 threads over the rows (block reductions), so a step of a matrix with
 * 300 pairs uses 300 workgroups instead of one.  Pairs, order, rotation
 * and stopping rule are those of k_svd_jacobi; a matrix whose `active`
 * flag is 0 (no rotation in its previous sweep) is skipped, and every
 * rotation sets rot[b]. */
__global__ __launch_bounds__(256) void k_svd_step(const int32_t *__restrict__ big, const int64_t *__restrict__ aOff,
                                                  const int32_t *__restrict__ mm, const int32_t *__restrict__ nn,
                                                  double2 *__restrict__ A, const int64_t *__restrict__ vOff,
                                                  double2 *__restrict__ V, int step,
                                                  const int32_t *__restrict__ active, int32_t *__restrict__ rot) {
  __shared__ double red[4][4];
  const int bi = blockIdx.y;
  if (!active[bi]) return;
  const int b = big[bi];
  const int m = mm[b], n = nn[b];
  const int P = n + (n & 1), npairs = P / 2, pr = blockIdx.x;
  if (n < 2 || step >= P - 1 || pr >= npairs) return;
  const int p0 = rr_row(pr, step, P), p1 = rr_row(P - 1 - pr, step, P);
  const int i = min(p0, p1), j = max(p0, p1);
  if (j >= n) return;
  double2 *ci = A + aOff[b] + (int64_t)i * m, *cj = A + aOff[b] + (int64_t)j * m;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  double al = 0, be = 0, gx = 0, gy = 0;
  for (int r = tid; r < m; r += 256) {
    const double2 x = ci[r], y = cj[r];
    al = fma(x.x, x.x, fma(x.y, x.y, al));
    be = fma(y.x, y.x, fma(y.y, y.y, be));
    gx = fma(x.x, y.x, fma(x.y, y.y, gx));
    gy = fma(x.x, y.y, fma(-x.y, y.x, gy));
  }
  al = group_sum(al, 64);
  be = group_sum(be, 64);
  gx = group_sum(gx, 64);
  gy = group_sum(gy, 64);
  if (lane == 0) {
    red[w][0] = al;
    red[w][1] = be;
    red[w][2] = gx;
    red[w][3] = gy;
  }
  __syncthreads();
  al = red[0][0] + red[1][0] + red[2][0] + red[3][0];
  be = red[0][1] + red[1][1] + red[2][1] + red[3][1];
  gx = red[0][2] + red[1][2] + red[2][2] + red[3][2];
  gy = red[0][3] + red[1][3] + red[2][3] + red[3][3];
  const double eps = 2.220446049250313e-16;
  const double ag = sqrt(gx * gx + gy * gy);
  if (!(ag > 0 && ag > eps * sqrt(al * be))) return;
  if (tid == 0) rot[bi] = 1;
  const double ex = gx / ag, ey = -gy / ag;
  const double zeta = (be - al) / (2 * ag);
  const double t = (zeta >= 0 ? 1.0 : -1.0) / (fabs(zeta) + sqrt(1 + zeta * zeta));
  const double cs = 1 / sqrt(1 + t * t), sn = cs * t;
  for (int r = tid; r < m; r += 256) {
    const double2 x = ci[r], y = cj[r];
    const double ux = ex * y.x - ey * y.y, uy = ex * y.y + ey * y.x;
    ci[r] = make_double2(cs * x.x - sn * ux, cs * x.y - sn * uy);
    cj[r] = make_double2(sn * x.x + cs * ux, sn * x.y + cs * uy);
  }
  double2 *vi = V + vOff[b] + (int64_t)i * n, *vj = V + vOff[b] + (int64_t)j * n;
  for (int r = tid; r < n; r += 256) {
    const double2 x = vi[r], y = vj[r];
    const double ux = ex * y.x - ey * y.y, uy = ex * y.y + ey * y.x;
    vi[r] = make_double2(cs * x.x - sn * ux, cs * x.y - sn * uy);
    vj[r] = make_double2(sn * x.x + cs * ux, sn * x.y + cs * uy);
  }
}

/* V = I and, after the sweeps, sig[j] = |column j of A|, for big matrix
 * blockIdx.y (one workgroup row per matrix) */
__global__ __launch_bounds__(256) void k_svd_big_init(const int32_t *__restrict__ big, const int32_t *__restrict__ nn,
                                                      const int64_t *__restrict__ vOff, double2 *__restrict__ V) {
  const int b = big[blockIdx.y];
  const int n = nn[b];
  const int64_t cnt = (int64_t)n * n;
  for (int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x; e < cnt; e += (int64_t)gridDim.x * 256)
    V[vOff[b] + e] = make_double2((e % (n + 1)) == 0 ? 1.0 : 0.0, 0.0);
}

__global__ __launch_bounds__(256) void k_svd_big_norms(const int32_t *__restrict__ big,
                                                       const int64_t *__restrict__ aOff,
                                                       const int32_t *__restrict__ mm, const int32_t *__restrict__ nn,
                                                       const double2 *__restrict__ A,
                                                       const int64_t *__restrict__ sOff, double *__restrict__ sig) {
  __shared__ double red[4];
  const int b = big[blockIdx.y];
  const int m = mm[b], n = nn[b], c = blockIdx.x;
  if (c >= n) return;
  const double2 *col = A + aOff[b] + (int64_t)c * m;
  double q = 0;
  for (int r = threadIdx.x; r < m; r += 256) q = fma(col[r].x, col[r].x, fma(col[r].y, col[r].y, q));
  q = group_sum(q, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = q;
  __syncthreads();
  if (threadIdx.x == 0) sig[sOff[b] + c] = sqrt(red[0] + red[1] + red[2] + red[3]);
}

/* y -= v (v^H y) inv over rows [j, m) of one column by one wave, four
 * rows per lane in flight; with NORM, returns |y|^2 over rows (j, m) after
 * the update (the wave's total in every lane) */
template <bool NORM = false>
__device__ __forceinline__ double house_apply_wave(const double2 *__restrict__ x, double2 *__restrict__ y, int j, int m,
                                                   double inv, int lane) {
  double dx = 0, dy = 0, ex = 0, ey = 0; /* v^H y in two chains */
  int r = j + lane;
  for (; r + 192 < m; r += 256) {
    double2 v[4], t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { v[u] = x[r + 64 * u]; t[u] = y[r + 64 * u]; }
#pragma unroll
    for (int u = 0; u < 4; u += 2) {
      dx = fma(v[u].x, t[u].x, fma(v[u].y, t[u].y, dx));
      dy = fma(v[u].x, t[u].y, fma(-v[u].y, t[u].x, dy));
      ex = fma(v[u + 1].x, t[u + 1].x, fma(v[u + 1].y, t[u + 1].y, ex));
      ey = fma(v[u + 1].x, t[u + 1].y, fma(-v[u + 1].y, t[u + 1].x, ey));
    }
  }
  for (; r < m; r += 64) {
    const double2 v = x[r], t = y[r];
    dx = fma(v.x, t.x, fma(v.y, t.y, dx));
    dy = fma(v.x, t.y, fma(-v.y, t.x, dy));
  }
  dx = group_sum(dx + ex, 64) * inv;
  dy = group_sum(dy + ey, 64) * inv;
  r = j + lane;
  double nq = 0;
  for (; r + 192 < m; r += 256) {
    double2 v[4], t[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) { v[u] = x[r + 64 * u]; t[u] = y[r + 64 * u]; }
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      t[u].x -= v[u].x * dx - v[u].y * dy;
      t[u].y -= v[u].x * dy + v[u].y * dx;
      y[r + 64 * u] = t[u];
      if (NORM && r + 64 * u > j) nq = fma(t[u].x, t[u].x, fma(t[u].y, t[u].y, nq));
    }
  }
  for (; r < m; r += 64) {
    const double2 v = x[r];
    double2 t = y[r];
    t.x -= v.x * dx - v.y * dy;
    t.y -= v.x * dy + v.y * dx;
    y[r] = t;
    if (NORM && r > j) nq = fma(t.x, t.x, fma(t.y, t.y, nq));
  }
  return NORM ? group_sum(nq, 64) : 0.0;
}

/* SVD of a tall matrix (m >= 2n, n <= SVD_TALL_MAX_N), in three launches:
 *   k_tall_qr      Householder QR with column pivoting of a copy S of A
 *                  (global memory; each wave takes every 16th column of a
 *                  step), A P = Q R; R, the pivots and the reflectors' scales
 *                  go to a per-matrix scratch slot
 *   k_tall_jacobi  one wave: one-sided Jacobi on the rows of [R | I] in LDS
 *                  (jacobi_rows, as for the re-expansion blocks): G R = R'
 *                  with orthogonal rows sigma_i v_i^H, G accumulated
 *   k_tall_w       sigma_i = |row i of R'|, V = P [conj(R') / sigma]^T, and
 *                  W = Q [sigma_i conj(G[i, :]); 0] with Q in compact WY
 *                  form (one pass over the reflectors), so A V = W and W's
 *                  columns keep their mutual orthogonality relative to
 *                  their own norms
 * The pivoted QR grades R's rows, so the row Jacobi needs few sweeps; the
 * rotations touch n-row operands instead of m. */
#define SVD_TALL_MAX_N 70
#define SVD_TALL_T 1024 /* threads of the QR and W launches */
__host__ __device__ __forceinline__ int tall_ld(int n) { return 2 * n + 1; }
__host__ __device__ __forceinline__ size_t svd_tall_lds(int n) { return (size_t)n * tall_ld(n) * 16; }
__host__ __device__ __forceinline__ size_t svd_tall_w_lds(int n) { return (size_t)2 * n * n * 16; }
/* scratch slot of one matrix: [R | I] then [R' | G], n x (2n + 1) */
#define TALL_SLOT ((int64_t)SVD_TALL_MAX_N * (2 * SVD_TALL_MAX_N + 1))

__global__ __launch_bounds__(SVD_TALL_T) void k_tall_qr(const int32_t *__restrict__ idx,
                                                        const int64_t *__restrict__ aOff,
                                                        const int32_t *__restrict__ mm,
                                                        const int32_t *__restrict__ nn, const double2 *__restrict__ A,
                                                        double2 *__restrict__ S, double2 *__restrict__ RG,
                                                        int32_t *__restrict__ pivg, double *__restrict__ hinvg) {
  __shared__ double cn[SVD_TALL_MAX_N];
  __shared__ double2 rdiag[SVD_TALL_MAX_N];
  __shared__ int pv[SVD_TALL_MAX_N];
  __shared__ int qsel;
  __shared__ double qnorm;
  const int t = blockIdx.x, b = idx[t];
  const int m = mm[b], n = nn[b];
  const double2 *a = A + aOff[b];
  double2 *s = S + aOff[b];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  constexpr int NW = SVD_TALL_T / 64;
  for (int64_t e = tid; e < (int64_t)m * n; e += SVD_TALL_T) s[e] = a[e];
  for (int j = tid; j < n; j += SVD_TALL_T) pv[j] = j;
  __syncthreads();
  /* squared norms of the columns over rows j.. : taken here for j = 0, then
   * by the updates of step j - 1 */
  for (int c = w; c < n; c += NW) {
    const double2 *col = s + (int64_t)c * m;
    double q = 0;
    for (int r = lane; r < m; r += 64) q = fma(col[r].x, col[r].x, fma(col[r].y, col[r].y, q));
    q = group_sum(q, 64);
    if (lane == 0) cn[c] = q;
  }
  __syncthreads();
  for (int j = 0; j < n; ++j) {
    if (tid == 0) {
      int best = j;
      for (int c = j + 1; c < n; ++c)
        if (cn[c] > cn[best]) best = c;
      qsel = best;
      qnorm = cn[best];
    }
    __syncthreads();
    const int q = qsel;
    if (q != j) { /* whole columns: the rows above j hold R */
      double2 *cj = s + (int64_t)j * m, *cq = s + (int64_t)q * m;
      for (int r = tid; r < m; r += SVD_TALL_T) {
        const double2 u = cj[r];
        cj[r] = cq[r];
        cq[r] = u;
      }
      if (tid == 0) {
        const int u = pv[j];
        pv[j] = pv[q];
        pv[q] = u;
        cn[q] = cn[j];
      }
    }
    __syncthreads();
    double2 *x = s + (int64_t)j * m;
    const double sg = sqrt(qnorm);
    const double2 x0 = x[j];
    const double ax0 = hypot(x0.x, x0.y);
    const double phx = ax0 > 0 ? x0.x / ax0 : 1.0, phy = ax0 > 0 ? x0.y / ax0 : 0.0;
    __syncthreads();
    if (tid == 0) {
      hinvg[(int64_t)t * SVD_TALL_MAX_N + j] = sg > 0 ? 1.0 / (sg * (sg + ax0)) : 0.0;
      rdiag[j] = sg > 0 ? make_double2(-phx * sg, -phy * sg) : x0;
    }
    if (sg > 0) {
      if (tid == 0) x[j] = make_double2(x0.x + phx * sg, x0.y + phy * sg);
      __syncthreads();
      const double inv = 1.0 / (sg * (sg + ax0));
      for (int c = j + 1 + w; c < n; c += NW) {
        const double q2 = house_apply_wave<true>(x, s + (int64_t)c * m, j, m, inv, lane);
        if (lane == 0) cn[c] = q2;
      }
    } else {
      for (int c = j + 1 + w; c < n; c += NW) /* no reflector: drop row j from the norms */
        if (lane == 0) cn[c] = fmax(cn[c] - (s[(int64_t)c * m + j].x * s[(int64_t)c * m + j].x +
                                             s[(int64_t)c * m + j].y * s[(int64_t)c * m + j].y), 0.0);
    }
    __syncthreads();
  }
  /* [R | I], row-major with leading dimension 2n + 1 */
  const int ld = tall_ld(n);
  double2 *rg = RG + (int64_t)t * TALL_SLOT;
  for (int e = tid; e < n * 2 * n; e += SVD_TALL_T) {
    const int i = e / (2 * n), c = e - i * 2 * n;
    double2 v;
    if (c < n) v = c > i ? s[(int64_t)c * m + i] : (c == i ? rdiag[i] : make_double2(0, 0));
    else v = make_double2(c - n == i ? 1.0 : 0.0, 0.0);
    rg[i * ld + c] = v;
  }
  for (int j = tid; j < n; j += SVD_TALL_T) pivg[(int64_t)t * SVD_TALL_MAX_N + j] = pv[j];
}

__global__ __launch_bounds__(64) void k_tall_jacobi(const int32_t *__restrict__ idx, const int32_t *__restrict__ nn,
                                                    double2 *__restrict__ RG) {
  extern __shared__ double2 lds[];
  const int t = blockIdx.x, n = nn[idx[t]], ld = tall_ld(n);
  double2 *rg = RG + (int64_t)t * TALL_SLOT;
  for (int e = threadIdx.x; e < n * ld; e += 64) lds[e] = rg[e];
  __syncthreads();
  const int npairs = (n + (n & 1)) / 2;
  int Lp = 64;
  while (Lp > 1 && Lp * npairs > 64) Lp >>= 1;
  switch (Lp) {
    case 8: jacobi_rows<8, 64>(lds, n, 2 * n, ld, nullptr); break;
    case 4: jacobi_rows<4, 64>(lds, n, 2 * n, ld, nullptr); break;
    case 2: jacobi_rows<2, 64>(lds, n, 2 * n, ld, nullptr); break;
    case 1: jacobi_rows<1, 64>(lds, n, 2 * n, ld, nullptr); break;
    case 16: jacobi_rows<16, 64>(lds, n, 2 * n, ld, nullptr); break;
    case 32: jacobi_rows<32, 64>(lds, n, 2 * n, ld, nullptr); break;
    default: jacobi_rows<64, 64>(lds, n, 2 * n, ld, nullptr); break;
  }
  __syncthreads();
  for (int e = threadIdx.x; e < n * ld; e += 64) rg[e] = lds[e];
}

__global__ __launch_bounds__(SVD_TALL_T) void k_tall_w(const int32_t *__restrict__ idx,
                                                       const int64_t *__restrict__ aOff,
                                                       const int32_t *__restrict__ mm, const int32_t *__restrict__ nn,
                                                       double2 *__restrict__ A, const double2 *__restrict__ S,
                                                       const double2 *__restrict__ RG, const int32_t *__restrict__ pivg,
                                                       const double *__restrict__ hinvg,
                                                       const int64_t *__restrict__ vOff, double2 *__restrict__ Vout,
                                                       const int64_t *__restrict__ sOff, double *__restrict__ sig) {
  __shared__ double sgm[SVD_TALL_MAX_N];
  extern __shared__ double2 lds[]; /* 2 n^2: T^{-1} and Z */
  const int t = blockIdx.x, b = idx[t];
  const int m = mm[b], n = nn[b], ld = tall_ld(n);
  double2 *a = A + aOff[b];
  const double2 *s = S + aOff[b], *rg = RG + (int64_t)t * TALL_SLOT;
  const int32_t *pv = pivg + (int64_t)t * SVD_TALL_MAX_N;
  const double *hinv = hinvg + (int64_t)t * SVD_TALL_MAX_N;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  constexpr int NW = SVD_TALL_T / 64;
  for (int i = w; i < n; i += NW) {
    double q = 0;
    for (int c = lane; c < n; c += 64) q = fma(rg[i * ld + c].x, rg[i * ld + c].x, fma(rg[i * ld + c].y, rg[i * ld + c].y, q));
    q = group_sum(q, 64);
    if (lane == 0) {
      sgm[i] = sqrt(q);
      sig[sOff[b] + i] = sqrt(q);
    }
  }
  __syncthreads();
  /* V column i = P conj(row i of R') / sigma_i */
  double2 *vo = Vout + vOff[b];
  for (int e = tid; e < n * n; e += SVD_TALL_T) {
    const int i = e / n, j = e - i * n;
    const double inv = sgm[i] > 0 ? 1.0 / sgm[i] : 0.0;
    const double2 r = rg[i * ld + j];
    vo[(int64_t)i * n + pv[j]] = make_double2(r.x * inv, -r.y * inv);
  }
  /* W = Q X, X = [sigma_c conj(G[c, :])^T; 0] (m x n, zero below row n),
   * Q = H_0 ... H_{n-1}, H_j = I - tau_j v_j v_j^H, v_j = column j of S on
   * rows >= j, tau_j = hinv[j] (0: no reflector).  Compact WY form:
   * Q = I - Y T Y^H with T^{-1} = diag(1/tau) + the strict upper part of
   * Y^H Y, so W = X - Y Z' with T^{-1} Z' = Y^H X.  Y^H X only needs the
   * first n rows of Y; each column of W goes through the same linear steps
   * on its own, so it keeps its accuracy relative to its own norm. */
  double2 *Ti = lds, *Z = lds + n * n; /* row-major n x n each */
#define TALL_X(r, c) make_double2(sgm[c] * rg[(c) * ld + n + (r)].x, -sgm[c] * rg[(c) * ld + n + (r)].y)
  for (int e = tid; e < n * n; e += SVD_TALL_T) { /* Z = Y^H X */
    const int j = e / n, c = e - j * n;
    double zx = 0, zy = 0;
    if (hinv[j] != 0)
      for (int r = j; r < n; ++r) {
        const double2 v = s[(int64_t)j * m + r], x = TALL_X(r, c);
        zx = fma(v.x, x.x, fma(v.y, x.y, zx)); /* conj(v) x */
        zy = fma(v.x, x.y, fma(-v.y, x.x, zy));
      }
    Z[e] = make_double2(zx, zy);
    const int i = j, jj = c; /* diagonal and lower part of T^{-1} */
    if (i >= jj) Ti[e] = make_double2(i == jj ? (hinv[i] != 0 ? 1.0 / hinv[i] : 1.0) : 0.0, 0.0);
  }
  { /* strict upper part of Y^H Y: 4 x 4 tiles of (i, j), one wave each */
    const int nb4 = (n + 3) / 4, ntiles = nb4 * (nb4 + 1) / 2;
    for (int tl = w; tl < ntiles; tl += NW) {
      int ib = 0, rem = tl; /* tile tl -> (ib, jb), ib <= jb, row by row */
      while (rem >= nb4 - ib) {
        rem -= nb4 - ib;
        ++ib;
      }
      const int jb = ib + rem, i0 = 4 * ib, j0 = 4 * jb;
      double ax[4][4], ay[4][4];
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) ax[p][q] = ay[p][q] = 0;
      for (int r = j0 + lane; r < m; r += 64) {
        double2 vi[4], vj[4];
#pragma unroll
        for (int p = 0; p < 4; ++p) {
          const int i = i0 + p, j = j0 + p;
          vi[p] = (i < n && r >= i) ? s[(int64_t)i * m + r] : make_double2(0, 0);
          vj[p] = (j < n && r >= j) ? s[(int64_t)j * m + r] : make_double2(0, 0);
        }
#pragma unroll
        for (int p = 0; p < 4; ++p)
#pragma unroll
          for (int q = 0; q < 4; ++q) {
            ax[p][q] = fma(vi[p].x, vj[q].x, fma(vi[p].y, vj[q].y, ax[p][q]));
            ay[p][q] = fma(vi[p].x, vj[q].y, fma(-vi[p].y, vj[q].x, ay[p][q]));
          }
      }
#pragma unroll
      for (int p = 0; p < 4; ++p)
#pragma unroll
        for (int q = 0; q < 4; ++q) {
          const double gx = group_sum(ax[p][q], 64), gy = group_sum(ay[p][q], 64);
          const int i = i0 + p, j = j0 + q;
          if (lane == 0 && i < j && j < n)
            Ti[i * n + j] = (hinv[i] != 0 && hinv[j] != 0) ? make_double2(gx, gy) : make_double2(0, 0);
        }
    }
  }
  __syncthreads();
  /* T^{-1} Z' = Z, upper triangular with a real diagonal: one thread per
   * column of Z, in place */
  for (int c = tid; c < n; c += SVD_TALL_T)
    for (int j = n - 1; j >= 0; --j) {
      double zx = Z[j * n + c].x, zy = Z[j * n + c].y;
      for (int k = j + 1; k < n; ++k) {
        const double2 tk = Ti[j * n + k], zk = Z[k * n + c];
        zx -= tk.x * zk.x - tk.y * zk.y;
        zy -= tk.x * zk.y + tk.y * zk.x;
      }
      const double d = 1.0 / Ti[j * n + j].x;
      Z[j * n + c] = make_double2(zx * d, zy * d);
    }
  __syncthreads();
  /* W = X - Y Z': a thread per (row, chunk of TALL_CH columns); the lanes
   * of a wave share the column chunk, so the Z' reads are broadcasts */
  constexpr int TALL_CH = 14;
  const int nch = (n + TALL_CH - 1) / TALL_CH;
  for (int64_t e = tid; e < (int64_t)m * nch; e += SVD_TALL_T) {
    const int ch = (int)(e / m), r = (int)(e - (int64_t)ch * m), c0 = ch * TALL_CH;
    double accx[TALL_CH], accy[TALL_CH];
#pragma unroll
    for (int q = 0; q < TALL_CH; ++q) accx[q] = accy[q] = 0;
    const int jmax = r < n - 1 ? r : n - 1;
    for (int j = 0; j <= jmax; ++j) {
      if (hinv[j] == 0) continue;
      const double2 v = s[(int64_t)j * m + r];
      const double2 *zr = Z + j * n + c0;
#pragma unroll
      for (int q = 0; q < TALL_CH; ++q)
        if (c0 + q < n) {
          const double2 z = zr[q];
          accx[q] = fma(v.x, z.x, fma(-v.y, z.y, accx[q]));
          accy[q] = fma(v.x, z.y, fma(v.y, z.x, accy[q]));
        }
    }
#pragma unroll
    for (int q = 0; q < TALL_CH; ++q) {
      const int c = c0 + q;
      if (c < n) {
        const double2 x = r < n ? TALL_X(r, c) : make_double2(0, 0);
        a[(int64_t)c * m + r] = make_double2(x.x - accx[q], x.y - accy[q]);
      }
    }
  }
#undef TALL_X
}

/* ---------------------------------------------------- orthogonalization */

/* LDS of one orthogonalization job, in 16-byte units: the stacked column
 * (rows x n, column-major, odd leading dimension), the diagonal of R and
 * the Householder scales (n each). */
__host__ __device__ __forceinline__ int ortho_ld(int rows) { return rows | 1; }
__host__ __device__ __forceinline__ size_t ortho_units(int rows, int n) {
  return (size_t)ortho_ld(rows) * n + 2 * (size_t)n;
}

__device__ __forceinline__ double wave_sum(double v) { return group_sum(v, 64); }

/* Columns j+1 .. n-1 of the column-major M (leading dimension ld) minus
 * v (v^H y) iv over rows [j, rows), v = column j (a Householder reflector
 * I - iv v v^H): S lanes per column (qr_split), each over every S-th row,
 * the partial products combined by lane swaps; loads are issued in
 * batches before the waits. */
template <int S>
__device__ __forceinline__ void ortho_reflect(double2 *__restrict__ M, int ld, int j, int rows, int n, double iv) {
  const double2 *__restrict__ v = M + j * ld;
  for (int t = threadIdx.x; t < (n - j - 1) * S; t += 64) {
    double2 *__restrict__ y = M + (j + 1 + t / S) * ld;
    const int r0 = j + (t & (S - 1));
    double sr = 0, si = 0, tr = 0, ti = 0;
    int r = r0;
    for (; r + S < rows; r += 2 * S) {
      const double2 v0 = v[r], y0 = y[r], v1 = v[r + S], y1 = y[r + S];
      sr = fma(v0.x, y0.x, fma(v0.y, y0.y, sr));
      si = fma(v0.x, y0.y, fma(-v0.y, y0.x, si));
      tr = fma(v1.x, y1.x, fma(v1.y, y1.y, tr));
      ti = fma(v1.x, y1.y, fma(-v1.y, y1.x, ti));
    }
    if (r < rows) {
      const double2 v0 = v[r], y0 = y[r];
      sr = fma(v0.x, y0.x, fma(v0.y, y0.y, sr));
      si = fma(v0.x, y0.y, fma(-v0.y, y0.x, si));
    }
    sr = group_sum(sr + tr, S) * iv;
    si = group_sum(si + ti, S) * iv;
    r = r0;
    for (; r + 3 * S < rows; r += 4 * S) {
      double2 vv[4], yy[4];
#pragma unroll
      for (int u = 0; u < 4; ++u) { vv[u] = v[r + u * S]; yy[u] = y[r + u * S]; }
#pragma unroll
      for (int u = 0; u < 4; ++u) {
        yy[u].x -= sr * vv[u].x - si * vv[u].y;
        yy[u].y -= sr * vv[u].y + si * vv[u].x;
        y[r + u * S] = yy[u];
      }
    }
    for (; r < rows; r += S) {
      const double2 vr = v[r];
      double2 yr = y[r];
      yr.x -= sr * vr.x - si * vr.y;
      yr.y -= sr * vr.y + si * vr.x;
      y[r] = yr;
    }
  }
}
__device__ __forceinline__ void ortho_reflect_any(double2 *M, int ld, int j, int rows, int n, double iv) {
  switch (qr_split<64>(n - j - 1)) {
    case 4: ortho_reflect<4>(M, ld, j, rows, n, iv); break;
    case 2: ortho_reflect<2>(M, ld, j, rows, n, iv); break;
    default: ortho_reflect<1>(M, ld, j, rows, n, iv); break;
  }
}

/* col <- R col in place for the n columns of one block (R upper
 * triangular m x m, column-major in global memory): row r is
 * sum_{k >= r} R[r][k] col[k], rows ascending, so each row reads only rows
 * not yet overwritten; S lanes per column split the sum over k. */
template <int S>
__device__ __forceinline__ void ortho_rmul(double2 *__restrict__ M, int ld, int off, int m, int n,
                                           const double2 *__restrict__ R) {
  for (int t = threadIdx.x; t < n * S; t += 64) {
    double2 *col = M + (t / S) * ld + off;
    const int h = t & (S - 1);
    for (int r = 0; r < m; ++r) {
      double2 a0 = make_double2(0, 0), a1 = make_double2(0, 0);
      int k = r + h;
      for (; k + S < m; k += 2 * S) {
        const double2 r0 = R[k * m + r], r1 = R[(k + S) * m + r], c0 = col[k], c1 = col[k + S];
        a0 = cfma(r0, c0, a0);
        a1 = cfma(r1, c1, a1);
      }
      if (k < m) a0 = cfma(R[k * m + r], col[k], a0);
      const double sx = group_sum(a0.x + a1.x, S), sy = group_sum(a0.y + a1.y, S);
      if (h == 0) col[r] = make_double2(sx, sy);
    }
  }
}

/* One column block of one factor per wave: load its blocks, multiply each
 * by its row's R_in (upper triangular, in place, rows ascending), then for
 * rOut >= 0 a Householder QR of the stacked column: R goes to rOut and the
 * blocks are overwritten by the matching rows of Q, formed in place from
 * the reflectors (last reflector first). */
__global__ __launch_bounds__(64) void k_ortho(const BfkOrthoJob *__restrict__ jobs, const int32_t *__restrict__ order,
                                              const BfkOrthoItem *__restrict__ items, double2 *__restrict__ arena,
                                              const double2 *__restrict__ rIn, double2 *__restrict__ rOut) {
  extern __shared__ double2 lds[];
  const BfkOrthoJob J = jobs[order[blockIdx.x]];
  const int n = J.n, rows = J.rows, ld = ortho_ld(rows);
  const int lane = threadIdx.x;
  double2 *M = lds;
  double2 *dg = M + (size_t)ld * n;
  double *inv = (double *)(dg + n);

  int off = 0;
  for (int it = J.i0; it < J.i1; ++it) {
    const BfkOrthoItem I = items[it];
    const int m = I.m;
    const double2 *B = arena + I.aOff;
    for (int e = lane; e < m * n; e += 64) {
      const int c = e / m, r = e - c * m;
      M[c * ld + off + r] = B[e];
    }
    __syncthreads();
    if (I.rIn >= 0) {
      const double2 *R = rIn + I.rIn;
      switch (qr_split<64>(n)) {
        case 4: ortho_rmul<4>(M, ld, off, m, n, R); break;
        case 2: ortho_rmul<2>(M, ld, off, m, n, R); break;
        default: ortho_rmul<1>(M, ld, off, m, n, R); break;
      }
      __syncthreads();
    }
    off += m;
  }

  if (J.rOut >= 0) {
    /* Householder QR, reflector j: v = x + ph sig e_j, H = I - v v^H / (sig (sig + |x_j|)) */
    for (int j = 0; j < n; ++j) {
      double s = 0;
      for (int r = j + lane; r < rows; r += 64) {
        double2 a = M[j * ld + r];
        s = fma(a.x, a.x, fma(a.y, a.y, s));
      }
      s = wave_sum(s);
      const double sig = sqrt(s);
      const double2 x0 = M[j * ld + j];
      const double ax0 = hypot(x0.x, x0.y);
      const double phx = ax0 > 0 ? x0.x / ax0 : 1.0, phy = ax0 > 0 ? x0.y / ax0 : 0.0;
      const double iv = sig > 0 ? 1.0 / (sig * (sig + ax0)) : 0.0;
      __syncthreads();
      if (lane == 0) {
        if (sig > 0) M[j * ld + j] = make_double2(x0.x + phx * sig, x0.y + phy * sig);
        dg[j] = make_double2(-phx * sig, -phy * sig);
        inv[j] = iv;
      }
      __syncthreads();
      if (iv != 0) ortho_reflect_any(M, ld, j, rows, n, iv);
      __syncthreads();
    }
    double2 *Ro = rOut + J.rOut;
    for (int e = lane; e < n * n; e += 64) {
      const int c = e / n, r = e - c * n;
      Ro[e] = r < c ? M[c * ld + r] : (r == c ? dg[c] : make_double2(0, 0));
    }
    __syncthreads();
    for (int e = lane; e < n * n; e += 64) {
      const int c = e / n, r = e - c * n;
      if (r < c) M[c * ld + r] = make_double2(0, 0);
    }
    __syncthreads();
    for (int j = n - 1; j >= 0; --j) {
      const double iv = inv[j];
      if (iv != 0) ortho_reflect_any(M, ld, j, rows, n, iv);
      __syncthreads();
      const double2 v0 = M[j * ld + j];
      __syncthreads();
      for (int r = j + lane; r < rows; r += 64) {
        const double2 vr = M[j * ld + r];
        const double tx = (vr.x * v0.x + vr.y * v0.y) * iv, ty = (vr.y * v0.x - vr.x * v0.y) * iv;
        M[j * ld + r] = make_double2((r == j ? 1.0 : 0.0) - tx, -ty);
      }
      __syncthreads();
    }
  }

  off = 0;
  for (int it = J.i0; it < J.i1; ++it) {
    const BfkOrthoItem I = items[it];
    const int m = I.m;
    double2 *B = arena + I.aOff;
    for (int e = lane; e < m * n; e += 64) {
      const int c = e / m, r = e - c * m;
      B[e] = M[c * ld + off + r];
    }
    off += m;
  }
}

/* ---------------------------------------------------------------- apply */

struct BfkBufs {
  const double2 *in[4];
  double2 *out[4];
};

struct BfkBufsF {
  const float2 *in[4];
  float2 *out[4];
};

__device__ __forceinline__ float2 cfmaf(float2 a, float2 b, float2 c) {
  return make_float2(fmaf(a.x, b.x, fmaf(-a.y, b.y, c.x)), fmaf(a.x, b.y, fmaf(a.y, b.x, c.y)));
}
__device__ __forceinline__ double2 cfmaT(double2 a, double2 b, double2 c) { return cfma(a, b, c); }
__device__ __forceinline__ float2 cfmaT(float2 a, float2 b, float2 c) { return cfmaf(a, b, c); }
template <typename V> __device__ __forceinline__ V czero();
template <> __device__ __forceinline__ double2 czero<double2>() { return make_double2(0, 0); }
template <> __device__ __forceinline__ float2 czero<float2>() { return make_float2(0, 0); }
template <typename V> __device__ __forceinline__ V cadd(V a, V b) {
  a.x += b.x;
  a.y += b.y;
  return a;
}

template <typename V> struct BufsOf;
template <> struct BufsOf<double2> { typedef BfkBufs T; };
template <> struct BufsOf<float2> { typedef BfkBufsF T; };

/* compensated (Kahan) accumulation: s + (-c) carries the running sum to
 * about one rounding of the final value however long the chain */
/* add an FP32 accumulator tile into an FP64 running sum and clear it */
typedef double mf_d4 __attribute__((ext_vector_type(4)));
template <class A> __device__ __forceinline__ void fold64(mf_d4 &s, A &p) {
  s += __builtin_convertvector(p, mf_d4);
  p = A{};
}
__device__ __forceinline__ float2 cmulf(float2 a, float2 b) {
  return make_float2(fmaf(a.x, b.x, -a.y * b.y), fmaf(a.x, b.y, a.y * b.x));
}

/* Launch shape of the 1-RHS apply: waves (segments) per block and the lane
 * map (0 = rows = segment height, 1 = the next power of two); set by
 * bfk_set_apply_config. */
static int g_applyWaves = 1;
static int g_applyUnroll = 0;
/* the MFMA apply's shape: right-hand-side tiles of 16 per wave | 16 for
 * two row tiles per wave (0 = the vector-unit kernel, -1 = the defaults
 * below, the shapes the tuner picks at N = 2^20) */
static int g_mfmaCT = -1;
#define MFMA_DEFAULT_F64 (2 | 16)
#define MFMA_DEFAULT_F32 2
/* the default shape with no more column tiles than nrhs fills */
static inline int mfma_auto(int cfg, int nrhs) {
  int ct = cfg & 15;
  while (ct > 1 && 16 * (ct / 2) >= nrhs) ct /= 2;
  return (cfg & 16) | ct;
}
/* accumulation in the FP32 applies (bfk_set_compensation): 2 = FP32
 * matrix-core tiles folded into FP64 sums */
static int g_comp32 = 2;
/* the MFMA apply folds its partial FP32 tile into the FP64 sum every
 * (mask + 1) 4-column steps */
static int g_foldMask = 0;
__constant__ int g_foldMaskDev = 0;

/* One stage of the apply for one right-hand side.  Each wave of the block
 * owns one output row segment: its lanes split into R rows x G = 64/R
 * column groups (R = the segment height, at most 64; UNROLL selects the
 * next power of two instead); a lane walks its columns of every term
 * (column-major blocks, so the R lanes of a group read R consecutive
 * entries), then the G partial sums meet through lane reads and the row is
 * written once.  The walk is software-pipelined: a lane's columns go in
 * batches of two, a batch never spanning two terms, and the loads of the
 * next batch are issued before the products of the current one, so a wave
 * has two batches of loads in flight across term boundaries.  Loads never
 * branch: columns past a term's width read its last column and are masked
 * out by a select.  With LO the factors are an FP32 pair hi + lo.  With
 * COMP (FP32 only) every lane sums its products in FP64 (two chains) and
 * the lanes of a row meet in FP64, so the rounding of a stage no longer
 * grows with the length of its rows; with LO and COMP the pair is summed
 * and multiplied in FP64. */
template <typename V, bool LO, bool UNROLL, bool COMP = false>
__global__ __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(7))) void k_apply1(const BfkSeg *__restrict__ segs, int64_t numSegs,
                                                const BfkTerm *__restrict__ terms, const V *__restrict__ hi,
                                                const V *__restrict__ lo, typename BufsOf<V>::T bufs) {
  const int lane = threadIdx.x & 63;
  /* the wave's segment index as a wave-uniform value, so the segment and
   * term descriptors come through scalar loads */
  const int64_t si = (int64_t)blockIdx.x * (blockDim.x >> 6) + __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  if (si >= numSegs) return; /* whole wave; no block barrier below */
  const BfkSeg S = segs[si];
  const int m = S.m;
  /* lane map: Rm rows x Gm column groups; the tight map (rows = segment
   * height) keeps 51 of 64 lanes busy on a 17-row segment where the
   * power-of-two map keeps 34 (4.4% faster at N = 2^20) */
  int Rm;
  if (UNROLL && !COMP) {
    Rm = 64;
    while (Rm > 1 && (Rm >> 1) >= m) Rm >>= 1;
  } else {
    Rm = m < 64 ? m : 64;
  }
  const int Gm = 64 / Rm, rl = lane % Rm, gl = lane / Rm;
  /* columns per lane per batch: two FP64 batches of two fit seven waves
   * per SIMD without spills; the compensated FP32 form measured 9.2 ms
   * with two against 10.5 ms with three at N = 2^20 */
  constexpr int APPLY1_BATCH = 2;
  V *y = bufs.out[S.ybuf] + S.yOff;
  /* batches over the block terms: ceil(n / (APPLY1_BATCH Gm)) per term */
  const int span = APPLY1_BATCH * Gm;
  int batches = 0;
  for (int t = S.t0; t < S.t1; ++t) {
    const BfkTerm T = terms[t];
    if (T.aOff >= 0) batches += (T.n + span - 1) / span;
  }
  for (int rb = 0; rb < m; rb += Rm) {
    const int i = rb + rl;
    const bool live = i < m && gl < Gm;
    const int ic = min(i, m - 1); /* idle lanes load row m - 1, unused */
    V a0 = czero<V>();
    double sx0 = 0, sy0 = 0, sx1 = 0, sy1 = 0;
    for (int t = S.t0; t < S.t1; ++t) { /* identity terms */
      const BfkTerm T = terms[t];
      if (T.aOff >= 0 || !live || gl != 0) continue;
      const V v = bufs.in[T.xbuf][T.xOff + i];
      if constexpr (COMP) {
        sx0 += (double)v.x;
        sy0 += (double)v.y;
      } else {
        a0 = cadd(a0, v);
      }
    }
    if (batches > 0) {
      /* the load cursor: term ct, batch offset cj (uniform) */
      int ct = S.t0; /* the first block term with columns (batches > 0: one exists) */
      while (terms[ct].aOff < 0 || terms[ct].n <= 0) ++ct;
      BfkTerm CT = terms[ct];
      int cj = 0, issued = 0;
      V xa[2][APPLY1_BATCH], aa[2][APPLY1_BATCH], la[2][APPLY1_BATCH];
      int sn[2], sj[2];
      auto issue = [&](int s) {
        const V *x = bufs.in[CT.xbuf] + CT.xOff;
        const int64_t ab = CT.aOff + CT.rowOff + ic, lda = CT.lda;
        const int n = CT.n;
#pragma unroll
        for (int u = 0; u < APPLY1_BATCH; ++u) {
          const int j = min(cj + gl + u * Gm, n - 1);
          xa[s][u] = x[j];
          aa[s][u] = hi[ab + (int64_t)j * lda];
          if (LO) la[s][u] = lo[ab + (int64_t)j * lda];
        }
        sn[s] = n;
        sj[s] = cj;
        if (++issued < batches) { /* past the last batch the cursor stays put */
          cj += span;
          if (cj >= n) {
            cj = 0;
            do ++ct;
            while (terms[ct].aOff < 0 || terms[ct].n <= 0);
            CT = terms[ct];
          }
        }
      };
      auto compute = [&](int s) {
#pragma unroll
        for (int u = 0; u < APPLY1_BATCH; ++u) { /* selects, not branches: the loads stay above */
          const bool ok = live && sj[s] + gl + u * Gm < sn[s];
          if constexpr (COMP) {
            double px, py;
            if constexpr (LO) { /* hi + lo and the product in FP64: the pair's low part survives */
              const double ar = (double)aa[s][u].x + (double)la[s][u].x;
              const double ai = (double)aa[s][u].y + (double)la[s][u].y;
              const double xr = (double)xa[s][u].x, xi = (double)xa[s][u].y;
              px = fma(ar, xr, -ai * xi);
              py = fma(ar, xi, ai * xr);
            } else {
              const V p = cmulf(aa[s][u], xa[s][u]);
              px = (double)p.x;
              py = (double)p.y;
            }
            px = ok ? px : 0.0;
            py = ok ? py : 0.0;
            if (u & 1) {
              sx1 += px;
              sy1 += py;
            } else {
              sx0 += px;
              sy0 += py;
            }
          } else {
            V t = cfmaT(aa[s][u], xa[s][u], a0);
            if (LO) t = cfmaT(la[s][u], xa[s][u], t);
            a0.x = ok ? t.x : a0.x;
            a0.y = ok ? t.y : a0.y;
          }
        }
      };
      /* whole turns (two batches each), then the last batch if the count
       * is odd: no exit inside a turn */
      issue(0);
      const int turns = batches >> 1;
      for (int u = 0; u < turns; ++u) {
        issue(1);
        compute(0);
        issue(0);
        compute(1);
      }
      if (batches & 1) compute(0);
    }
    if constexpr (COMP) {
      const double px = sx0 + sx1, py = sy0 + sy1;
      double ax = px, ay = py;
      for (int k = 1; k < Gm; ++k) { /* lanes rl + k Rm hold the other groups */
        ax += __shfl(px, rl + k * Rm, 64);
        ay += __shfl(py, rl + k * Rm, 64);
      }
      if (gl == 0 && i < m) {
        V out;
        out.x = (decltype(out.x))ax;
        out.y = (decltype(out.y))ay;
        y[i] = out;
      }
    } else {
      V acc = a0;
      for (int k = 1; k < Gm; ++k) {
        acc.x += __shfl(a0.x, rl + k * Rm, 64);
        acc.y += __shfl(a0.y, rl + k * Rm, 64);
      }
      if (gl == 0 && i < m) y[i] = acc;
    }
  }
}

#define APPLYN_ROWS 4

__global__ __launch_bounds__(64) void k_applyN(const BfkSeg *__restrict__ segs, const BfkTerm *__restrict__ terms,
                                               const double2 *__restrict__ arena, BfkBufs bufs, int nrhs) {
  const BfkSeg S = segs[blockIdx.x];
  const int c = blockIdx.y * 64 + threadIdx.x;
  const bool live = c < nrhs;
  const int m = S.m;
  double2 *y = bufs.out[S.ybuf];
  for (int rb = 0; rb < m; rb += APPLYN_ROWS) {
    double2 acc[APPLYN_ROWS];
    for (int q = 0; q < APPLYN_ROWS; ++q) acc[q] = make_double2(0, 0);
    const int nr = min(APPLYN_ROWS, m - rb);
    for (int t = S.t0; t < S.t1; ++t) {
      const BfkTerm T = terms[t];
      const double2 *x = bufs.in[T.xbuf];
      if (T.aOff < 0) {
        for (int q = 0; q < nr; ++q)
          if (live) {
            double2 v = x[(T.xOff + rb + q) * nrhs + c];
            acc[q].x += v.x;
            acc[q].y += v.y;
          }
      } else {
        const double2 *A = arena + T.aOff + T.rowOff + rb;
        const int64_t lda = T.lda;
        for (int j = 0; j < T.n; ++j) {
          double2 xv = live ? x[(T.xOff + j) * nrhs + c] : make_double2(0, 0);
          for (int q = 0; q < nr; ++q) acc[q] = cfma(A[j * lda + q], xv, acc[q]);
        }
      }
    }
    if (live)
      for (int q = 0; q < nr; ++q) y[(S.yOff + rb + q) * nrhs + c] = acc[q];
  }
}

/* ------------------------------------------- many right-hand sides: MFMA */

typedef float mf_f4 __attribute__((ext_vector_type(4)));
#define MFMA_SEG_GROUP 4
/* workgroups of one many-RHS stage: every (segment, column chunk) pair */
static inline unsigned mfma_grid(int64_t numSegs, int chunks) {
  const int64_t g = 8 * MFMA_SEG_GROUP;
  return (unsigned)(((numSegs + g - 1) / g) * g * chunks);
}

/* 16x16x4 MFMA on one wave, real operands.  Operand lanes: A[l%16][l/16],
 * B[l/16][l%16].  Result register r of lane l holds D[row(l,r)][l%16]
 * with row = 4r + l/16 for FP64 and 4(l/16) + r for FP32 (measured on
 * gfx950). */
struct MfmaF64 {
  typedef double S;
  typedef double2 V;
  typedef mf_d4 Acc;
  static __device__ __forceinline__ Acc mma(S a, S b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * r + (l >> 4); }
};
/* FP32 storage (factors and vectors), FP64 arithmetic on the matrix cores:
 * operands widened on load, sums kept in FP64, one rounding to FP32 per
 * output.  On MI355X the FP64 and FP32 matrix rates differ by 2x only, and
 * the many-RHS apply is bound by its vector traffic, which FP32 storage
 * halves either way. */
struct MfmaF64S32 {
  typedef double S;
  typedef float2 V;
  typedef mf_d4 Acc;
  static __device__ __forceinline__ Acc mma(S a, S b, Acc c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * r + (l >> 4); }
};
struct MfmaF32 {
  typedef float S;
  typedef float2 V;
  typedef mf_f4 Acc;
  static __device__ __forceinline__ Acc mma(S a, S b, Acc c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int l, int r) { return 4 * (l >> 4) + r; }
};

/* One stage of the apply for nrhs right-hand sides on the matrix cores:
 * one wave computes RT x 16 rows x 16*CT columns of a segment's output per
 * pass.  The complex product is four real MFMAs per 16x16x4 step (re*re -
 * im*im, re*im + im*re); with RT = 2 every input value a lane loads feeds
 * both row tiles (segments are mostly 17-22 rows high), which halves the
 * reads of the input vectors.  blockIdx.x picks the column chunk so the
 * chunks of one segment share its factor blocks in one L2.
 *
 * The operand loads are software-pipelined: a wave walks the 4-column
 * steps of all the segment's block terms as one sequence and keeps
 * MFMA_RING steps of operands in registers; the loads of the next
 * MFMA_RING - 1 steps are in flight while a step's MFMAs run, so a wave
 * waits on memory once per pass instead of once per step.  Loads never
 * branch: rows past the segment and columns past nrhs read a clamped
 * in-bounds address and only feed output entries that are not stored, and
 * the columns of a block's last step past its width read the block's last
 * column and are zeroed before use.  With COMP the matrix cores accumulate
 * FP32 tiles that start from zero at a fold group's first step and are
 * added into FP64 sums every (mask + 1) steps and at the end of every
 * term, so rounding no longer grows with the length of the rows. */
#define MFMA_RING 3
template <class M, int CT, int RT, bool LO, bool COMP>
__global__ __launch_bounds__(64) void k_applyN_mfma(const BfkSeg *__restrict__ segs,
                                                    const BfkTerm *__restrict__ terms,
                                                    const typename M::V *__restrict__ hi,
                                                    const typename M::V *__restrict__ lo,
                                                    typename BufsOf<typename M::V>::T bufs, int nrhs,
                                                    int numChunks, int64_t numSegs) {
  typedef typename M::V V;
  typedef typename M::Acc Acc;
  /* workgroups go to the 8 XCDs round-robin by id: ids b with the same
   * b % 8 share an XCD (and its L2).  A group of MFMA_SEG_GROUP adjacent
   * segments (rows that read the same input blocks) with all their column
   * chunks gets the ids 8 * slot + xcd, so both the blocks and the inputs
   * they share are read from one L2. */
  const int64_t slot = blockIdx.x >> 3;
  const int64_t per = (int64_t)MFMA_SEG_GROUP * numChunks;
  const int64_t grp = slot / per;
  const int64_t within = slot - grp * per;
  const int64_t si = (grp * 8 + (blockIdx.x & 7)) * MFMA_SEG_GROUP + within / numChunks;
  const int chunk = (int)(within - (within / numChunks) * numChunks);
  if (si >= numSegs) return;
  const BfkSeg Sg = segs[si];
  const int l = threadIdx.x;
  const int c0 = chunk * 16 * CT;
  const int m = Sg.m;
  const int lr = l & 15, lk = l >> 4;
  V *y = bufs.out[Sg.ybuf];
  int steps = 0; /* 4-column steps over the block terms */
  for (int t = Sg.t0; t < Sg.t1; ++t) {
    const BfkTerm T = terms[t];
    if (T.aOff >= 0) steps += (T.n + 3) >> 2;
  }
  int colc[CT];
  for (int q = 0; q < CT; ++q) colc[q] = min(c0 + 16 * q + lr, nrhs - 1);
  for (int rt = 0; rt < m; rt += 16 * RT) {
    /* row tiles of this pass that hold rows (uniform) */
    const int live = min(RT, (m - rt + 15) >> 4);
    Acc re[RT][CT], im[RT][CT], pre[RT][CT], pim[RT][CT];
    mf_d4 re64[RT][CT], im64[RT][CT];
    for (int u = 0; u < RT; ++u)
      for (int q = 0; q < CT; ++q)
        for (int r = 0; r < 4; ++r) {
          re[u][q][r] = im[u][q][r] = pre[u][q][r] = pim[u][q][r] = 0;
          re64[u][q][r] = im64[u][q][r] = 0;
        }
    for (int t = Sg.t0; t < Sg.t1; ++t) { /* identity terms */
      const BfkTerm T = terms[t];
      if (T.aOff >= 0) continue;
      const V *x = bufs.in[T.xbuf];
      for (int u = 0; u < RT; ++u)
        for (int q = 0; q < CT; ++q) {
          const int col = c0 + 16 * q + lr;
          for (int r = 0; r < 4; ++r) {
            const int i = rt + 16 * u + M::row(l, r);
            if (i < m && col < nrhs) {
              V v = x[(T.xOff + i) * (int64_t)nrhs + col];
              if constexpr (COMP) {
                pre[u][q][r] += v.x;
                pim[u][q][r] += v.y;
              } else {
                re[u][q][r] += v.x;
                im[u][q][r] += v.y;
              }
            }
          }
        }
      if constexpr (COMP)
        for (int u = 0; u < RT; ++u)
          for (int q = 0; q < CT; ++q) {
            fold64(re64[u][q], pre[u][q]);
            fold64(im64[u][q], pim[u][q]);
          }
    }
    if (steps > 0) {
      int arow[RT];
      for (int u = 0; u < RT; ++u) arow[u] = min(rt + 16 * u + lr, m - 1);
      /* the load cursor: term lt, its columns from lk0 */
      int lt = Sg.t0; /* the first block term with columns (steps > 0: one exists) */
      while (terms[lt].aOff < 0 || terms[lt].n <= 0) ++lt;
      BfkTerm LT = terms[lt];
      const V *LX = bufs.in[LT.xbuf];
      int lk0 = 0, issued = 0;
      V ra[MFMA_RING][RT], rl[MFMA_RING][RT], rb[MFMA_RING][CT];
      int sk0[MFMA_RING], sn[MFMA_RING];
      auto issue = [&](int s) {
        const int kk = min(lk0 + lk, LT.n - 1);
        const int64_t acol = LT.aOff + LT.rowOff + (int64_t)kk * LT.lda;
        for (int u = 0; u < RT; ++u) {
          ra[s][u] = hi[acol + arow[u]];
          if (LO) rl[s][u] = lo[acol + arow[u]];
        }
        const V *x = LX + (LT.xOff + kk) * (int64_t)nrhs;
        for (int q = 0; q < CT; ++q) rb[s][q] = x[colc[q]];
        sk0[s] = lk0;
        sn[s] = LT.n;
        if (++issued < steps) { /* past the last step the cursor stays put */
          lk0 += 4;
          if (lk0 >= LT.n) {
            lk0 = 0;
            do ++lt;
            while (terms[lt].aOff < 0 || terms[lt].n <= 0);
            LT = terms[lt];
            LX = bufs.in[LT.xbuf];
          }
        }
      };
      auto compute = [&](int s) {
        V a[RT], al[RT], b[CT];
        for (int u = 0; u < RT; ++u) {
          a[u] = ra[s][u];
          al[u] = rl[s][u];
        }
        for (int q = 0; q < CT; ++q) b[q] = rb[s][q];
        const int rem = sn[s] - sk0[s];
        if (rem < 4 && lk >= rem) {
          for (int u = 0; u < RT; ++u) {
            a[u] = czero<V>();
            if (LO) al[u] = czero<V>();
          }
          for (int q = 0; q < CT; ++q) b[q] = czero<V>();
        }
        /* with COMP a fold group's first step starts its FP32 tiles from
         * zero instead of clearing them after the fold */
        auto tile = [&](int u, auto freshTag) { /* straight-line MFMAs for either case */
          constexpr bool FRESH = decltype(freshTag)::value;
          /* with LO the pair hi + lo is summed in the matrix cores'
           * precision (FP64 for MfmaF64S32) before the products */
          typedef typename M::S S;
          const S ar = LO ? (S)a[u].x + (S)al[u].x : (S)a[u].x;
          const S ai = LO ? (S)a[u].y + (S)al[u].y : (S)a[u].y;
          for (int q = 0; q < CT; ++q) {
            Acc &dr = COMP ? pre[u][q] : re[u][q];
            Acc &di = COMP ? pim[u][q] : im[u][q];
            dr = M::mma(ar, b[q].x, FRESH ? Acc{} : dr);
            di = M::mma(ar, b[q].y, FRESH ? Acc{} : di);
            dr = M::mma(-ai, b[q].y, dr);
            di = M::mma(ai, b[q].x, di);
          }
          if constexpr (COMP) {
            if (((sk0[s] >> 2) & g_foldMaskDev) == g_foldMaskDev || rem <= 4)
              for (int q = 0; q < CT; ++q) {
                re64[u][q] += __builtin_convertvector(pre[u][q], mf_d4);
                im64[u][q] += __builtin_convertvector(pim[u][q], mf_d4);
              }
          }
        };
        const bool fresh = COMP && ((sk0[s] >> 2) & g_foldMaskDev) == 0;
        for (int u = 0; u < RT; ++u) {
          if (u >= live) break; /* uniform: a pass past the segment's rows */
          if (fresh)
            tile(u, std::true_type{});
          else
            tile(u, std::false_type{});
        }
      };
#pragma unroll
      for (int s = 0; s < MFMA_RING - 1; ++s) issue(s);
      /* whole turns of the ring, then the last steps (fewer than the
       * ring holds, already in flight): no exit inside a turn, so the loop
       * head sees the ring's loads in their steady-state order */
      const int turns = steps / MFMA_RING, tail = steps - turns * MFMA_RING;
      for (int u = 0; u < turns; ++u) {
#pragma unroll
        for (int s = 0; s < MFMA_RING; ++s) {
          issue((s + MFMA_RING - 1) % MFMA_RING);
          compute(s);
        }
      }
#pragma unroll
      for (int s = 0; s < MFMA_RING - 1; ++s)
        if (s < tail) compute(s);
    }
    for (int u = 0; u < RT; ++u)
      for (int q = 0; q < CT; ++q) {
        const int col = c0 + 16 * q + lr;
        for (int r = 0; r < 4; ++r) {
          const int i = rt + 16 * u + M::row(l, r);
          if (i < m && col < nrhs) {
            V v;
            if constexpr (COMP) {
              v.x = (decltype(v.x))re64[u][q][r];
              v.y = (decltype(v.y))im64[u][q][r];
            } else {
              v.x = (decltype(v.x))re[u][q][r];
              v.y = (decltype(v.y))im[u][q][r];
            }
            y[(Sg.yOff + i) * (int64_t)nrhs + col] = v;
          }
        }
      }
  }
}

/* the matrix-core kernel of one many-RHS stage: cfg = tiles of 16
 * right-hand sides per wave (1, 2, 4 or 8) | 16 for two row tiles per
 * wave */
template <class M, bool LO, bool COMP>
static void launch_mfma(int cfg, dim3 grid, const BfkSeg *segs, const BfkTerm *terms, const typename M::V *H,
                        const typename M::V *Lo, typename BufsOf<typename M::V>::T b, int nrhs, int chunks,
                        int64_t numSegs) {
#define BFK_MFMA_LAUNCH(C, R)                                                                                  \
  hipLaunchKernelGGL((k_applyN_mfma<M, C, R, LO, COMP>), grid, dim3(64), 0, g_stream, segs, terms, H, Lo, b, nrhs, \
                     chunks, numSegs)
  const int ct = cfg & 15;
  if ((cfg & 16) && ct != 8) {
    if (ct == 1) BFK_MFMA_LAUNCH(1, 2);
    else if (ct == 2) BFK_MFMA_LAUNCH(2, 2);
    else BFK_MFMA_LAUNCH(4, 2);
  } else {
    if (ct == 1) BFK_MFMA_LAUNCH(1, 1);
    else if (ct == 2) BFK_MFMA_LAUNCH(2, 1);
    else if (ct == 8) BFK_MFMA_LAUNCH(8, 1);
    else BFK_MFMA_LAUNCH(4, 1);
  }
#undef BFK_MFMA_LAUNCH
}


/* ------------------------------------------------------- plain copy */

/* a streaming copy of known size: the reference point for the HBM byte
 * counters (FETCH_SIZE / WRITE_SIZE) of the other kernels */
__global__ __launch_bounds__(256) void k_copy(const double2 *__restrict__ src, double2 *__restrict__ dst, int64_t n) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] = src[i];
}

/* ----------------------------------------------------------------- H0 */

__global__ void k_h0(int64_t n, const double *__restrict__ x, double2 *__restrict__ out) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) {
    double j, y;
    bf_h0(x[i], &j, &y);
    out[i] = make_double2(j, y);
  }
}

/* ============================================================ launchers */

extern "C" {

int bfk_init(void) {
  if (g_stream) return 0;
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    fprintf(stderr, "bf_kernels: no HIP device available\n");
    return -1;
  }
  BFK_CHECK(hipSetDevice(0));
  BFK_CHECK(hipGetSymbolAddress(&g_scratchPtr[0], HIP_SYMBOL(scratch)));
  BFK_CHECK(hipStreamCreateWithFlags(&g_stream, hipStreamNonBlocking));
  for (int i = 0; i < BFK_SIDE_STREAMS; ++i) {
    BFK_CHECK(hipStreamCreateWithFlags(&g_side[i], hipStreamNonBlocking));
    BFK_CHECK(hipEventCreateWithFlags(&g_join[i], hipEventDisableTiming));
  }
  BFK_CHECK(hipEventCreateWithFlags(&g_fork, hipEventDisableTiming));
  BFK_CHECK(hipStreamCreateWithFlags(&g_copy, hipStreamNonBlocking));
  BFK_CHECK(hipEventCreate(&g_ev0));
  BFK_CHECK(hipEventCreate(&g_ev1));
  BFK_CHECK(hipEventCreate(&g_evL0));
  BFK_CHECK(hipEventCreate(&g_evL1));
  hipDeviceProp_t prop;
  BFK_CHECK(hipGetDeviceProperties(&prop, 0));
  g_ldsCap = prop.sharedMemPerBlock;
  if (const char *e = getenv("BF_AMD_LSTSQ_WAVES")) {
    char *end = nullptr;
    g_lstsq2 = (size_t)strtoull(e, &end, 10);
    if (end && *end == ',') g_lstsq4 = (size_t)strtoull(end + 1, nullptr, 10);
  }
  if (g_ldsCap > 65536) {
    if (hipFuncSetAttribute((const void *)k_lstsq<64, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)g_ldsCap) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_lstsq_loaded, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)g_ldsCap) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_lstsq<128, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)g_ldsCap) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_lstsq<256, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)g_ldsCap) != hipSuccess ||
        hipFuncSetAttribute((const void *)k_ortho, hipFuncAttributeMaxDynamicSharedMemorySize,
                            (int)g_ldsCap) != hipSuccess) {
      (void)hipGetLastError();
      g_ldsCap = 65536;
    }
  }
  /* the tall SVD kernel asks only for what its largest case needs (it has
   * a few static LDS words besides) */
  /* k_tall_w also holds sigma in static LDS (SVD_TALL_MAX_N doubles) */
  g_tallOk = g_ldsCap >= svd_tall_lds(SVD_TALL_MAX_N) &&
             g_ldsCap >= svd_tall_w_lds(SVD_TALL_MAX_N) + SVD_TALL_MAX_N * 8 &&
             hipFuncSetAttribute((const void *)k_tall_jacobi, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)svd_tall_lds(SVD_TALL_MAX_N)) == hipSuccess &&
             hipFuncSetAttribute((const void *)k_tall_w, hipFuncAttributeMaxDynamicSharedMemorySize,
                                 (int)svd_tall_w_lds(SVD_TALL_MAX_N)) == hipSuccess;
  if (!g_tallOk) (void)hipGetLastError();
  return 0;
}

/* the device address and size of the current device's scratch array */
int bfk_scratch(void **ptr, size_t *bytes) {
  if (bfk_init()) return -1;
  int dev = 0;
  BFK_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= BFK_MAX_DEVICES) return -1;
  if (!g_scratchPtr[dev]) BFK_CHECK(hipGetSymbolAddress(&g_scratchPtr[dev], HIP_SYMBOL(scratch)));
  *ptr = g_scratchPtr[dev];
  *bytes = BFK_SCRATCH_BYTES;
  return 0;
}

int bfk_device_count(void) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) return 0;
  return n;
}

size_t bfk_lds_cap(void) { return g_ldsCap; }

/* the library's main stream, for the launchers of bf_rank.hip */
void *bfk_stream_handle(void) { return (void *)g_stream; }

size_t bfk_lstsq_lds_bytes(int p, int n) { return lstsq_units(p, n) * 16; }

/* One freed block of 16 GiB or more is kept for the next request it can
 * serve (one that needs at least half of it): a factorization's arena is
 * tens of GB, and a hipMalloc of that size right after the last one was
 * freed waited up to a second while the driver cleared the freed memory
 * (0.5-1.1 s at N = 2^20, against 0.4 ms).  A request that hipMalloc
 * cannot serve releases the kept block and tries again.  Every kernel
 * writes the arena entries it later reads, so a reused block needs no
 * clearing; where zeros matter the library writes them (bfk_memset). */
#define BFK_KEEP_MIN ((size_t)16 << 30)
#define BFK_MAX_BIG 16
struct BfkBigBlock {
  void *p;
  size_t bytes;
};
static BfkBigBlock g_big[BFK_MAX_BIG]; /* live blocks of BFK_KEEP_MIN and up */
static int g_numBig = 0;
static void *g_kept = nullptr;
static size_t g_keptBytes = 0;

static void big_track(void *p, size_t bytes) {
  if (g_numBig < BFK_MAX_BIG) g_big[g_numBig++] = BfkBigBlock{p, bytes};
}

void *bfk_malloc(size_t bytes) {
  if (bfk_init()) return nullptr;
  if (bytes >= BFK_KEEP_MIN && g_kept && g_keptBytes >= bytes && bytes >= g_keptBytes / 2) {
    void *p = g_kept;
    big_track(p, g_keptBytes);
    g_kept = nullptr;
    g_keptBytes = 0;
    return p;
  }
  void *p = nullptr;
  if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
    (void)hipGetLastError();
    if (!g_kept) return nullptr;
    (void)hipFree(g_kept);
    g_kept = nullptr;
    g_keptBytes = 0;
    if (hipMalloc(&p, bytes ? bytes : 16) != hipSuccess) {
      (void)hipGetLastError();
      return nullptr;
    }
  }
  if (bytes >= BFK_KEEP_MIN) big_track(p, bytes);
  return p;
}

/* hipFree, or for a large block: wait for the device (as hipFree would),
 * then keep it in place of the block kept before */
static void release(void *p) {
  for (int i = 0; i < g_numBig; ++i)
    if (g_big[i].p == p) {
      const size_t b = g_big[i].bytes;
      g_big[i] = g_big[--g_numBig];
      (void)hipDeviceSynchronize();
      if (g_kept) (void)hipFree(g_kept);
      g_kept = p;
      g_keptBytes = b;
      return;
    }
  (void)hipFree(p);
}

void bfk_free(void *p) {
  if (p) release(p);
}

/* released at the next bfk_sync, after the work queued before it is done;
 * the host does not wait here (hipFree would wait for the device) */
#define BFK_MAX_DEFERRED 16
static void *g_deferred[BFK_MAX_DEFERRED];
static int g_numDeferred = 0;
void bfk_free_async(void *p) {
  if (!p) return;
  if (g_numDeferred == BFK_MAX_DEFERRED) {
    (void)hipStreamSynchronize(g_stream);
    for (int i = 0; i < g_numDeferred; ++i) release(g_deferred[i]);
    g_numDeferred = 0;
  }
  g_deferred[g_numDeferred++] = p;
}
static void free_deferred(void) {
  for (int i = 0; i < g_numDeferred; ++i) release(g_deferred[i]);
  g_numDeferred = 0;
}

/* pageable uploads in pieces of 256 MiB: one multi-GB pageable copy ran at
 * a quarter of the rate of the same bytes in pieces */
int bfk_h2d(void *dst, const void *src, size_t bytes) {
  if (bfk_init()) return -1;
  const size_t piece = (size_t)256 << 20;
  for (size_t off = 0; off < bytes; off += piece) {
    const size_t n = bytes - off < piece ? bytes - off : piece;
    BFK_CHECK(hipMemcpyAsync((char *)dst + off, (const char *)src + off, n, hipMemcpyHostToDevice, g_stream));
  }
  BFK_CHECK(hipStreamSynchronize(g_stream));
  return 0;
}

/* host -> device on the copy stream: waits for this copy only, not for
 * kernels queued on the compute streams */
int bfk_h2d_side(void *dst, const void *src, size_t bytes) {
  if (bfk_init()) return -1;
  BFK_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, g_copy));
  BFK_CHECK(hipStreamSynchronize(g_copy));
  return 0;
}

int bfk_d2h(void *dst, const void *src, size_t bytes) {
  if (bfk_init()) return -1;
  BFK_CHECK(hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, g_stream));
  BFK_CHECK(hipStreamSynchronize(g_stream));
  return 0;
}

int bfk_memset(void *dst, size_t bytes) {
  if (bfk_init()) return -1;
  BFK_CHECK(hipMemsetAsync(dst, 0, bytes, g_stream));
  return 0;
}

static int collect_pending(void) {
  if (!g_lPending) return 0;
  g_lPending = 0;
  BFK_CHECK(hipEventSynchronize(g_evL1));
  float ms = 0;
  BFK_CHECK(hipEventElapsedTime(&ms, g_evL0, g_evL1));
  g_times[1] = ms;
  g_sum[1] += ms;
  g_cnt[1] += 1;
  return 0;
}

int bfk_sync(void) {
  if (bfk_init()) return -1;
  BFK_CHECK(hipStreamSynchronize(g_stream));
  free_deferred();
  return collect_pending();
}

void bfk_get_times(double out[8]) {
  (void)collect_pending();
  for (int i = 0; i < 8; ++i) out[i] = g_times[i];
}

void bfk_get_time_sums(double sums[8], double counts[8]) {
  (void)collect_pending();
  for (int i = 0; i < 8; ++i) {
    sums[i] = g_sum[i];
    counts[i] = g_cnt[i];
  }
}

int bfk_set_apply_config(int wavesPerBlock, int unroll, int mfmaTiles) {
  if (wavesPerBlock != 1 && wavesPerBlock != 2 && wavesPerBlock != 4) return -1;
  const int ct = mfmaTiles & 15;
  if (mfmaTiles != -1 && mfmaTiles != 0 && ((mfmaTiles & ~31) || (ct != 1 && ct != 2 && ct != 4 && ct != 8) ||
                                            mfmaTiles == (8 | 16)))
    return -1;
  g_applyWaves = wavesPerBlock;
  g_applyUnroll = unroll ? 1 : 0;
  g_mfmaCT = mfmaTiles;
  return 0;
}

void bfk_reset_times(void) {
  for (int i = 0; i < 8; ++i) g_sum[i] = g_cnt[i] = 0;
}

static int time_begin(void) {
  BFK_CHECK(hipEventRecord(g_ev0, g_stream));
  return 0;
}
static int time_end(int slot) {
  BFK_CHECK(hipEventRecord(g_ev1, g_stream));
  BFK_CHECK(hipEventSynchronize(g_ev1));
  float ms = 0;
  BFK_CHECK(hipEventElapsedTime(&ms, g_ev0, g_ev1));
  g_times[slot] = ms;
  g_sum[slot] += ms;
  g_cnt[slot] += 1;
  return 0;
}

/* kernel blocks; jobs already in device memory */
int bfk_sample(const BfkKernelJob *dJobs, int64_t numJobs, const double *dSrcXY, const double *dTgtXY,
               const double *dSrcN, const double *dTgtN, BfkLayer layer, double *dArena) {
  if (bfk_init()) return -1;
  if (numJobs == 0) {
    g_times[0] = 0;
    return 0;
  }
  if (time_begin()) return -1;
  const int64_t maxGrid = 1 << 30;
  for (int64_t j0 = 0; j0 < numJobs; j0 += maxGrid) {
    int64_t nj = numJobs - j0 < maxGrid ? numJobs - j0 : maxGrid;
    hipLaunchKernelGGL(k_sample, dim3((unsigned)nj), dim3(SAMPLE_THREADS), 0, g_stream, dJobs + j0,
                       (const double2 *)dSrcXY, (const double2 *)dTgtXY, (const double2 *)dSrcN,
                       (const double2 *)dTgtN, layer, (double2 *)dArena);
    BFK_CHECK(hipGetLastError());
  }
  return time_end(0);
}

/* Re-expansion blocks: bfk_lstsq_begin, then any number of
 * bfk_lstsq_launch calls (each over consecutive jobs of one LDS class,
 * already in device memory), then bfk_lstsq_end.  The launches return at
 * once; the time from begin to end is read at the next bfk_sync. */
/* scratch for the jobs above the LDS capacity: allocated once per
 * compression (hipMalloc between launches would stall them), reused chunk
 * after chunk by launches that all go to g_stream */
static double2 *g_lstsqWork = nullptr;
static size_t g_lstsqWorkStride = 0; /* bytes per job slot */
#define BFK_SCRATCH_CHUNK 4096

int bfk_lstsq_begin(size_t maxOversizeBytes, int64_t numOversize) {
  if (bfk_init()) return -1;
  if (collect_pending()) return -1;
  g_lstsqWork = nullptr;
  g_lstsqWorkStride = 0;
  if (numOversize > 0 && maxOversizeBytes > 0) {
    const int64_t slots = numOversize < BFK_SCRATCH_CHUNK ? numOversize : BFK_SCRATCH_CHUNK;
    BFK_CHECK(hipMalloc((void **)&g_lstsqWork, (size_t)slots * maxOversizeBytes));
    g_lstsqWorkStride = maxOversizeBytes;
  }
  BFK_CHECK(hipEventRecord(g_evL0, g_stream));
  /* the side streams start after everything queued so far on g_stream */
  BFK_CHECK(hipEventRecord(g_fork, g_stream));
  for (int i = 0; i < BFK_SIDE_STREAMS; ++i) BFK_CHECK(hipStreamWaitEvent(g_side[i], g_fork, 0));
  g_nextStream = 0;
  return 0;
}

/* One launch over `numJobs` jobs of one LDS class (ldsBytes per job):
 * dOrder[0..numJobs) indexes dJobs, or with dOrder NULL the jobs are
 * consecutive.  Launches rotate over g_stream and the side streams: a class of a
 * few large blocks runs for milliseconds at one block's latency and low
 * occupancy, and overlaps the classes that hold most of the blocks.  A job
 * gets 1, 2 or 4 waves by its class (g_lstsq2, g_lstsq4).  Jobs above the
 * LDS capacity run 256 threads each on the scratch of bfk_lstsq_begin,
 * chunk after chunk on g_stream. */
int bfk_lstsq_launch(const BfkLstsqJob *dJobs, const int32_t *dOrder, int64_t numJobs, size_t ldsBytes,
                     const double *dSrcXY, const double *dSrcN, BfkLayer layer, double *dArena) {
  if (numJobs <= 0) return 0;
  if (ldsBytes > g_ldsCap) {
    if (!g_lstsqWork || ldsBytes > g_lstsqWorkStride) return -3; /* bfk_lstsq_begin sized the scratch */
    const int64_t stride = (int64_t)(g_lstsqWorkStride / sizeof(double2));
    for (int64_t j0 = 0; j0 < numJobs; j0 += BFK_SCRATCH_CHUNK) {
      int64_t nj = numJobs - j0 < BFK_SCRATCH_CHUNK ? numJobs - j0 : BFK_SCRATCH_CHUNK;
      hipLaunchKernelGGL((k_lstsq<256, true>), dim3((unsigned)nj), dim3(256), 0, g_stream,
                         dOrder ? dJobs : dJobs + j0, dOrder ? dOrder + j0 : (const int32_t *)nullptr,
                         (const double2 *)dSrcXY, (const double2 *)dSrcN, layer, (double2 *)dArena, g_lstsqWork,
                         stride);
      BFK_CHECK(hipGetLastError());
    }
    return 0;
  }
  hipStream_t st = g_nextStream == 0 ? g_stream : g_side[g_nextStream - 1];
  g_nextStream = (g_nextStream + 1) % (BFK_SIDE_STREAMS + 1);
  const int64_t maxGrid = 1 << 30;
  for (int64_t j0 = 0; j0 < numJobs; j0 += maxGrid) {
    int64_t nj = numJobs - j0 < maxGrid ? numJobs - j0 : maxGrid;
    const BfkLstsqJob *jb = dOrder ? dJobs : dJobs + j0;
    const int32_t *ob = dOrder ? dOrder + j0 : (const int32_t *)nullptr;
    if (ldsBytes > g_lstsq4)
      hipLaunchKernelGGL((k_lstsq<256, false>), dim3((unsigned)nj), dim3(256), ldsBytes, st, jb, ob,
                         (const double2 *)dSrcXY, (const double2 *)dSrcN, layer, (double2 *)dArena,
                         (double2 *)nullptr, (int64_t)0);
    else if (ldsBytes > g_lstsq2)
      hipLaunchKernelGGL((k_lstsq<128, false>), dim3((unsigned)nj), dim3(128), ldsBytes, st, jb, ob,
                         (const double2 *)dSrcXY, (const double2 *)dSrcN, layer, (double2 *)dArena,
                         (double2 *)nullptr, (int64_t)0);
    else
      hipLaunchKernelGGL((k_lstsq<64, false>), dim3((unsigned)nj), dim3(64), ldsBytes, st, jb, ob,
                         (const double2 *)dSrcXY, (const double2 *)dSrcN, layer, (double2 *)dArena,
                         (double2 *)nullptr, (int64_t)0);
    BFK_CHECK(hipGetLastError());
  }
  return 0;
}

/* the LDS classes (bytes per job) above which the compression runs 2 and
 * 4 waves per job; a negative value leaves that threshold as it is.  The
 * previous values go to *old2 / *old4 when given. */
void bfk_set_lstsq_waves(int64_t two, int64_t four, int64_t *old2, int64_t *old4) {
  if (old2) *old2 = (int64_t)g_lstsq2;
  if (old4) *old4 = (int64_t)g_lstsq4;
  if (two >= 0) g_lstsq2 = (size_t)two;
  if (four >= 0) g_lstsq4 = (size_t)four;
}

int bfk_lstsq_end(void) {
  bfk_free_async(g_lstsqWork);
  g_lstsqWork = nullptr;
  for (int i = 0; i < BFK_SIDE_STREAMS; ++i) {
    BFK_CHECK(hipEventRecord(g_join[i], g_side[i]));
    BFK_CHECK(hipStreamWaitEvent(g_stream, g_join[i], 0));
  }
  BFK_CHECK(hipEventRecord(g_evL1, g_stream));
  g_lPending = 1;
  return 0;
}

static inline unsigned blocks_for(int64_t n, int t) { return (unsigned)((n + t - 1) / t); }

int bfk_pair_green(int64_t n, const double *dA, const double *dB, const double *dNA, double k, double *dOut) {
  if (bfk_init()) return -1;
  if (n <= 0) return 0;
  hipLaunchKernelGGL(k_pair_green, dim3(blocks_for(n, 256)), dim3(256), 0, g_stream, n, (const double2 *)dA,
                     (const double2 *)dB, (const double2 *)dNA, k, (double2 *)dOut);
  BFK_CHECK(hipGetLastError());
  return 0;
}

int bfk_scale_rows(int64_t n, int nrhs, const double *dW, const void *dX, void *dU, int fp32) {
  if (bfk_init()) return -1;
  const int64_t e = n * nrhs;
  if (e <= 0) return 0;
  if (fp32)
    hipLaunchKernelGGL(k_scale_rows<float2>, dim3(blocks_for(e, 256)), dim3(256), 0, g_stream, n, nrhs, dW,
                       (const float2 *)dX, (float2 *)dU);
  else
    hipLaunchKernelGGL(k_scale_rows<double2>, dim3(blocks_for(e, 256)), dim3(256), 0, g_stream, n, nrhs, dW,
                       (const double2 *)dX, (double2 *)dU);
  BFK_CHECK(hipGetLastError());
  return 0;
}

int bfk_bie_post(int64_t n, int nrhs, int ncorr, double alpha, const int32_t *dCol, const double *dC,
                 const void *dX, const void *dU, void *dY, int fp32) {
  if (bfk_init()) return -1;
  const int64_t e = n * nrhs;
  if (e <= 0) return 0;
  if (fp32)
    hipLaunchKernelGGL(k_bie_post<float2>, dim3(blocks_for(e, 256)), dim3(256), 0, g_stream, n, nrhs, ncorr,
                       alpha, dCol, (const double2 *)dC, (const float2 *)dX, (const float2 *)dU, (float2 *)dY);
  else
    hipLaunchKernelGGL(k_bie_post<double2>, dim3(blocks_for(e, 256)), dim3(256), 0, g_stream, n, nrhs, ncorr,
                       alpha, dCol, (const double2 *)dC, (const double2 *)dX, (const double2 *)dU, (double2 *)dY);
  BFK_CHECK(hipGetLastError());
  return 0;
}

/* h (host, nv complex) = V^H w; partial sums per block summed on the host
 * in block order, so the result does not depend on scheduling */
static double2 *g_part = nullptr, *g_h = nullptr;
static size_t g_partCap = 0, g_hCap = 0;
#define DOT_BLOCKS 512

int bfk_dotc_multi(int64_t n, int nv, const double *dV, int64_t ldv, const double *dW, double *hOut) {
  if (bfk_init()) return -1;
  if (nv <= 0) return 0;
  const size_t need = (size_t)DOT_BLOCKS * nv;
  if (need > g_partCap) {
    if (g_part) (void)hipFree(g_part);
    g_part = nullptr;
    BFK_CHECK(hipMalloc(&g_part, need * sizeof(double2)));
    g_partCap = need;
  }
  hipLaunchKernelGGL(k_dotc_multi, dim3(DOT_BLOCKS), dim3(256), 0, g_stream, n, nv, (const double2 *)dV, ldv,
                     (const double2 *)dW, g_part);
  BFK_CHECK(hipGetLastError());
  double2 *hp = (double2 *)malloc(need * sizeof(double2));
  if (!hp) return -2;
  if (hipMemcpyAsync(hp, g_part, need * sizeof(double2), hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
      hipStreamSynchronize(g_stream) != hipSuccess) {
    free(hp);
    return -1;
  }
  for (int j = 0; j < nv; ++j) {
    double sx = 0, sy = 0;
    for (int b = 0; b < DOT_BLOCKS; ++b) {
      sx += hp[(size_t)b * nv + j].x;
      sy += hp[(size_t)b * nv + j].y;
    }
    hOut[2 * j] = sx;
    hOut[2 * j + 1] = sy;
  }
  free(hp);
  return 0;
}

/* w = s w - V h, h given on the host */
int bfk_axpy_multi(int64_t n, int nv, const double *dV, int64_t ldv, const double *hH, double s, double *dW) {
  if (bfk_init()) return -1;
  if ((size_t)(nv > 0 ? nv : 1) > g_hCap) {
    if (g_h) (void)hipFree(g_h);
    g_h = nullptr;
    BFK_CHECK(hipMalloc(&g_h, (size_t)(nv > 0 ? nv : 1) * sizeof(double2)));
    g_hCap = (size_t)(nv > 0 ? nv : 1);
  }
  if (nv > 0) BFK_CHECK(hipMemcpyAsync(g_h, hH, (size_t)nv * sizeof(double2), hipMemcpyHostToDevice, g_stream));
  hipLaunchKernelGGL(k_axpy_multi, dim3(blocks_for(n, 256)), dim3(256), 0, g_stream, n, nv, (const double2 *)dV,
                     ldv, g_h, s, (double2 *)dW);
  BFK_CHECK(hipGetLastError());
  return 0;
}

int bfk_convert(int64_t n, const void *src, void *dst, int to32) {
  if (bfk_init()) return -1;
  if (n <= 0) return 0;
  if (to32)
    hipLaunchKernelGGL(k_f64_to_f32, dim3(blocks_for(n, 256)), dim3(256), 0, g_stream, n, (const double2 *)src,
                       (float2 *)dst);
  else
    hipLaunchKernelGGL(k_f32_to_f64, dim3(blocks_for(n, 256)), dim3(256), 0, g_stream, n, (const float2 *)src,
                       (double2 *)dst);
  BFK_CHECK(hipGetLastError());
  return 0;
}

/* batched SVD: tables and matrices already in device memory */
int bfk_svd_jacobi(int64_t batch, const int64_t *dAOff, const int32_t *dM, const int32_t *dN, double *dA,
                   const int64_t *dVOff, double *dV, const int64_t *dSOff, double *dSig, int maxSweeps) {
  if (bfk_init()) return -1;
  if (batch <= 0) return 0;
  const int64_t maxGrid = 1 << 30;
  for (int64_t b0 = 0; b0 < batch; b0 += maxGrid) {
    int64_t nb = batch - b0 < maxGrid ? batch - b0 : maxGrid;
    hipLaunchKernelGGL(k_svd_jacobi, dim3((unsigned)nb), dim3(256), 0, g_stream, dAOff + b0, dM + b0, dN + b0,
                       (double2 *)dA, dVOff + b0, (double2 *)dV, dSOff + b0, dSig, maxSweeps);
    BFK_CHECK(hipGetLastError());
  }
  BFK_CHECK(hipStreamSynchronize(g_stream));
  return 0;
}

/* the big matrices big[0 .. nb) of a batch (indices into the tables):
 * sweeps of per-step launches until a sweep rotates nothing in any of
 * them; maxN is their largest column count */
int bfk_svd_jacobi_big(int64_t nb, const int32_t *dBig, const int64_t *dAOff, const int32_t *dM, const int32_t *dN,
                       double *dA, const int64_t *dVOff, double *dV, const int64_t *dSOff, double *dSig,
                       int maxSweeps, int maxN) {
  if (bfk_init()) return -1;
  if (nb <= 0) return 0;
  if (nb > 65535) return -3;
  int32_t *dAct = nullptr, *dRot = nullptr;
  BFK_CHECK(hipMalloc((void **)&dAct, (size_t)nb * 4));
  BFK_CHECK(hipMalloc((void **)&dRot, (size_t)nb * 4));
  int32_t *act = (int32_t *)malloc((size_t)nb * 4), *rot = (int32_t *)malloc((size_t)nb * 4);
  int rc = (act && rot) ? 0 : -2;
  for (int64_t i = 0; i < nb && !rc; ++i) act[i] = 1;
  const int P = maxN + (maxN & 1);
  const unsigned vBlocks = (unsigned)((((int64_t)maxN * maxN) + 255) / 256 < 1024 ? (((int64_t)maxN * maxN) + 255) / 256 : 1024);
  if (!rc) {
    hipLaunchKernelGGL(k_svd_big_init, dim3(vBlocks > 0 ? vBlocks : 1, (unsigned)nb), dim3(256), 0, g_stream, dBig,
                       dN, dVOff, (double2 *)dV);
    if (hipGetLastError() != hipSuccess) rc = -1;
  }
  for (int sweep = 0; sweep < maxSweeps && !rc && P > 2; ++sweep) {
    int any = 0;
    for (int64_t i = 0; i < nb; ++i) any |= act[i];
    if (!any) break;
    if (hipMemcpyAsync(dAct, act, (size_t)nb * 4, hipMemcpyHostToDevice, g_stream) != hipSuccess ||
        hipMemsetAsync(dRot, 0, (size_t)nb * 4, g_stream) != hipSuccess) {
      rc = -1;
      break;
    }
    for (int step = 0; step < P - 1; ++step)
      hipLaunchKernelGGL(k_svd_step, dim3((unsigned)(P / 2), (unsigned)nb), dim3(256), 0, g_stream, dBig, dAOff, dM,
                         dN, (double2 *)dA, dVOff, (double2 *)dV, step, dAct, dRot);
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(rot, dRot, (size_t)nb * 4, hipMemcpyDeviceToHost, g_stream) != hipSuccess ||
        hipStreamSynchronize(g_stream) != hipSuccess) {
      rc = -1;
      break;
    }
    for (int64_t i = 0; i < nb; ++i) act[i] = act[i] && rot[i];
  }
  if (!rc) {
    hipLaunchKernelGGL(k_svd_big_norms, dim3((unsigned)maxN, (unsigned)nb), dim3(256), 0, g_stream, dBig, dAOff, dM,
                       dN, (const double2 *)dA, dSOff, dSig);
    if (hipGetLastError() != hipSuccess || hipStreamSynchronize(g_stream) != hipSuccess) rc = -1;
  }
  free(act);
  free(rot);
  (void)hipFree(dAct);
  (void)hipFree(dRot);
  return rc;
}

int bfk_svd_tall_max_n(void) { return g_tallOk ? SVD_TALL_MAX_N : 0; }

/* tall matrices idx[0 .. nb) of a batch (m >= 2n, n <= bfk_svd_tall_max_n()):
 * k_tall_qr, k_tall_jacobi, k_tall_w, one workgroup per matrix each; dS is
 * scratch of the size of dA; maxSweeps is not used (the row Jacobi stops by
 * its own test, at most 40 sweeps) */
int bfk_svd_jacobi_tall(int64_t nb, const int32_t *dIdx, const int64_t *dAOff, const int32_t *dM, const int32_t *dN,
                        double *dA, double *dS, const int64_t *dVOff, double *dV, const int64_t *dSOff, double *dSig,
                        int maxSweeps, int maxN) {
  (void)maxSweeps;
  if (bfk_init()) return -1;
  if (nb <= 0) return 0;
  if (!g_tallOk || maxN > SVD_TALL_MAX_N || nb > (1 << 30)) return -3;
  double2 *rg = nullptr;
  int32_t *piv = nullptr;
  double *hinv = nullptr;
  BFK_CHECK(hipMalloc((void **)&rg, (size_t)nb * TALL_SLOT * 16));
  BFK_CHECK(hipMalloc((void **)&piv, (size_t)nb * SVD_TALL_MAX_N * 4));
  BFK_CHECK(hipMalloc((void **)&hinv, (size_t)nb * SVD_TALL_MAX_N * 8));
  int rc = 0;
  hipLaunchKernelGGL(k_tall_qr, dim3((unsigned)nb), dim3(SVD_TALL_T), 0, g_stream, dIdx, dAOff, dM, dN,
                     (const double2 *)dA, (double2 *)dS, rg, piv, hinv);
  if (hipGetLastError() != hipSuccess) rc = -1;
  if (!rc) {
    hipLaunchKernelGGL(k_tall_jacobi, dim3((unsigned)nb), dim3(64), svd_tall_lds(maxN), g_stream, dIdx, dN, rg);
    if (hipGetLastError() != hipSuccess) rc = -1;
  }
  if (!rc) {
    hipLaunchKernelGGL(k_tall_w, dim3((unsigned)nb), dim3(SVD_TALL_T), svd_tall_w_lds(maxN), g_stream, dIdx, dAOff,
                       dM, dN,
                       (double2 *)dA, (const double2 *)dS, (const double2 *)rg, piv, hinv, dVOff, (double2 *)dV, dSOff,
                       dSig);
    if (hipGetLastError() != hipSuccess) rc = -1;
  }
  if (hipStreamSynchronize(g_stream) != hipSuccess) rc = -1;
  (void)hipFree(rg);
  (void)hipFree(piv);
  (void)hipFree(hinv);
  return rc;
}

size_t bfk_ortho_units(int rows, int n) { return ortho_units(rows, n); }

/* orthogonalization jobs of one stage; `dOrder` lists job indices grouped
 * by LDS need as for bfk_lstsq, and every group must fit in LDS */
int bfk_ortho(const BfkOrthoJob *dJobs, const int32_t *dOrder, int numGroups, const int64_t *groupStart,
              const size_t *groupBytes, const BfkOrthoItem *dItems, double *dArena, const double *dRin,
              double *dRout) {
  if (bfk_init()) return -1;
  for (int gi = 0; gi < numGroups; ++gi)
    if (groupStart[gi + 1] > groupStart[gi] && groupBytes[gi] > g_ldsCap) return -3;
  /* the classes of one stage are independent: largest first, spread over
   * g_stream and the side streams, joined back into g_stream */
  BFK_CHECK(hipEventRecord(g_fork, g_stream));
  for (int i = 0; i < BFK_SIDE_STREAMS; ++i) BFK_CHECK(hipStreamWaitEvent(g_side[i], g_fork, 0));
  int next = 0;
  for (int gi = numGroups - 1; gi >= 0; --gi) {
    int64_t s = groupStart[gi], e = groupStart[gi + 1];
    if (e <= s) continue;
    hipStream_t st = next == 0 ? g_stream : g_side[next - 1];
    next = (next + 1) % (BFK_SIDE_STREAMS + 1);
    const int64_t maxGrid = 1 << 30;
    for (int64_t j0 = s; j0 < e; j0 += maxGrid) {
      int64_t nj = e - j0 < maxGrid ? e - j0 : maxGrid;
      hipLaunchKernelGGL(k_ortho, dim3((unsigned)nj), dim3(64), groupBytes[gi], st, dJobs, dOrder + j0, dItems,
                         (double2 *)dArena, (const double2 *)dRin, (double2 *)dRout);
      BFK_CHECK(hipGetLastError());
    }
  }
  for (int i = 0; i < BFK_SIDE_STREAMS; ++i) {
    BFK_CHECK(hipEventRecord(g_join[i], g_side[i]));
    BFK_CHECK(hipStreamWaitEvent(g_stream, g_join[i], 0));
  }
  return 0;
}

/* The launch shape of stage s: cfg < 0 takes the global settings; for one
 * right-hand side cfg = waves per block | unroll << 4, for 16 and more
 * cfg = right-hand-side tiles of 16 per MFMA wave | 16 for two row tiles
 * per wave (0 = the vector-unit kernel); between 2 and 15 right-hand sides
 * it is not used. */
static int g_capturing = 0; /* launches go into a graph: no timing events */

static int launch_stage_f64(const BfkSeg *segs, int64_t n, const BfkTerm *dTerms, const double *dArena,
                            const BfkBufs &b, int nrhs, int cfg) {
  if (n == 0) return 0;
  if (nrhs == 1) {
    const int wv = cfg >= 0 ? (cfg & 15) : g_applyWaves;
    const int unroll = cfg >= 0 ? (cfg >> 4) & 1 : g_applyUnroll;
    dim3 grid((unsigned)((n + wv - 1) / wv)), block(64 * wv);
    if (unroll)
      hipLaunchKernelGGL((k_apply1<double2, false, true>), grid, block, 0, g_stream, segs, n, dTerms,
                         (const double2 *)dArena, (const double2 *)nullptr, b);
    else
      hipLaunchKernelGGL((k_apply1<double2, false, false>), grid, block, 0, g_stream, segs, n, dTerms,
                         (const double2 *)dArena, (const double2 *)nullptr, b);
  } else if (nrhs >= 16 && (cfg >= 0 ? cfg : g_mfmaCT != 0)) {
    const int mc = cfg >= 0 ? cfg : g_mfmaCT > 0 ? g_mfmaCT : mfma_auto(MFMA_DEFAULT_F64, nrhs);
    const int ct = mc & 15;
    const int chunks = (nrhs + 16 * ct - 1) / (16 * ct);
    dim3 grid(mfma_grid(n, chunks));
    launch_mfma<MfmaF64, false, false>(mc, grid, segs, dTerms, (const double2 *)dArena, nullptr, b, nrhs, chunks, n);
  } else {
    hipLaunchKernelGGL(k_applyN, dim3((unsigned)n, (unsigned)((nrhs + 63) / 64)), dim3(64), 0, g_stream, segs,
                       dTerms, (const double2 *)dArena, b, nrhs);
  }
  BFK_CHECK(hipGetLastError());
  return 0;
}

static BfkBufs bufs_f64(const double *dX, double *dPing, double *dPong, double *dY) {
  BfkBufs b;
  b.in[BFK_BUF_X] = (const double2 *)dX;
  b.in[BFK_BUF_PING] = (const double2 *)dPing;
  b.in[BFK_BUF_PONG] = (const double2 *)dPong;
  b.in[BFK_BUF_Y] = (const double2 *)dY;
  b.out[BFK_BUF_X] = nullptr;
  b.out[BFK_BUF_PING] = (double2 *)dPing;
  b.out[BFK_BUF_PONG] = (double2 *)dPong;
  b.out[BFK_BUF_Y] = (double2 *)dY;
  return b;
}

/* one apply: `numStages` launches; stage s covers segments
 * [stageSeg[s], stageSeg[s+1]) with launch shape stageCfg[s] (NULL: the
 * global settings for every stage) */
int bfk_apply(const BfkSeg *dSegs, const int64_t *stageSeg, int numStages, const BfkTerm *dTerms,
              const double *dArena, const double *dX, double *dPing, double *dPong, double *dY, int nrhs,
              const int32_t *stageCfg) {
  if (bfk_init()) return -1;
  const BfkBufs b = bufs_f64(dX, dPing, dPong, dY);
  if (!g_capturing && time_begin()) return -1;
  for (int s = 0; s < numStages; ++s)
    if (launch_stage_f64(dSegs + stageSeg[s], stageSeg[s + 1] - stageSeg[s], dTerms, dArena, b, nrhs,
                         stageCfg ? stageCfg[s] : -1))
      return -1;
  return g_capturing ? 0 : time_end(2);
}

/* device ms per launch of stage s alone with launch shape cfg, averaged
 * over `reps` launches after one untimed launch (for the tuner; the
 * buffers hold whatever the last apply left) */
int bfk_apply_stage_ms(const BfkSeg *dSegs, const int64_t *stageSeg, int s, const BfkTerm *dTerms,
                       const double *dArena, const double *dX, double *dPing, double *dPong, double *dY, int nrhs,
                       int cfg, int reps, double *ms) {
  if (bfk_init()) return -1;
  const BfkBufs b = bufs_f64(dX, dPing, dPong, dY);
  const int64_t n = stageSeg[s + 1] - stageSeg[s];
  if (launch_stage_f64(dSegs + stageSeg[s], n, dTerms, dArena, b, nrhs, cfg)) return -1;
  BFK_CHECK(hipEventRecord(g_ev0, g_stream));
  for (int r = 0; r < reps; ++r)
    if (launch_stage_f64(dSegs + stageSeg[s], n, dTerms, dArena, b, nrhs, cfg)) return -1;
  BFK_CHECK(hipEventRecord(g_ev1, g_stream));
  BFK_CHECK(hipEventSynchronize(g_ev1));
  float t = 0;
  BFK_CHECK(hipEventElapsedTime(&t, g_ev0, g_ev1));
  *ms = (double)t / (reps > 0 ? reps : 1);
  return 0;
}

/* graphs of apply chains: begin capturing g_stream, end into an
 * executable graph, launch one (timed like an apply in `slot`), destroy */
int bfk_graph_begin(void) {
  if (bfk_init()) return -1;
  BFK_CHECK(hipStreamBeginCapture(g_stream, hipStreamCaptureModeThreadLocal));
  g_capturing = 1;
  return 0;
}

int bfk_graph_end(void **exec) {
  g_capturing = 0;
  *exec = nullptr;
  hipGraph_t g = nullptr;
  BFK_CHECK(hipStreamEndCapture(g_stream, &g));
  hipGraphExec_t e = nullptr;
  const hipError_t err = hipGraphInstantiate(&e, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  BFK_CHECK(err);
  *exec = (void *)e;
  return 0;
}

int bfk_graph_launch(void *exec, int slot) {
  if (time_begin()) return -1;
  BFK_CHECK(hipGraphLaunch((hipGraphExec_t)exec, g_stream));
  return time_end(slot);
}

void bfk_graph_destroy(void *exec) {
  if (exec) (void)hipGraphExecDestroy((hipGraphExec_t)exec);
}

/* the global launch settings as one number: a captured apply is stale once
 * it changes */
int bfk_apply_config_stamp(void) {
  return g_applyWaves | g_applyUnroll << 4 | (g_mfmaCT & 31) << 5 | g_comp32 << 10 | g_foldMask << 14;
}

int bfk_copy(const void *src, void *dst, size_t bytes) {
  if (bfk_init()) return -1;
  int64_t n = (int64_t)(bytes / 16);
  if (n == 0) return 0;
  hipLaunchKernelGGL(k_copy, dim3(8192), dim3(256), 0, g_stream, (const double2 *)src, (double2 *)dst, n);
  BFK_CHECK(hipGetLastError());
  BFK_CHECK(hipStreamSynchronize(g_stream));
  return 0;
}

int bfk_h0(int64_t n, const double *x, double *out) {
  if (bfk_init()) return -1;
  double *dx = nullptr, *dout = nullptr;
  BFK_CHECK(hipMalloc(&dx, (n ? n : 1) * sizeof(double)));
  BFK_CHECK(hipMalloc(&dout, (n ? n : 1) * 2 * sizeof(double)));
  BFK_CHECK(hipMemcpy(dx, x, n * sizeof(double), hipMemcpyHostToDevice));
  if (n > 0) {
    hipLaunchKernelGGL(k_h0, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, g_stream, n, dx, (double2 *)dout);
    BFK_CHECK(hipGetLastError());
  }
  BFK_CHECK(hipStreamSynchronize(g_stream));
  BFK_CHECK(hipMemcpy(out, dout, n * 2 * sizeof(double), hipMemcpyDeviceToHost));
  BFK_CHECK(hipFree(dx));
  BFK_CHECK(hipFree(dout));
  return 0;
}

int bfk_lstsq_loaded(int64_t batch, int p, int n, const double *A, const double *B, double *X) {
  if (bfk_init()) return -1;
  size_t lds = lstsq_units(p, n) * 16;
  if (lds > g_ldsCap || p < 1 || n < 1) return -3;
  size_t sa = (size_t)batch * p * p * 16, sb = (size_t)batch * p * n * 16;
  double *dA = nullptr, *dB = nullptr, *dX = nullptr;
  BFK_CHECK(hipMalloc(&dA, sa));
  BFK_CHECK(hipMalloc(&dB, sb));
  BFK_CHECK(hipMalloc(&dX, sb));
  BFK_CHECK(hipMemcpy(dA, A, sa, hipMemcpyHostToDevice));
  BFK_CHECK(hipMemcpy(dB, B, sb, hipMemcpyHostToDevice));
  hipLaunchKernelGGL(k_lstsq_loaded, dim3((unsigned)batch), dim3(64), lds, g_stream, (const double2 *)dA,
                     (const double2 *)dB, p, n, (double2 *)dX);
  BFK_CHECK(hipGetLastError());
  BFK_CHECK(hipStreamSynchronize(g_stream));
  BFK_CHECK(hipMemcpy(X, dX, sb, hipMemcpyDeviceToHost));
  BFK_CHECK(hipFree(dA));
  BFK_CHECK(hipFree(dB));
  BFK_CHECK(hipFree(dX));
  return 0;
}

} /* extern "C" */

/* ====================================================== FP32 apply path */

__global__ __launch_bounds__(64) void k_applyN_f32(const BfkSeg *__restrict__ segs, const BfkTerm *__restrict__ terms,
                                                   const float2 *__restrict__ hi, const float2 *__restrict__ lo,
                                                   BfkBufsF bufs, int nrhs) {
  const BfkSeg S = segs[blockIdx.x];
  const int c = blockIdx.y * 64 + threadIdx.x;
  const bool live = c < nrhs;
  const int m = S.m;
  float2 *y = bufs.out[S.ybuf];
  for (int rb = 0; rb < m; rb += APPLYN_ROWS) {
    float2 acc[APPLYN_ROWS];
    for (int q = 0; q < APPLYN_ROWS; ++q) acc[q] = make_float2(0, 0);
    const int nr = min(APPLYN_ROWS, m - rb);
    for (int t = S.t0; t < S.t1; ++t) {
      const BfkTerm T = terms[t];
      const float2 *x = bufs.in[T.xbuf];
      if (T.aOff < 0) {
        for (int q = 0; q < nr; ++q)
          if (live) {
            float2 v = x[(T.xOff + rb + q) * nrhs + c];
            acc[q].x += v.x;
            acc[q].y += v.y;
          }
      } else {
        const int64_t base = T.aOff + T.rowOff + rb;
        const int64_t lda = T.lda;
        for (int j = 0; j < T.n; ++j) {
          float2 xv = live ? x[(T.xOff + j) * nrhs + c] : make_float2(0, 0);
          for (int q = 0; q < nr; ++q) {
            if (lo) acc[q] = cfmaf(lo[base + j * lda + q], xv, acc[q]);
            acc[q] = cfmaf(hi[base + j * lda + q], xv, acc[q]);
          }
        }
      }
    }
    if (live)
      for (int q = 0; q < nr; ++q) y[(S.yOff + rb + q) * nrhs + c] = acc[q];
  }
}

__global__ void k_to_f32(const double2 *__restrict__ a, int64_t n, float2 *__restrict__ hi, float2 *__restrict__ lo) {
  int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (; i < n; i += stride) {
    double2 v = a[i];
    float2 h = make_float2((float)v.x, (float)v.y);
    hi[i] = h;
    if (lo) lo[i] = make_float2((float)(v.x - (double)h.x), (float)(v.y - (double)h.y));
  }
}

template <bool LO, bool COMP>
static void launch_apply1_f32(dim3 grid, dim3 block, const BfkSeg *segs, int64_t n, const BfkTerm *terms,
                              const float2 *H, const float2 *Lo, BfkBufsF b) {
  hipLaunchKernelGGL((k_apply1<float2, LO, false, COMP>), grid, block, 0, g_stream, segs, n, terms, H, Lo, b);
}



extern "C" {

/* accumulation of the FP32 applies: 0 plain FP32; 1 FP64 sums -- the
 * one-RHS kernel sums FP32 products in FP64, the matrix-core kernel widens
 * its FP32 operands and runs the FP64 matrix cores; 2 + mask (2 is the
 * default) the one-RHS kernel as 1, the matrix-core kernel accumulates
 * FP32 tiles and folds them into FP64 sums every (mask + 1) 4-column steps
 * (mask 0, 1, 3 or 7) */
int bfk_set_compensation(int on) {
  const int mask = on >= 2 ? on - 2 : 0;
  if (on < 0 || (mask != 0 && mask != 1 && mask != 3 && mask != 7)) return -1;
  g_comp32 = on;
  if (mask != g_foldMask) {
    if (bfk_init()) return -1;
    BFK_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(g_foldMaskDev), &mask, sizeof(int)));
    g_foldMask = mask;
  }
  return 0;
}

int bfk_to_f32(const double *dArena, int64_t numEntries, float *dHi, float *dLo) {
  if (bfk_init()) return -1;
  if (numEntries == 0) return 0;
  hipLaunchKernelGGL(k_to_f32, dim3(4096), dim3(256), 0, g_stream, (const double2 *)dArena, numEntries,
                     (float2 *)dHi, (float2 *)dLo);
  BFK_CHECK(hipGetLastError());
  return 0;
}

static int launch_stage_f32(const BfkSeg *segs, int64_t n, const BfkTerm *dTerms, const float *dArenaHi,
                            const float *dArenaLo, const BfkBufsF &b, int nrhs, int cfg) {
  if (n == 0) return 0;
  const float2 *H = (const float2 *)dArenaHi, *Lo = (const float2 *)dArenaLo;
  if (nrhs == 1) {
    const int wv = cfg >= 0 ? (cfg & 15) : g_applyWaves;
    const int unroll = cfg >= 0 ? (cfg >> 4) & 1 : g_applyUnroll;
    dim3 grid((unsigned)((n + wv - 1) / wv)), block(64 * wv);
    if (g_comp32) {
      if (Lo) launch_apply1_f32<true, true>(grid, block, segs, n, dTerms, H, Lo, b);
      else launch_apply1_f32<false, true>(grid, block, segs, n, dTerms, H, Lo, b);
    } else if (Lo) {
      launch_apply1_f32<true, false>(grid, block, segs, n, dTerms, H, Lo, b);
    } else if (unroll) {
      hipLaunchKernelGGL((k_apply1<float2, false, true>), grid, block, 0, g_stream, segs, n, dTerms, H, Lo, b);
    } else {
      launch_apply1_f32<false, false>(grid, block, segs, n, dTerms, H, Lo, b);
    }
  } else if (nrhs >= 16 && (cfg >= 0 ? cfg : g_mfmaCT != 0)) {
    const int mc = cfg >= 0 ? cfg : g_mfmaCT > 0 ? g_mfmaCT : mfma_auto(MFMA_DEFAULT_F32, nrhs);
    const int ct = mc & 15;
    const int chunks = (nrhs + 16 * ct - 1) / (16 * ct);
    dim3 grid(mfma_grid(n, chunks));
    if (Lo) { /* hi + lo summed in FP64 on load, the FP64 matrix cores */
      launch_mfma<MfmaF64S32, true, false>(mc, grid, segs, dTerms, H, Lo, b, nrhs, chunks, n);
    } else if (g_comp32 == 1) {
      launch_mfma<MfmaF64S32, false, false>(mc, grid, segs, dTerms, H, nullptr, b, nrhs, chunks, n);
    } else {
      if (g_comp32) launch_mfma<MfmaF32, false, true>(mc, grid, segs, dTerms, H, Lo, b, nrhs, chunks, n);
      else launch_mfma<MfmaF32, false, false>(mc, grid, segs, dTerms, H, Lo, b, nrhs, chunks, n);
    }
  } else {
    hipLaunchKernelGGL(k_applyN_f32, dim3((unsigned)n, (unsigned)((nrhs + 63) / 64)), dim3(64), 0, g_stream, segs,
                       dTerms, H, Lo, b, nrhs);
  }
  BFK_CHECK(hipGetLastError());
  return 0;
}

static BfkBufsF bufs_f32(const float *dX, float *dPing, float *dPong, float *dY) {
  BfkBufsF b;
  b.in[BFK_BUF_X] = (const float2 *)dX;
  b.in[BFK_BUF_PING] = (const float2 *)dPing;
  b.in[BFK_BUF_PONG] = (const float2 *)dPong;
  b.in[BFK_BUF_Y] = (const float2 *)dY;
  b.out[BFK_BUF_X] = nullptr;
  b.out[BFK_BUF_PING] = (float2 *)dPing;
  b.out[BFK_BUF_PONG] = (float2 *)dPong;
  b.out[BFK_BUF_Y] = (float2 *)dY;
  return b;
}

int bfk_apply_f32(const BfkSeg *dSegs, const int64_t *stageSeg, int numStages, const BfkTerm *dTerms,
                  const float *dArenaHi, const float *dArenaLo, const float *dX, float *dPing, float *dPong, float *dY,
                  int nrhs, const int32_t *stageCfg) {
  if (bfk_init()) return -1;
  const BfkBufsF b = bufs_f32(dX, dPing, dPong, dY);
  if (!g_capturing && time_begin()) return -1;
  for (int s = 0; s < numStages; ++s)
    if (launch_stage_f32(dSegs + stageSeg[s], stageSeg[s + 1] - stageSeg[s], dTerms, dArenaHi, dArenaLo, b, nrhs,
                         stageCfg ? stageCfg[s] : -1))
      return -1;
  return g_capturing ? 0 : time_end(3);
}

int bfk_apply_f32_stage_ms(const BfkSeg *dSegs, const int64_t *stageSeg, int s, const BfkTerm *dTerms,
                           const float *dArenaHi, const float *dArenaLo, const float *dX, float *dPing, float *dPong,
                           float *dY, int nrhs, int cfg, int reps, double *ms) {
  if (bfk_init()) return -1;
  const BfkBufsF b = bufs_f32(dX, dPing, dPong, dY);
  const int64_t n = stageSeg[s + 1] - stageSeg[s];
  if (launch_stage_f32(dSegs + stageSeg[s], n, dTerms, dArenaHi, dArenaLo, b, nrhs, cfg)) return -1;
  BFK_CHECK(hipEventRecord(g_ev0, g_stream));
  for (int r = 0; r < reps; ++r)
    if (launch_stage_f32(dSegs + stageSeg[s], n, dTerms, dArenaHi, dArenaLo, b, nrhs, cfg)) return -1;
  BFK_CHECK(hipEventRecord(g_ev1, g_stream));
  BFK_CHECK(hipEventSynchronize(g_ev1));
  float t = 0;
  BFK_CHECK(hipEventElapsedTime(&t, g_ev0, g_ev1));
  *ms = (double)t / (reps > 0 ? reps : 1);
  return 0;
}

} /* extern "C" */
