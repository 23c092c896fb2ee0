/*
 * Rank prediction and randomized truncated SVDs for gfx950.
 *
 *   k_mlp_grad     one mini-batch of the rank model (FP32 MLP 6 -> 64 -> 64
 *                  -> 64 -> 1, GELU): forward, backward and one gradient
 *                  partial per workgroup, 64-sample tiles staged in LDS
 *   k_mlp_adam     sums the partials in a fixed order and takes one Adam step
 *   k_mlp_predict  forward pass only
 *   k_sketch_*     the randomized range finder behind the streaming
 *                  factorization's truncated SVDs: a Gaussian test matrix
 *                  Omega, Y = A Omega, Q = the normalized columns of Y after
 *                  one-sided Jacobi (the SVD kernel of bf_kernels.hip),
 *                  B^H = A^H Q and the residual ||A - Q B||_F per matrix
 *
 * The model is small (8833 parameters) and its layers are 64 x 64, so the
 * training and inference kernels keep the weights and every activation
 * of a tile in LDS (139 KB of the 160 KB) and run on the VALU; a launch
 * covers a whole mini-batch.
 */
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <stdio.h>

#define RK_CHECK(call)                                                         \
  do {                                                                         \
    hipError_t e_ = (call);                                                    \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "bf_rank: %s failed: %s\n", #call, hipGetErrorString(e_)); \
      return (e_ == hipErrorOutOfMemory || e_ == hipErrorMemoryAllocation) ? -2 : -1; \
    }                                                                          \
  } while (0)

extern "C" void *bfk_stream_handle(void);
extern "C" int bfk_init(void);

/* ---------------------------------------------------------------- MLP */

#define NF 6        /* features */
#define NH 64       /* hidden width */
#define TILE 64     /* samples per tile */
#define LDH 65      /* padded row stride of the 64-wide LDS arrays */
/* parameters in global memory (unpadded, row-major W[out][in]) */
#define P_W1 0
#define P_B1 (P_W1 + NH * NF)
#define P_W2 (P_B1 + NH)
#define P_B2 (P_W2 + NH * NH)
#define P_W3 (P_B2 + NH)
#define P_B3 (P_W3 + NH * NH)
#define P_W4 (P_B3 + NH)
#define P_B4 (P_W4 + NH)
#define NPARAM (P_B4 + 1)
/* LDS layout (floats) */
#define L_W1 0
#define L_B1 (L_W1 + NH * NF)
#define L_W2 (L_B1 + NH)
#define L_B2 (L_W2 + NH * LDH)
#define L_W3 (L_B2 + NH)
#define L_B3 (L_W3 + NH * LDH)
#define L_W4 (L_B3 + NH)
#define L_B4 (L_W4 + NH)
#define L_X (L_B4 + 4)
#define L_Z1 (L_X + TILE * NF)
#define L_A1 (L_Z1 + TILE * LDH)
#define L_Z2 (L_A1 + TILE * LDH)
#define L_A2 (L_Z2 + TILE * LDH)
#define L_Z3 (L_A2 + TILE * LDH)
#define L_A3 (L_Z3 + TILE * LDH)
#define L_DY (L_A3 + TILE * LDH)
#define L_RED (L_DY + TILE)
#define L_END (L_RED + 256)
#define MLP_LDS_BYTES (L_END * 4)

static_assert(NPARAM == 8833, "parameter layout");

__device__ __forceinline__ float gelu(float z) { return 0.5f * z * (1.0f + erff(z * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_d(float z) {
  return 0.5f * (1.0f + erff(z * 0.70710678118654752f)) + z * 0.39894228040143268f * expf(-0.5f * z * z);
}

__device__ void mlp_load_params(const float *__restrict__ P, float *L) {
  const int t = threadIdx.x;
  for (int e = t; e < NH * NF; e += 256) L[L_W1 + e] = P[P_W1 + e];
  for (int e = t; e < NH * NH; e += 256) {
    const int j = e >> 6, i = e & 63;
    L[L_W2 + j * LDH + i] = P[P_W2 + e];
    L[L_W3 + j * LDH + i] = P[P_W3 + e];
  }
  if (t < NH) {
    L[L_B1 + t] = P[P_B1 + t];
    L[L_B2 + t] = P[P_B2 + t];
    L[L_B3 + t] = P[P_B3 + t];
    L[L_W4 + t] = P[P_W4 + t];
  }
  if (t == 0) L[L_B4] = P[P_B4];
}

/* forward pass of one tile whose features sit in L[L_X]: pre-activations
 * and activations of the three hidden layers; returns the output of sample
 * t for threads t < TILE */
__device__ float mlp_forward_tile(float *L) {
  const int t = threadIdx.x;
  for (int e = t; e < TILE * NH; e += 256) {
    const int s = e >> 6, j = e & 63;
    float z = L[L_B1 + j];
#pragma unroll
    for (int f = 0; f < NF; ++f) z = fmaf(L[L_W1 + j * NF + f], L[L_X + s * NF + f], z);
    L[L_Z1 + s * LDH + j] = z;
    L[L_A1 + s * LDH + j] = gelu(z);
  }
  __syncthreads();
  for (int layer = 0; layer < 2; ++layer) {
    const int w = layer ? L_W3 : L_W2, b = layer ? L_B3 : L_B2;
    const int in = layer ? L_A2 : L_A1, zo = layer ? L_Z3 : L_Z2, ao = layer ? L_A3 : L_A2;
    for (int e = t; e < TILE * NH; e += 256) {
      const int s = e >> 6, j = e & 63;
      float z = L[b + j];
#pragma unroll 16
      for (int i = 0; i < NH; ++i) z = fmaf(L[w + j * LDH + i], L[in + s * LDH + i], z);
      L[zo + s * LDH + j] = z;
      L[ao + s * LDH + j] = gelu(z);
    }
    __syncthreads();
  }
  float out = 0;
  if (t < TILE) {
    out = L[L_B4];
#pragma unroll 16
    for (int j = 0; j < NH; ++j) out = fmaf(L[L_W4 + j], L[L_A3 + t * LDH + j], out);
  }
  return out;
}

/* Backward through one hidden layer: dZ (in LDS at DZ) and the layer's
 * input activations AIN give the weight and bias gradients (this thread's
 * 16 weights e = t + 256 q and, for t < 64, bias t); dZ of the layer below
 * replaces its pre-activations ZB. */
template <int DZ, int AIN, int W, int ZB>
__device__ __forceinline__ void back_hidden(float *L, float (&gW)[16], float &gB) {
  const int t = threadIdx.x;
  /* weight e = t + 256 q is W[j0 + 4 q][i] with j0 = t / 64, i = t % 64 */
  const int j0 = t >> 6, i0 = t & 63;
#pragma unroll 2
  for (int s = 0; s < TILE; ++s) {
    const float a = L[AIN + s * LDH + i0];
#pragma unroll
    for (int q = 0; q < 16; ++q) gW[q] = fmaf(L[DZ + s * LDH + j0 + 4 * q], a, gW[q]);
  }
  if (t < NH) {
    float a = 0;
    for (int s = 0; s < TILE; ++s) a += L[DZ + s * LDH + t];
    gB += a;
  }
  for (int e = t; e < TILE * NH; e += 256) {
    const int s = e >> 6, i = e & 63;
    float d = 0;
#pragma unroll 16
    for (int j = 0; j < NH; ++j) d = fmaf(L[DZ + s * LDH + j], L[W + j * LDH + i], d);
    L[ZB + s * LDH + i] = d * gelu_d(L[ZB + s * LDH + i]);
  }
}

/* One mini-batch: samples idx[b0 .. b0 + B) of (X, y) (features already
 * standardized, targets too).  Workgroup g takes tiles g, g + G, ...;
 * its gradient of the mean squared error goes to part[g * NPARAM ...]
 * and its summed squared error to lossPart[g]. */
__global__ __launch_bounds__(256) void k_mlp_grad(const float *__restrict__ P, const float *__restrict__ X,
                                                  const float *__restrict__ Y, const int32_t *__restrict__ idx,
                                                  int64_t b0, int B, float *__restrict__ part,
                                                  float *__restrict__ lossPart) {
  extern __shared__ float L[];
  const int t = threadIdx.x, G = gridDim.x;
  mlp_load_params(P, L);
  float gW2[16], gW3[16], gW1[2] = {0, 0};
  for (int q = 0; q < 16; ++q) gW2[q] = gW3[q] = 0;
  float gB1 = 0, gB2 = 0, gB3 = 0, gW4 = 0, gB4 = 0, loss = 0;
  const float scale = 2.0f / (float)B;
  const int numTiles = (B + TILE - 1) / TILE;
  for (int tile = blockIdx.x; tile < numTiles; tile += G) {
    const int s0 = tile * TILE, ns = min(TILE, B - s0);
    __syncthreads();
    for (int e = t; e < TILE * NF; e += 256) {
      const int s = e / NF, f = e - s * NF;
      L[L_X + e] = s < ns ? X[(int64_t)idx[b0 + s0 + s] * NF + f] : 0.0f;
    }
    __syncthreads();
    const float out = mlp_forward_tile(L);
    if (t < TILE) {
      float d = 0;
      if (t < ns) {
        d = out - Y[idx[b0 + s0 + t]];
        loss = fmaf(d, d, loss);
      }
      L[L_DY + t] = scale * d;
    }
    __syncthreads();
    /* output layer: dZ3 = dy w4 gelu'(z3), in place of Z3 */
    for (int e = t; e < TILE * NH; e += 256) {
      const int s = e >> 6, j = e & 63;
      L[L_Z3 + s * LDH + j] = L[L_DY + s] * L[L_W4 + j] * gelu_d(L[L_Z3 + s * LDH + j]);
    }
    if (t < NH) {
      float a = 0;
      for (int s = 0; s < TILE; ++s) a = fmaf(L[L_DY + s], L[L_A3 + s * LDH + t], a);
      gW4 += a;
    } else if (t == NH) {
      float a = 0;
      for (int s = 0; s < TILE; ++s) a += L[L_DY + s];
      gB4 += a;
    }
    __syncthreads();
    /* hidden layers 3 and 2: weight gradients from dZ and the layer input,
     * then dZ of the layer below (in place of its Z) */
    back_hidden<L_Z3, L_A2, L_W3, L_Z2>(L, gW3, gB3);
    __syncthreads();
    back_hidden<L_Z2, L_A1, L_W2, L_Z1>(L, gW2, gB2);
    __syncthreads();
    /* first layer: dZ1 is in Z1 */
    for (int q = 0; q < 2; ++q) {
      const int e = t + 256 * q;
      if (e < NH * NF) {
        const int j = e / NF, f = e - j * NF;
        float a = 0;
        for (int s = 0; s < TILE; ++s) a = fmaf(L[L_Z1 + s * LDH + j], L[L_X + s * NF + f], a);
        gW1[q] += a;
      }
    }
    if (t < NH) {
      float a = 0;
      for (int s = 0; s < TILE; ++s) a += L[L_Z1 + s * LDH + t];
      gB1 += a;
    }
  }
  float *g = part + (int64_t)blockIdx.x * NPARAM;
#pragma unroll
  for (int q = 0; q < 16; ++q) {
    g[P_W2 + t + 256 * q] = gW2[q];
    g[P_W3 + t + 256 * q] = gW3[q];
  }
  for (int q = 0; q < 2; ++q)
    if (t + 256 * q < NH * NF) g[P_W1 + t + 256 * q] = gW1[q];
  if (t < NH) {
    g[P_B1 + t] = gB1;
    g[P_B2 + t] = gB2;
    g[P_B3 + t] = gB3;
    g[P_W4 + t] = gW4;
  } else if (t == NH) {
    g[P_B4] = gB4;
  }
  /* the tile loss of this workgroup */
  __syncthreads();
  L[L_RED + t] = loss;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) L[L_RED + t] += L[L_RED + t + o];
    __syncthreads();
  }
  if (t == 0) lossPart[blockIdx.x] = L[L_RED];
}

/* Adam on every parameter: the gradient is the sum of the G partials,
 * taken in workgroup order so that a run is reproducible.  lossAcc
 * accumulates the mini-batch losses (FP64) for the caller. */
__global__ __launch_bounds__(256) void k_mlp_adam(float *__restrict__ P, float *__restrict__ M, float *__restrict__ V,
                                                  const float *__restrict__ part, int G, float lr, float c1, float c2,
                                                  const float *__restrict__ lossPart, int B,
                                                  double *__restrict__ lossAcc) {
  const int p = blockIdx.x * blockDim.x + threadIdx.x;
  if (p == 0) {
    double l = 0;
    for (int g = 0; g < G; ++g) l += lossPart[g];
    lossAcc[0] += l / B;
    lossAcc[1] += 1.0;
  }
  if (p >= NPARAM) return;
  float g = 0;
  for (int w = 0; w < G; ++w) g += part[(int64_t)w * NPARAM + p];
  const float m = 0.9f * M[p] + 0.1f * g;
  const float v = 0.999f * V[p] + 0.001f * g * g;
  M[p] = m;
  V[p] = v;
  P[p] -= lr * (m * c1) / (sqrtf(v * c2) + 1e-8f);
}

__global__ __launch_bounds__(256) void k_mlp_predict(const float *__restrict__ P, const float *__restrict__ X,
                                                     int64_t n, float *__restrict__ out) {
  extern __shared__ float L[];
  const int t = threadIdx.x;
  mlp_load_params(P, L);
  const int64_t numTiles = (n + TILE - 1) / TILE;
  for (int64_t tile = blockIdx.x; tile < numTiles; tile += gridDim.x) {
    const int64_t s0 = tile * TILE;
    const int ns = (int)min((int64_t)TILE, n - s0);
    __syncthreads();
    for (int e = t; e < TILE * NF; e += 256) {
      const int s = e / NF;
      L[L_X + e] = s < ns ? X[s0 * NF + e] : 0.0f;
    }
    __syncthreads();
    const float o = mlp_forward_tile(L);
    if (t < ns) out[s0 + t] = o;
  }
}

/* ------------------------------------------------ randomized range finder */

/* One matrix of a sketch batch: A (m x n, column-major at aOff), the test
 * matrix Omega (n x w at oOff), Y then Q (m x w at yOff), B^H (n x w at
 * bOff); 1 <= w <= min(m, n). */
typedef struct {
  int32_t m, n, w, pad;
  int64_t aOff, oOff, yOff, bOff;
} BfkSketchJob;

__device__ __forceinline__ uint64_t mix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

/* Omega: i.i.d. standard complex Gaussians (Box-Muller on a counter hash
 * of (seed, matrix, entry)) */
__global__ __launch_bounds__(256) void k_sketch_omega(const BfkSketchJob *__restrict__ jobs, uint64_t seed,
                                                      double2 *__restrict__ Om) {
  const BfkSketchJob J = jobs[blockIdx.x];
  const uint64_t key = mix64(seed ^ mix64((uint64_t)blockIdx.x));
  const int64_t cnt = (int64_t)J.n * J.w;
  for (int64_t e = threadIdx.x; e < cnt; e += 256) {
    const uint64_t h = mix64(key + (uint64_t)e);
    const double u1 = ((double)(h >> 40) + 1.0) * 0x1.0p-24; /* (0, 1] */
    const double u2 = (double)(h & 0xFFFFFFull) * 0x1.0p-24;
    const double r = sqrt(-log(u1)); /* each part has variance 1/2 */
    double sn, cs;
    sincos(6.283185307179586 * u2, &sn, &cs);
    Om[J.oOff + e] = make_double2(r * cs, r * sn);
  }
}

/* Y = A Omega: grid (chunks of 256 entries of Y, matrices) */
__global__ __launch_bounds__(256) void k_sketch_mul(const BfkSketchJob *__restrict__ jobs,
                                                    const double2 *__restrict__ A, const double2 *__restrict__ Om,
                                                    double2 *__restrict__ Y) {
  const BfkSketchJob J = jobs[blockIdx.y];
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)J.m * J.w) return;
  const int c = (int)(e / J.m), r = (int)(e - (int64_t)c * J.m);
  const double2 *a = A + J.aOff + r, *o = Om + J.oOff + (int64_t)c * J.n;
  double sx = 0, sy = 0;
  for (int j = 0; j < J.n; ++j) {
    const double2 x = a[(int64_t)j * J.m], y = o[j];
    sx = fma(x.x, y.x, fma(-x.y, y.y, sx));
    sy = fma(x.x, y.y, fma(x.y, y.x, sy));
  }
  Y[J.yOff + e] = make_double2(sx, sy);
}

/* Q = the columns sigma_c u_c of Y (after one-sided Jacobi) divided by
 * sigma_c; directions below 1e-14 sigma_max carry no information about A
 * and are set to zero.  Grid (chunks of 256 entries, matrices). */
__global__ __launch_bounds__(256) void k_sketch_q(const BfkSketchJob *__restrict__ jobs,
                                                  const int64_t *__restrict__ sOff, const double *__restrict__ sig,
                                                  double2 *__restrict__ Y) {
  const BfkSketchJob J = jobs[blockIdx.y];
  const int64_t e = (int64_t)blockIdx.x * 256 + threadIdx.x;
  if (e >= (int64_t)J.m * J.w) return;
  const double *s = sig + sOff[blockIdx.y];
  double smax = 0;
  for (int c = 0; c < J.w; ++c) smax = fmax(smax, s[c]);
  const int c = (int)(e / J.m);
  const double sc = s[c] > 1e-14 * smax ? 1.0 / s[c] : 0.0;
  const double2 y = Y[J.yOff + e];
  Y[J.yOff + e] = make_double2(y.x * sc, y.y * sc);
}

/* B^H = A^H Q (n x w): one wave per entry, the lanes over the rows of A and
 * Q (coalesced), a wave sum; grid (entries / 4, matrices) */
__global__ __launch_bounds__(256) void k_sketch_mulh(const BfkSketchJob *__restrict__ jobs,
                                                     const double2 *__restrict__ A, const double2 *__restrict__ Q,
                                                     double2 *__restrict__ BH) {
  const BfkSketchJob J = jobs[blockIdx.y];
  const int64_t e = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (e >= (int64_t)J.n * J.w) return;
  const int lane = threadIdx.x & 63;
  const int c = (int)(e / J.n), j = (int)(e - (int64_t)c * J.n);
  const double2 *a = A + J.aOff + (int64_t)j * J.m, *q = Q + J.yOff + (int64_t)c * J.m;
  double sx = 0, sy = 0;
  for (int r = lane; r < J.m; r += 64) {
    const double2 x = a[r], y = q[r]; /* conj(x) y */
    sx = fma(x.x, y.x, fma(x.y, y.y, sx));
    sy = fma(x.x, y.y, fma(-x.y, y.x, sy));
  }
  for (int o = 32; o > 0; o >>= 1) {
    sx += __shfl_xor(sx, o, 64);
    sy += __shfl_xor(sy, o, 64);
  }
  if (lane == 0) BH[J.bOff + e] = make_double2(sx, sy);
}

#define RESID_CHUNK 4096 /* entries of A per workgroup of the residual */

/* part[b * maxChunks + chunk] = the squared entries of A - Q (B^H)^H in
 * chunk `chunk` of matrix b (the host sums them in order) */
__global__ __launch_bounds__(256) void k_sketch_resid(const BfkSketchJob *__restrict__ jobs,
                                                      const double2 *__restrict__ A, const double2 *__restrict__ Q,
                                                      const double2 *__restrict__ BH, int maxChunks,
                                                      double *__restrict__ part) {
  __shared__ double red[4];
  const BfkSketchJob J = jobs[blockIdx.y];
  const int64_t cnt = (int64_t)J.m * J.n, e0 = (int64_t)blockIdx.x * RESID_CHUNK;
  if (e0 >= cnt) return;
  const int64_t e1 = e0 + RESID_CHUNK < cnt ? e0 + RESID_CHUNK : cnt;
  double acc = 0;
  for (int64_t e = e0 + threadIdx.x; e < e1; e += 256) {
    const int j = (int)(e / J.m), r = (int)(e - (int64_t)j * J.m);
    double2 d = A[J.aOff + e];
    for (int c = 0; c < J.w; ++c) {
      const double2 q = Q[J.yOff + (int64_t)c * J.m + r], b = BH[J.bOff + (int64_t)c * J.n + j];
      /* d -= q conj(b) */
      d.x -= q.x * b.x + q.y * b.y;
      d.y -= q.y * b.x - q.x * b.y;
    }
    acc = fma(d.x, d.x, fma(d.y, d.y, acc));
  }
  for (int o = 32; o > 0; o >>= 1) acc += __shfl_xor(acc, o, 64);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = acc;
  __syncthreads();
  if (threadIdx.x == 0) part[(int64_t)blockIdx.y * maxChunks + blockIdx.x] = red[0] + red[1] + red[2] + red[3];
}

/* ============================================================ launchers */

static int g_mlpReady = 0;

static int mlp_setup(void) {
  if (g_mlpReady) return 0;
  if (bfk_init()) return -1;
  RK_CHECK(hipFuncSetAttribute((const void *)k_mlp_grad, hipFuncAttributeMaxDynamicSharedMemorySize,
                               MLP_LDS_BYTES));
  RK_CHECK(hipFuncSetAttribute((const void *)k_mlp_predict, hipFuncAttributeMaxDynamicSharedMemorySize,
                               MLP_LDS_BYTES));
  g_mlpReady = 1;
  return 0;
}

extern "C" {

int bfk_mlp_nparam(void) { return NPARAM; }

/* workgroups of one mini-batch of B samples (the partial buffer holds
 * this many gradient copies) */
int bfk_mlp_groups(int B) {
  int tiles = (B + TILE - 1) / TILE;
  return tiles < 256 ? tiles : 256;
}

/* one optimizer step over samples idx[b0 .. b0 + B); c1, c2 are Adam's
 * bias corrections 1 / (1 - beta^t) */
int bfk_mlp_step(float *dP, float *dM, float *dV, const float *dX, const float *dY, const int32_t *dIdx, int64_t b0,
                 int B, float lr, float c1, float c2, float *dPart, float *dLossPart, double *dLossAcc) {
  if (mlp_setup()) return -1;
  hipStream_t s = (hipStream_t)bfk_stream_handle();
  const int G = bfk_mlp_groups(B);
  hipLaunchKernelGGL(k_mlp_grad, dim3(G), dim3(256), MLP_LDS_BYTES, s, dP, dX, dY, dIdx, b0, B, dPart, dLossPart);
  RK_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_mlp_adam, dim3((NPARAM + 255) / 256), dim3(256), 0, s, dP, dM, dV, dPart, G, lr, c1, c2,
                     dLossPart, B, dLossAcc);
  RK_CHECK(hipGetLastError());
  return 0;
}

/* the gradient partials of samples idx[0 .. B) without a step */
int bfk_mlp_grad(const float *dP, const float *dX, const float *dY, const int32_t *dIdx, int B, float *dPart,
                 float *dLossPart) {
  if (mlp_setup()) return -1;
  hipStream_t s = (hipStream_t)bfk_stream_handle();
  hipLaunchKernelGGL(k_mlp_grad, dim3(bfk_mlp_groups(B)), dim3(256), MLP_LDS_BYTES, s, dP, dX, dY, dIdx, (int64_t)0,
                     B, dPart, dLossPart);
  RK_CHECK(hipGetLastError());
  return 0;
}

int bfk_mlp_predict(const float *dP, const float *dX, int64_t n, float *dOut) {
  if (mlp_setup()) return -1;
  if (n <= 0) return 0;
  hipStream_t s = (hipStream_t)bfk_stream_handle();
  int64_t tiles = (n + TILE - 1) / TILE;
  int G = (int)(tiles < 1024 ? tiles : 1024);
  hipLaunchKernelGGL(k_mlp_predict, dim3(G), dim3(256), MLP_LDS_BYTES, s, dP, dX, n, dOut);
  RK_CHECK(hipGetLastError());
  return 0;
}

int bfk_sketch_jobs_bytes(void) { return (int)sizeof(BfkSketchJob); }

int bfk_sketch_resid_chunks(int64_t maxEntries) { return (int)((maxEntries + RESID_CHUNK - 1) / RESID_CHUNK); }

/* the range finder up to Y: Omega and Y = A Omega (maxMW = the largest
 * m w of the batch) */
int bfk_sketch_range(int64_t batch, const void *dJobs, uint64_t seed, const double *dA, double *dOm, double *dY,
                     int64_t maxMW) {
  if (bfk_init()) return -1;
  if (batch <= 0) return 0;
  if (batch > 65535) return -3;
  hipStream_t s = (hipStream_t)bfk_stream_handle();
  const BfkSketchJob *J = (const BfkSketchJob *)dJobs;
  hipLaunchKernelGGL(k_sketch_omega, dim3((unsigned)batch), dim3(256), 0, s, J, seed, (double2 *)dOm);
  RK_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_sketch_mul, dim3((unsigned)((maxMW + 255) / 256), (unsigned)batch), dim3(256), 0, s, J,
                     (const double2 *)dA, (const double2 *)dOm, (double2 *)dY);
  RK_CHECK(hipGetLastError());
  return 0;
}

/* after the Jacobi of Y: Q, B^H and the residual partials (maxMW, maxNW,
 * maxMN the largest m w, n w, m n of the batch; part holds batch x
 * bfk_sketch_resid_chunks(maxMN) doubles) */
int bfk_sketch_project(int64_t batch, const void *dJobs, const int64_t *dSOff, const double *dSig, const double *dA,
                       double *dY, double *dBH, double *dPart, int64_t maxMW, int64_t maxNW, int64_t maxMN) {
  if (bfk_init()) return -1;
  if (batch <= 0) return 0;
  if (batch > 65535) return -3;
  hipStream_t s = (hipStream_t)bfk_stream_handle();
  const BfkSketchJob *J = (const BfkSketchJob *)dJobs;
  hipLaunchKernelGGL(k_sketch_q, dim3((unsigned)((maxMW + 255) / 256), (unsigned)batch), dim3(256), 0, s, J, dSOff,
                     dSig, (double2 *)dY);
  RK_CHECK(hipGetLastError());
  hipLaunchKernelGGL(k_sketch_mulh, dim3((unsigned)((maxNW + 3) / 4), (unsigned)batch), dim3(256), 0, s, J,
                     (const double2 *)dA, (const double2 *)dY, (double2 *)dBH);
  RK_CHECK(hipGetLastError());
  const int chunks = bfk_sketch_resid_chunks(maxMN);
  hipLaunchKernelGGL(k_sketch_resid, dim3((unsigned)chunks, (unsigned)batch), dim3(256), 0, s, J,
                     (const double2 *)dA, (const double2 *)dY, (const double2 *)dBH, chunks, dPart);
  RK_CHECK(hipGetLastError());
  return 0;
}

} /* extern "C" */
