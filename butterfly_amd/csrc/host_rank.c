/*
 * The rank model: a small FP32 MLP that predicts the numerical rank of a
 * block of the streaming factorization from the geometry of its row and
 * column nodes, trained and evaluated by the HIP kernels of bf_rank.hip.
 *
 * Inputs (bfhipRankFeatures): log column-node diameter, log row-node
 * diameter, k times the larger diameter, centre separation over the larger
 * diameter, row-node depth, log10 tol.  Three hidden layers of 64 with
 * GELU, one output.  Features and targets are standardized with the mean
 * and deviation of the first training set (kept with the weights).
 * Training is mini-batch Adam (beta 0.9 / 0.999, eps 1e-8) on the mean
 * squared error; the sample order of each epoch is a seeded shuffle, so a
 * run is reproducible.
 */
#include <math.h>

#include "bf_host.h"

#define RANK_NF 6

struct BfhipRankModel {
  float *dP, *dM, *dV; /* parameters and Adam moments (device) */
  float norm[14];      /* feature mean[6], deviation[6], target mean, deviation */
  int normSet;
  int64_t step;        /* Adam steps taken */
  uint64_t rng;        /* shuffles */
};

static uint64_t sm64(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static double unif(uint64_t *s) { return ((double)(sm64(s) >> 11) + 0.5) * 0x1.0p-53; }

void bfhipRankFeatures(BfReal const colBox[4], BfReal const rowBox[4], BfReal k, BfSize level, BfReal tol,
                       float out[6]) {
  const double dc = hypot(colBox[2] - colBox[0], colBox[3] - colBox[1]);
  const double dr = hypot(rowBox[2] - rowBox[0], rowBox[3] - rowBox[1]);
  const double d = dc > dr ? dc : dr;
  const double cx = 0.5 * (colBox[0] + colBox[2]) - 0.5 * (rowBox[0] + rowBox[2]);
  const double cy = 0.5 * (colBox[1] + colBox[3]) - 0.5 * (rowBox[1] + rowBox[3]);
  out[0] = (float)log(dc > 1e-300 ? dc : 1e-300);
  out[1] = (float)log(dr > 1e-300 ? dr : 1e-300);
  out[2] = (float)(k * d);
  out[3] = (float)(d > 0 ? hypot(cx, cy) / d : 0);
  out[4] = (float)level;
  out[5] = (float)log10(tol > 0 ? tol : 1e-300);
}

BfhipRankModel *bfhipRankModelNew(uint64_t seed) {
  const int np = bfk_mlp_nparam();
  BfhipRankModel *m = calloc(1, sizeof(BfhipRankModel));
  float *h = calloc((size_t)np, sizeof(float));
  if (!m || !h) {
    free(m);
    free(h);
    bfSetError(BF_ERROR_MEMORY_ERROR);
    return NULL;
  }
  m->rng = seed ^ 0x5DEECE66Dull;
  uint64_t s = seed;
  /* weights N(0, 2 / fan_in) (He), biases 0; layout W1 64x6, b1, W2 64x64,
   * b2, W3 64x64, b3, W4 1x64, b4 */
  const int fan[4] = {6, 64, 64, 64}, rows[4] = {64, 64, 64, 1};
  int off = 0;
  for (int l = 0; l < 4; ++l) {
    const double sd = sqrt(2.0 / fan[l]) * (l == 3 ? 0.5 : 1.0);
    for (int e = 0; e < rows[l] * fan[l]; e += 2) {
      const double r = sqrt(-2 * log(unif(&s))), th = 6.283185307179586 * unif(&s);
      h[off + e] = (float)(sd * r * cos(th));
      if (e + 1 < rows[l] * fan[l]) h[off + e + 1] = (float)(sd * r * sin(th));
    }
    off += rows[l] * fan[l] + rows[l];
  }
  m->dP = bfk_malloc((size_t)np * 4);
  m->dM = bfk_malloc((size_t)np * 4);
  m->dV = bfk_malloc((size_t)np * 4);
  int rc = (!m->dP || !m->dM || !m->dV) ? -2 : 0;
  if (!rc) rc = bfk_h2d(m->dP, h, (size_t)np * 4);
  if (!rc) rc = bfk_memset(m->dM, (size_t)np * 4);
  if (!rc) rc = bfk_memset(m->dV, (size_t)np * 4);
  if (!rc) rc = bfk_sync();
  free(h);
  if (rc) {
    bfhipRankModelDelete(&m);
    bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
    return NULL;
  }
  return m;
}

void bfhipRankModelDelete(BfhipRankModel **model) {
  if (!model || !*model) return;
  bfk_free((*model)->dP);
  bfk_free((*model)->dM);
  bfk_free((*model)->dV);
  free(*model);
  *model = NULL;
}

/* standardized copy of n feature rows */
static float *standardize(BfhipRankModel const *m, BfSize n, float const *x) {
  float *o = malloc((n ? n : 1) * RANK_NF * sizeof(float));
  if (!o) return NULL;
  for (BfSize i = 0; i < n; ++i)
    for (int f = 0; f < RANK_NF; ++f) o[i * RANK_NF + f] = (x[i * RANK_NF + f] - m->norm[f]) / m->norm[6 + f];
  return o;
}

int bfhipRankModelTrain(BfhipRankModel *model, BfSize n, float const *features, float const *ranks, BfSize epochs,
                        BfSize batch, float lr, float *lastLoss) {
  /* sample indices are 32-bit on the device */
  if (!model || !features || !ranks || n == 0 || n > (BfSize)INT32_MAX || batch == 0 || batch > (1u << 30)) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  if (!model->normSet) { /* mean and deviation of this first training set */
    for (int f = 0; f < RANK_NF; ++f) {
      double s = 0, q = 0;
      for (BfSize i = 0; i < n; ++i) s += features[i * RANK_NF + f];
      s /= (double)n;
      for (BfSize i = 0; i < n; ++i) q += (features[i * RANK_NF + f] - s) * (features[i * RANK_NF + f] - s);
      q = sqrt(q / (double)n);
      model->norm[f] = (float)s;
      model->norm[6 + f] = (float)(q > 1e-6 ? q : 1.0);
    }
    double s = 0, q = 0;
    for (BfSize i = 0; i < n; ++i) s += ranks[i];
    s /= (double)n;
    for (BfSize i = 0; i < n; ++i) q += (ranks[i] - s) * (ranks[i] - s);
    q = sqrt(q / (double)n);
    model->norm[12] = (float)s;
    model->norm[13] = (float)(q > 1e-6 ? q : 1.0);
    model->normSet = 1;
  }
  float *hx = standardize(model, n, features);
  float *hy = malloc(n * sizeof(float));
  int32_t *idx = malloc(n * sizeof(int32_t));
  if (!hx || !hy || !idx) {
    free(hx); free(hy); free(idx);
    bfSetError(BF_ERROR_MEMORY_ERROR);
    return -2;
  }
  for (BfSize i = 0; i < n; ++i) hy[i] = (ranks[i] - model->norm[12]) / model->norm[13];
  const int B = (int)(batch < n ? batch : n);
  const int G = bfk_mlp_groups(B);
  const int np = bfk_mlp_nparam();
  float *dX = bfk_malloc(n * RANK_NF * 4), *dY = bfk_malloc(n * 4), *dPart = bfk_malloc((size_t)G * np * 4);
  float *dLossPart = bfk_malloc((size_t)G * 4);
  int32_t *dIdx = bfk_malloc(n * 4);
  double *dLoss = bfk_malloc(16);
  int rc = (!dX || !dY || !dPart || !dLossPart || !dIdx || !dLoss) ? -2 : 0;
  if (!rc) rc = bfk_h2d(dX, hx, n * RANK_NF * 4);
  if (!rc) rc = bfk_h2d(dY, hy, n * 4);
  for (BfSize ep = 0; ep < epochs && !rc; ++ep) {
    for (BfSize i = 0; i < n; ++i) idx[i] = (int32_t)i;
    for (BfSize i = n - 1; i > 0; --i) { /* Fisher-Yates */
      BfSize j = (BfSize)(sm64(&model->rng) % (uint64_t)(i + 1));
      int32_t t = idx[i];
      idx[i] = idx[j];
      idx[j] = t;
    }
    rc = bfk_h2d(dIdx, idx, n * 4);
    if (!rc && ep + 1 == epochs) rc = bfk_memset(dLoss, 16);
    for (BfSize b0 = 0; b0 < n && !rc; b0 += (BfSize)B) {
      const int nb = (int)(n - b0 < (BfSize)B ? n - b0 : (BfSize)B);
      model->step += 1;
      const float c1 = (float)(1.0 / (1.0 - pow(0.9, (double)model->step)));
      const float c2 = (float)(1.0 / (1.0 - pow(0.999, (double)model->step)));
      rc = bfk_mlp_step(model->dP, model->dM, model->dV, dX, dY, dIdx, (int64_t)b0, nb, lr, c1, c2, dPart, dLossPart,
                        dLoss);
    }
  }
  double loss[2] = {0, 0};
  if (!rc && epochs) rc = bfk_d2h(loss, dLoss, 16);
  if (!rc) rc = bfk_sync();
  if (lastLoss) *lastLoss = loss[1] > 0 ? (float)(loss[0] / loss[1] * model->norm[13] * model->norm[13]) : 0.0f;
  bfk_free(dX); bfk_free(dY); bfk_free(dPart); bfk_free(dLossPart); bfk_free(dIdx); bfk_free(dLoss);
  free(hx); free(hy); free(idx);
  if (rc) bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
  return rc;
}

/* the gradient of the mean squared error (in standardized units, as the
 * optimizer sees it) over all n samples at the current parameters */
int bfhipRankModelGradient(BfhipRankModel const *model, BfSize n, float const *features, float const *ranks,
                           float *grad, float *loss) {
  if (!model || !features || !ranks || !grad || n == 0 || n > (1u << 30) || !model->normSet) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  const int np = bfk_mlp_nparam(), G = bfk_mlp_groups((int)n);
  float *hx = standardize(model, n, features), *hy = malloc(n * sizeof(float));
  int32_t *idx = malloc(n * sizeof(int32_t));
  float *part = malloc((size_t)G * np * sizeof(float)), *lp = malloc((size_t)G * sizeof(float));
  float *dX = bfk_malloc(n * RANK_NF * 4), *dY = bfk_malloc(n * 4), *dPart = bfk_malloc((size_t)G * np * 4);
  float *dLp = bfk_malloc((size_t)G * 4);
  int32_t *dIdx = bfk_malloc(n * 4);
  int rc = (!hx || !hy || !idx || !part || !lp || !dX || !dY || !dPart || !dLp || !dIdx) ? -2 : 0;
  if (!rc) {
    for (BfSize i = 0; i < n; ++i) {
      hy[i] = (ranks[i] - model->norm[12]) / model->norm[13];
      idx[i] = (int32_t)i;
    }
    rc = bfk_h2d(dX, hx, n * RANK_NF * 4);
  }
  if (!rc) rc = bfk_h2d(dY, hy, n * 4);
  if (!rc) rc = bfk_h2d(dIdx, idx, n * 4);
  if (!rc) rc = bfk_mlp_grad(model->dP, dX, dY, dIdx, (int)n, dPart, dLp);
  if (!rc) rc = bfk_d2h(part, dPart, (size_t)G * np * 4);
  if (!rc) rc = bfk_d2h(lp, dLp, (size_t)G * 4);
  if (!rc) {
    for (int p = 0; p < np; ++p) {
      double g = 0;
      for (int w = 0; w < G; ++w) g += part[(size_t)w * np + p];
      grad[p] = (float)g;
    }
    double l = 0;
    for (int w = 0; w < G; ++w) l += lp[w];
    if (loss) *loss = (float)(l / (double)n);
  }
  bfk_free(dX); bfk_free(dY); bfk_free(dPart); bfk_free(dLp); bfk_free(dIdx);
  free(hx); free(hy); free(idx); free(part); free(lp);
  if (rc) bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
  return rc;
}

int bfhipRankModelPredict(BfhipRankModel const *model, BfSize n, float const *features, float *ranks) {
  if (!model || (n && (!features || !ranks)) || !model->normSet) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  if (n == 0) return 0;
  float *hx = standardize(model, n, features);
  float *dX = bfk_malloc(n * RANK_NF * 4), *dO = bfk_malloc(n * 4);
  int rc = (!hx || !dX || !dO) ? -2 : 0;
  if (!rc) rc = bfk_h2d(dX, hx, n * RANK_NF * 4);
  if (!rc) rc = bfk_mlp_predict(model->dP, dX, (int64_t)n, dO);
  if (!rc) rc = bfk_d2h(ranks, dO, n * 4);
  if (!rc)
    for (BfSize i = 0; i < n; ++i) ranks[i] = ranks[i] * model->norm[13] + model->norm[12];
  bfk_free(dX);
  bfk_free(dO);
  free(hx);
  if (rc) bfSetError(rc == -2 ? BF_ERROR_MEMORY_ERROR : BF_ERROR_RUNTIME_ERROR);
  return rc;
}

int bfhipRankModelGetParams(BfhipRankModel const *model, float *params, float norm[14]) {
  if (!model || !params || !norm) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  int rc = bfk_d2h(params, model->dP, (size_t)bfk_mlp_nparam() * 4);
  memcpy(norm, model->norm, sizeof(model->norm));
  if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}

int bfhipRankModelSetParams(BfhipRankModel *model, float const *params, float const norm[14]) {
  if (!model || !params || !norm) {
    bfSetError(BF_ERROR_INVALID_ARGUMENTS);
    return -1;
  }
  const size_t bytes = (size_t)bfk_mlp_nparam() * 4;
  int rc = bfk_h2d(model->dP, params, bytes);
  if (!rc) rc = bfk_memset(model->dM, bytes);
  if (!rc) rc = bfk_memset(model->dV, bytes);
  if (!rc) rc = bfk_sync();
  memcpy(model->norm, norm, sizeof(model->norm));
  model->normSet = 1;
  model->step = 0;
  if (rc) bfSetError(BF_ERROR_RUNTIME_ERROR);
  return rc;
}

BfSize bfhipRankModelNumParams(void) { return (BfSize)bfk_mlp_nparam(); }
