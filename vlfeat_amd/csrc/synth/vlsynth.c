/* vlsynth.c -- seeded synthetic data of the benchmark workloads.
 *
 * Random numbers: SplitMix64 (Steele, Lea and Flood, "Fast splittable
 * pseudorandom number generators", OOPSLA 2014): state += 0x9e3779b97f4a7c15,
 * output = the state through two xor-shift-multiply rounds.  Each (seed,
 * stream, row) starts its own state, hashed from the three with the same
 * mixing function, so rows are independent of each other and of the order in
 * which they are generated.  Uniforms take the top 53 bits.
 *
 * Normals: Marsaglia's polar method (u, v uniform in (-1, 1), s = u^2 + v^2
 * accepted in (0, 1), z = u sqrt(-2 log s / s), v sqrt(-2 log s / s)).
 * log and exp are computed here from basic IEEE operations (frexp/ldexp are
 * exact): log m for m in [sqrt(1/2), sqrt(2)) as 2 atanh((m - 1)/(m + 1)) by
 * its odd series to the 21st power, exp by reduction to |r| <= ln2/2 and the
 * Taylor series to r^13, both with ln 2 split into a 32-bit head and a tail.
 * Compiled without floating-point contraction. */
#include "vlsynth.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* widest mixture row (a row's normals live on the stack) */
#define VLSYNTH_MAX_DIM 1024

enum { STREAM_SIFT = 1, STREAM_PROJ = 2, STREAM_MIX_PARAM = 3, STREAM_MIX_ROW = 4, STREAM_COUNTS = 5 };

static inline uint64_t mix64(uint64_t z) {
  z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
  z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
  return z ^ (z >> 31);
}

static inline uint64_t stream_start(uint64_t seed, uint64_t stream, uint64_t row) {
  return mix64(seed ^ mix64(stream * 0x9e3779b97f4a7c15ULL ^ mix64(row + 1)));
}

static inline uint64_t next64(uint64_t *s) {
  *s += 0x9e3779b97f4a7c15ULL;
  return mix64(*s);
}

static inline double next_uniform(uint64_t *s) { return (double)(next64(s) >> 11) * 0x1.0p-53; }

/* ln 2 = LN2_HI + LN2_LO, LN2_HI with its low 32 bits zero (k * LN2_HI exact for |k| < 2^20) */
static const double LN2_HI = 0x1.62e42fefa0000p-1;
static const double LN2_LO = 0x1.cf79abc9e3b3ap-40;

double vlsynth_log(double x) {
  if (!(x > 0.0)) return x == 0.0 ? -INFINITY : NAN;
  if (isinf(x)) return x;
  int e;
  double m = frexp(x, &e); /* x = m 2^e, m in [0.5, 1) */
  if (m < 0.70710678118654752440) {
    m *= 2.0;
    e -= 1;
  }
  const double s = (m - 1.0) / (m + 1.0), s2 = s * s;
  /* 2 (s + s^3/3 + ... + s^21/21) */
  double p = 1.0 / 21.0;
  p = p * s2 + 1.0 / 19.0;
  p = p * s2 + 1.0 / 17.0;
  p = p * s2 + 1.0 / 15.0;
  p = p * s2 + 1.0 / 13.0;
  p = p * s2 + 1.0 / 11.0;
  p = p * s2 + 1.0 / 9.0;
  p = p * s2 + 1.0 / 7.0;
  p = p * s2 + 1.0 / 5.0;
  p = p * s2 + 1.0 / 3.0;
  const double lm = 2.0 * s + 2.0 * s * (s2 * p);
  return ((double)e * LN2_HI + lm) + (double)e * LN2_LO;
}

double vlsynth_exp(double x) {
  if (isnan(x)) return x;
  if (x > 709.78) return INFINITY;
  if (x < -745.2) return 0.0;
  const double kd = floor(x / 0.69314718055994530942 + 0.5);
  const int k = (int)kd;
  const double r = (x - kd * LN2_HI) - kd * LN2_LO;
  /* e^r = 1 + r + r^2/2! + ... + r^13/13! */
  double p = 1.0 / 6227020800.0;
  p = p * r + 1.0 / 479001600.0;
  p = p * r + 1.0 / 39916800.0;
  p = p * r + 1.0 / 3628800.0;
  p = p * r + 1.0 / 362880.0;
  p = p * r + 1.0 / 40320.0;
  p = p * r + 1.0 / 5040.0;
  p = p * r + 1.0 / 720.0;
  p = p * r + 1.0 / 120.0;
  p = p * r + 1.0 / 24.0;
  p = p * r + 1.0 / 6.0;
  p = p * r + 0.5;
  p = p * r + 1.0;
  p = p * r + 1.0;
  return ldexp(p, k);
}

double vlsynth_uniform(uint64_t seed, uint64_t stream, uint64_t row, uint64_t k) {
  uint64_t s = stream_start(seed, stream, row);
  double u = 0.0;
  for (uint64_t i = 0; i <= k; ++i) u = next_uniform(&s);
  return u;
}

/* n normals (n even) from the stream s, by pairs */
static void normals(uint64_t *s, double *z, int n) {
  for (int i = 0; i < n; i += 2) {
    double u, v, q;
    do {
      u = 2.0 * next_uniform(s) - 1.0;
      v = 2.0 * next_uniform(s) - 1.0;
      q = u * u + v * v;
    } while (q >= 1.0 || q == 0.0);
    const double f = sqrt(-2.0 * vlsynth_log(q) / q);
    z[i] = u * f;
    if (i + 1 < n) z[i + 1] = v * f;
  }
}

static void normalise(double *z, int n) {
  double nn = 0.0;
  for (int i = 0; i < n; ++i) nn += z[i] * z[i];
  double l = sqrt(nn);
  if (l < 1e-12) l = 1e-12;
  for (int i = 0; i < n; ++i) z[i] /= l;
}

void vlsynth_projection(double *P, int dimOut, int dimIn, uint64_t seed) {
  if (dimOut < 1 || dimIn < dimOut || dimIn > 128) return;
  /* Gram-Schmidt, twice, over Gaussian rows of length dimIn */
  for (int r = 0; r < dimOut; ++r) {
    uint64_t s = stream_start(seed, STREAM_PROJ, (uint64_t)r);
    double g[128];
    normals(&s, g, 128);
    double *p = P + (size_t)r * dimIn;
    for (int j = 0; j < dimIn; ++j) p[j] = g[j];
    for (int pass = 0; pass < 2; ++pass) {
      for (int q = 0; q < r; ++q) {
        const double *o = P + (size_t)q * dimIn;
        double d = 0.0;
        for (int j = 0; j < dimIn; ++j) d += p[j] * o[j];
        for (int j = 0; j < dimIn; ++j) p[j] -= d * o[j];
      }
      normalise(p, dimIn);
    }
  }
}

void vlsynth_sift_like(float *out, size_t n, int dim, uint64_t seed, uint64_t rowOffset) {
  if (dim < 1 || dim > 128 || n == 0) return;
  double *PT = NULL; /* the projection transposed: PT[j][d] */
  if (dim < 128) {
    double *P = (double *)malloc(sizeof(double) * (size_t)dim * 128);
    PT = (double *)malloc(sizeof(double) * (size_t)dim * 128);
    if (!P || !PT) {
      free(P);
      free(PT);
      return;
    }
    vlsynth_projection(P, dim, 128, seed);
    for (int d = 0; d < dim; ++d)
      for (int j = 0; j < 128; ++j) PT[(size_t)j * dim + d] = P[(size_t)d * 128 + j];
    free(P);
  }
#pragma omp parallel for schedule(static)
  for (long long i = 0; i < (long long)n; ++i) {
    uint64_t s = stream_start(seed, STREAM_SIFT, rowOffset + (uint64_t)i);
    double z[128], y[128];
    normals(&s, z, 128);
    for (int j = 0; j < 128; ++j) z[j] = z[j] > 0.0 ? z[j] : 0.0;
    normalise(z, 128);
    for (int j = 0; j < 128; ++j) z[j] = z[j] < 0.2 ? z[j] : 0.2;
    normalise(z, 128);
    float *o = out + (size_t)i * dim;
    if (!PT) {
      for (int j = 0; j < 128; ++j) o[j] = (float)z[j];
      continue;
    }
    /* y_d = sum_j P[d][j] z_j, j in order (zero terms skipped: adding +0 is exact) */
    for (int d = 0; d < dim; ++d) y[d] = 0.0;
    for (int j = 0; j < 128; ++j) {
      const double zj = z[j];
      if (zj == 0.0) continue;
      const double *pt = PT + (size_t)j * dim;
      for (int d = 0; d < dim; ++d) y[d] += pt[d] * zj;
    }
    for (int d = 0; d < dim; ++d) o[d] = (float)y[d];
  }
  free(PT);
}

void vlsynth_mixture(float *out, size_t n, int dim, int kTrue, uint64_t seed, uint64_t rowOffset, double *means,
                     double *sigmas, double *weights) {
  if (dim < 1 || dim > VLSYNTH_MAX_DIM || kTrue < 1) return;
  double *mu = (double *)malloc(sizeof(double) * (size_t)kTrue * dim);
  double *sg = (double *)malloc(sizeof(double) * (size_t)kTrue * dim);
  double *cw = (double *)malloc(sizeof(double) * (size_t)kTrue);
  if (!mu || !sg || !cw) {
    free(mu);
    free(sg);
    free(cw);
    return;
  }
  /* parameters: one stream per component (means, then deviations, then the weight draw) */
  double wsum = 0.0;
  for (int k = 0; k < kTrue; ++k) {
    uint64_t s = stream_start(seed, STREAM_MIX_PARAM, (uint64_t)k);
    for (int d = 0; d < dim; ++d) mu[(size_t)k * dim + d] = 2.0 * next_uniform(&s) - 1.0;
    for (int d = 0; d < dim; ++d) sg[(size_t)k * dim + d] = 0.1 + 0.4 * next_uniform(&s);
    double u;
    do u = next_uniform(&s);
    while (u == 0.0);
    cw[k] = -vlsynth_log(u);
    wsum += cw[k];
  }
  for (int k = 0; k < kTrue; ++k) cw[k] /= wsum;
  if (means) memcpy(means, mu, sizeof(double) * (size_t)kTrue * dim);
  if (sigmas) memcpy(sigmas, sg, sizeof(double) * (size_t)kTrue * dim);
  if (weights) memcpy(weights, cw, sizeof(double) * (size_t)kTrue);
  /* cumulative weights for the component draw */
  for (int k = 1; k < kTrue; ++k) cw[k] += cw[k - 1];
  if (out && n) {
    const int dz = dim + (dim & 1);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < (long long)n; ++i) {
      uint64_t s = stream_start(seed, STREAM_MIX_ROW, rowOffset + (uint64_t)i);
      const double u = next_uniform(&s);
      int c = 0;
      while (c < kTrue - 1 && u >= cw[c]) ++c;
      double z[VLSYNTH_MAX_DIM];
      normals(&s, z, dz);
      float *o = out + (size_t)i * dim;
      for (int d = 0; d < dim; ++d) o[d] = (float)(mu[(size_t)c * dim + d] + sg[(size_t)c * dim + d] * z[d]);
    }
  }
  free(mu);
  free(sg);
  free(cw);
}

void vlsynth_lognormal_counts(int64_t *counts, size_t numImages, double median, double sigma, int64_t lo,
                              int64_t hi, uint64_t seed) {
  const double lm = vlsynth_log(median);
  for (size_t i = 0; i < numImages; ++i) {
    uint64_t s = stream_start(seed, STREAM_COUNTS, (uint64_t)i);
    double z[2];
    normals(&s, z, 2);
    double v = floor(vlsynth_exp(lm + sigma * z[0]) + 0.5);
    if (v < (double)lo) v = (double)lo;
    if (v > (double)hi) v = (double)hi;
    counts[i] = (int64_t)v;
  }
}
