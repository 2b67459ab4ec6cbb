/* vlsynth.h -- seeded synthetic data of the benchmark workloads (host C).
 *
 * Every value is a function of (seed, row) only: each row draws from its own
 * SplitMix64 stream, normals come from Marsaglia's polar method with a
 * logarithm built from basic IEEE operations (no libm transcendental), and all
 * arithmetic is double without contraction, so the same call gives the same
 * bits on any IEEE-754 host, for any thread count and for any split of the
 * rows into chunks (rowOffset).  The rows are generated in parallel with
 * OpenMP.
 *
 * Workloads:
 *   SIFT-like   z ~ N(0, I_128), rectified, L2-normalised, clipped at 0.2,
 *               renormalised, then projected to dim < 128 by the fixed
 *               orthonormal rows of vlsynth_projection (dim = 128: no
 *               projection);
 *   mixture     x ~ sum_k pi_k N(mu_k, diag s_k^2), mu ~ U[-1,1]^dim,
 *               s ~ U[0.1,0.5]^dim, pi ~ Dirichlet(1) (normalised -log U);
 *   counts      n_i = clamp(round(exp(ln median + sigma z_i)), lo, hi). */
#ifndef VLSYNTH_H
#define VLSYNTH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#if defined(__GNUC__)
#define VLSYNTH_API __attribute__((visibility("default")))
#else
#define VLSYNTH_API
#endif

/* rows [rowOffset, rowOffset + n) of the SIFT-like workload, dim in [1, 128],
 * into out[n][dim] (float). */
VLSYNTH_API void vlsynth_sift_like(float *out, size_t n, int dim, uint64_t seed, uint64_t rowOffset);

/* The projection of the SIFT-like workload: dimOut orthonormal rows of length
 * dimIn (Gram-Schmidt, twice, over the rows of a seeded Gaussian matrix), into
 * P[dimOut][dimIn] (double).  dimOut <= dimIn <= 128. */
VLSYNTH_API void vlsynth_projection(double *P, int dimOut, int dimIn, uint64_t seed);

/* rows [rowOffset, rowOffset + n) of a kTrue-component Gaussian mixture in dim
 * <= 1024 dimensions (its parameters drawn from the seed), into out[n][dim] (float);
 * when the pointers are not NULL, the mixture's means and standard deviations
 * [kTrue][dim] and weights [kTrue] (double). */
VLSYNTH_API void vlsynth_mixture(float *out, size_t n, int dim, int kTrue, uint64_t seed, uint64_t rowOffset,
                                 double *means, double *sigmas, double *weights);

/* numImages descriptor counts of the uneven-image workload. */
VLSYNTH_API void vlsynth_lognormal_counts(int64_t *counts, size_t numImages, double median, double sigma,
                                          int64_t lo, int64_t hi, uint64_t seed);

/* Building blocks, exported for tests: uniform [0, 1) draw k of stream
 * (seed, stream, row); the logarithm and exponential used above. */
VLSYNTH_API double vlsynth_uniform(uint64_t seed, uint64_t stream, uint64_t row, uint64_t k);
VLSYNTH_API double vlsynth_log(double x);
VLSYNTH_API double vlsynth_exp(double x);

#ifdef __cplusplus
}
#endif

#endif /* VLSYNTH_H */
