/* oracle.c -- FP64 CPU restatement of VLFeat 0.9.21's GMM-EM, Fisher, VLAD,
 * exact k-means quantizer and homogeneous kernel map, used ONLY as the
 * checker of the GPU library
 * (tests, smoke test and the CPU leg of the benchmark).  Nothing in the
 * product links or calls it.
 *
 * Every routine runs in double precision on double inputs, single-threaded,
 * in the reference's loop order, and cites the reference lines it restates.
 * How each routine is pinned (closed forms, known answers, the compiled
 * reference, brute force) is listed in the tests that exercise it.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define OR_API __attribute__((visibility("default")))

static const double kPi = 3.141592653589793 ;

/* gmm.c:711-798: q[i*K+k] = posterior, returns sum_i log sum_k pi_k N(x_i). */
OR_API double
or_gmm_posteriors (double * q, size_t K, size_t N, double const * priors, double const * means,
                   size_t D, double const * covs, double const * x)
{
  double LL = 0 ;
  double halfDimLog2Pi = (D / 2.0) * log(2.0 * kPi) ;
  double * logCov = (double*) malloc(sizeof(double) * K) ;
  double * logW = (double*) malloc(sizeof(double) * K) ;
  double * invCov = (double*) malloc(sizeof(double) * K * D) ;
  size_t i, k, d ;
  for (k = 0 ; k < K ; ++k) {
    double s = 0 ;
    logW[k] = (priors[k] < 1e-6) ? -INFINITY : log(priors[k]) ;
    for (d = 0 ; d < D ; ++d) {
      s += log(covs[k * D + d]) ;
      invCov[k * D + d] = 1.0 / covs[k * D + d] ;
    }
    logCov[k] = s ;
  }
  for (i = 0 ; i < N ; ++i) {
    double mx = -INFINITY, sum = 0 ;
    for (k = 0 ; k < K ; ++k) {
      double m = 0, p ;
      for (d = 0 ; d < D ; ++d) {
        double t = x[i * D + d] - means[k * D + d] ;
        m += t * t * invCov[k * D + d] ;
      }
      p = logW[k] - halfDimLog2Pi - 0.5 * logCov[k] - 0.5 * m ;
      q[i * K + k] = p ;
      if (p > mx) mx = p ;
    }
    for (k = 0 ; k < K ; ++k) {
      double p = exp(q[i * K + k] - mx) ;
      q[i * K + k] = p ;
      sum += p ;
    }
    LL += log(sum) + mx ;
    for (k = 0 ; k < K ; ++k) q[i * K + k] /= sum ;
  }
  free(logCov) ; free(logW) ; free(invCov) ;
  return LL ;
}

/* gmm.c:1005-1030 */
static void
apply_bounds (double * covs, size_t K, size_t D, double const * lbs)
{
  size_t k, d ;
  for (k = 0 ; k < K ; ++k)
    for (d = 0 ; d < D ; ++d)
      if (covs[k * D + d] < lbs[d]) covs[k * D + d] = lbs[d] ;
}

/* gmm.c:1036-1213: posterior-weighted moments about the old means, skipping
 * posteriors below 0.01/K, division by the mass when it is at least 1e-6/K,
 * old-to-new mean correction, variance floor, prior normalisation. */
OR_API void
or_gmm_maximization (double const * q, double const * x, size_t N, size_t D, size_t K,
                     double * priors, double * means, double * covs, double const * lbs)
{
  double * oldMeans = (double*) malloc(sizeof(double) * K * D) ;
  size_t i, k, d ;
  double sum = 0 ;
  memcpy(oldMeans, means, sizeof(double) * K * D) ;
  memset(priors, 0, sizeof(double) * K) ;
  memset(means, 0, sizeof(double) * K * D) ;
  memset(covs, 0, sizeof(double) * K * D) ;
  for (i = 0 ; i < N ; ++i) {
    for (k = 0 ; k < K ; ++k) {
      double p = q[i * K + k] ;
      if (p < 1e-2 / K) continue ;
      priors[k] += p ;
      for (d = 0 ; d < D ; ++d) {
        double xv = x[i * D + d], diff = xv - oldMeans[k * D + d] ;
        means[k * D + d] += p * xv ;
        covs[k * D + d] += p * (diff * diff) ;
      }
    }
  }
  for (k = 0 ; k < K ; ++k) {
    double mass = priors[k] ;
    if (mass >= 1e-6 / K) {
      for (d = 0 ; d < D ; ++d) {
        means[k * D + d] /= mass ;
        covs[k * D + d] /= mass ;
      }
    }
  }
  for (k = 0 ; k < K ; ++k) {
    if (priors[k] >= 1e-6 / K) {
      for (d = 0 ; d < D ; ++d) {
        double diff = means[k * D + d] - oldMeans[k * D + d] ;
        covs[k * D + d] -= diff * diff ;
      }
    }
  }
  apply_bounds(covs, K, D, lbs) ;
  for (k = 0 ; k < K ; ++k) sum += priors[k] ;
  if (sum < 1e-12) sum = 1e-12 ;
  for (k = 0 ; k < K ; ++k) priors[k] /= sum ;
  free(oldMeans) ;
}

/* gmm.c:815-999: split nearly empty modes along the largest-variance
 * dimension of the mode with the largest entropy contribution. */
OR_API size_t
or_gmm_restart (double * q, double const * x, size_t N, size_t D, size_t K,
                double * priors, double * means, double * covs, double const * lbs)
{
  double * mass ;
  size_t i, splits = 0 ;
  long ic, jc, d ;
  if (K <= 1) return 0 ;
  mass = (double*) calloc(K, sizeof(double)) ;
  for (i = 0 ; i < N ; ++i)
    for (ic = 0 ; ic < (long)K ; ++ic) mass[ic] += q[i * K + ic] ;
  for (ic = 0 ; ic < (long)K ; ++ic) {
    double size = -INFINITY ;
    long best = -1 ;
    double thr = 1e-2 * ((double)N / K > 1.0 ? (double)N / K : 1.0) ;
    if (mass[ic] >= thr) continue ;
    for (jc = 0 ; jc < (long)K ; ++jc) {
      double s ;
      if (priors[jc] < 1e-6) continue ;
      s = 0.5 * D * (1.0 + log(2 * kPi)) ;
      for (d = 0 ; d < (long)D ; ++d) s += 0.5 * log(covs[jc * D + d]) ;
      s = priors[jc] * (s - log(priors[jc])) ;
      if (s > size) { size = s ; best = jc ; }
    }
    jc = best ;
    if (jc == ic || jc < 0) continue ;
    size = -INFINITY ;
    best = -1 ;
    for (d = 0 ; d < (long)D ; ++d) {
      if (covs[jc * D + d] > size) { size = covs[jc * D + d] ; best = d ; }
    }
    {
      double mu = means[best + jc * D] ;
      for (i = 0 ; i < N ; ++i) {
        double p = q[jc + K * i], qq = q[ic + K * i] ;
        if (x[best + i * D] < mu) { q[ic + K * i] = p + qq ; q[jc + K * i] = 0 ; }
        else { q[ic + K * i] = 0 ; q[jc + K * i] = p + qq ; }
      }
    }
    or_gmm_maximization(q, x, N, D, K, priors, means, covs, lbs) ;
    splits ++ ;
  }
  free(mass) ;
  return splits ;
}

/* gmm.c:1220-1308: EM from the given parameters; q (N x K) receives the
 * posteriors of the last E-step; returns the last log-likelihood. */
OR_API double
or_gmm_em (double const * x, size_t N, size_t D, size_t K, size_t maxIter,
           double * priors, double * means, double * covs, double const * lbs, double * q)
{
  double prevLL = -INFINITY, LL = -INFINITY ;
  size_t it ;
  apply_bounds(covs, K, D, lbs) ;
  for (it = 0 ; 1 ; ++it) {
    double eps ;
    LL = or_gmm_posteriors(q, K, N, priors, means, D, covs, x) ;
    if (it >= maxIter) break ;
    eps = fabs((LL - prevLL) / LL) ;
    if (it > 0 && eps < 1e-5) break ;
    prevLL = LL ;
    if (it > 1) or_gmm_restart(q, x, N, D, K, priors, means, covs, lbs) ;
    or_gmm_maximization(q, x, N, D, K, priors, means, covs, lbs) ;
  }
  return LL ;
}

/* fisher.c:363-488. flags: 1 square root, 2 normalised, 4 fast.
 * Returns the number of (component, datum) terms accumulated. */
OR_API size_t
or_fisher_encode (double * enc, double const * means, size_t D, size_t K, double const * covs,
                  double const * priors, double const * x, size_t N, int flags)
{
  double * q = (double*) malloc(sizeof(double) * K * (N ? N : 1)) ;
  double * isig = (double*) malloc(sizeof(double) * K * D) ;
  size_t numTerms = 0, i, k, d ;
  memset(enc, 0, sizeof(double) * 2 * D * K) ;
  for (k = 0 ; k < K * D ; ++k) isig[k] = sqrt(1.0 / covs[k]) ;
  or_gmm_posteriors(q, K, N, priors, means, D, covs, x) ;
  if (flags & 4) {
    for (i = 0 ; i < N ; ++i) {
      size_t best = 0 ;
      double bv = q[i * K] ;
      for (k = 1 ; k < K ; ++k) if (q[i * K + k] > bv) { bv = q[i * K + k] ; best = k ; }
      for (k = 0 ; k < K ; ++k) q[i * K + k] = (k == best) ;
    }
  }
  for (k = 0 ; k < K ; ++k) {
    double * uk = enc + k * D ;
    double * vk = enc + k * D + K * D ;
    if (priors[k] < 1e-6) continue ;
    for (i = 0 ; i < N ; ++i) {
      double p = q[i * K + k] ;
      if (p < 1e-6) continue ;
      numTerms ++ ;
      for (d = 0 ; d < D ; ++d) {
        double diff = (x[i * D + d] - means[k * D + d]) * isig[k * D + d] ;
        uk[d] += p * diff ;
        vk[d] += p * (diff * diff - 1) ;
      }
    }
    if (N > 0) {
      double up = 1 / (N * sqrt(priors[k])), vp = 1 / (N * sqrt(2 * priors[k])) ;
      for (d = 0 ; d < D ; ++d) { uk[d] *= up ; vk[d] *= vp ; }
    }
  }
  if (flags & 1) {
    for (d = 0 ; d < 2 * D * K ; ++d) enc[d] = enc[d] >= 0 ? sqrt(enc[d]) : -sqrt(-enc[d]) ;
  }
  if (flags & 2) {
    double n = 0 ;
    for (d = 0 ; d < 2 * D * K ; ++d) n += enc[d] * enc[d] ;
    n = sqrt(n) ;
    if (n < 1e-12) n = 1e-12 ;
    for (d = 0 ; d < 2 * D * K ; ++d) enc[d] /= n ;
  }
  free(q) ;
  free(isig) ;
  return numTerms ;
}

/* vlad.c:163-242. flags: 1 normalise components, 2 square root,
 * 4 unnormalised, 8 normalise mass. assign is N x K (row-major). */
OR_API void
or_vlad_encode (double * enc, double const * means, size_t D, size_t K, double const * x, size_t N,
                double const * assign, int flags)
{
  size_t i, k, d ;
  memset(enc, 0, sizeof(double) * D * K) ;
  for (k = 0 ; k < K ; ++k) {
    double mass = 0 ;
    for (i = 0 ; i < N ; ++i) {
      double q = assign[i * K + k] ;
      if (q > 0) {
        mass += q ;
        for (d = 0 ; d < D ; ++d) enc[k * D + d] += q * x[i * D + d] ;
      }
    }
    if (mass > 0) {
      for (d = 0 ; d < D ; ++d) {
        if (flags & 8) enc[k * D + d] = enc[k * D + d] / mass - means[k * D + d] ;
        else enc[k * D + d] -= mass * means[k * D + d] ;
      }
    }
    if (flags & 2) {
      for (d = 0 ; d < D ; ++d) {
        double z = enc[k * D + d] ;
        enc[k * D + d] = z >= 0 ? sqrt(z) : -sqrt(-z) ;
      }
    }
    if (flags & 1) {
      double n = 0 ;
      for (d = 0 ; d < D ; ++d) n += enc[k * D + d] * enc[k * D + d] ;
      n = sqrt(n) ;
      if (n < 1e-12) n = 1e-12 ;
      for (d = 0 ; d < D ; ++d) enc[k * D + d] /= n ;
    }
  }
  if (!(flags & 4)) {
    double n = 0 ;
    for (d = 0 ; d < D * K ; ++d) n += enc[d] * enc[d] ;
    n = sqrt(n) ;
    if (n < 1e-12) n = 1e-12 ;
    for (d = 0 ; d < D * K ; ++d) enc[d] /= n ;
  }
}

/* kmeans.c:655-702 in exact arithmetic order-free form: nearest center by
 * squared L2, strict '<' so the first index wins ties. */
OR_API void
or_kmeans_quantize (uint32_t * assign, double * dist, double const * x, double const * centers,
                    size_t N, size_t D, size_t K)
{
  size_t i, k, d ;
  for (i = 0 ; i < N ; ++i) {
    double best = INFINITY ;
    uint32_t bi = 0 ;
    for (k = 0 ; k < K ; ++k) {
      double s = 0 ;
      for (d = 0 ; d < D ; ++d) { double t = x[i * D + d] - centers[k * D + d] ; s += t * t ; }
      if (s < best) { best = s ; bi = (uint32_t)k ; }
    }
    assign[i] = bi ;
    if (dist) dist[i] = best ;
  }
}

/* ------------------------------------------------------------------------- */
/* Learned gate (no counterpart in the reference): the perceptron
 *   a1 = x W1 + b1, h1 = gelu(a1), a2 = h1 W2 + b2, h2 = gelu(a2), z = h2 W3 + b3,
 *   gelu(a) = a Phi(a) = a (1 + erf(a / sqrt 2)) / 2,
 * loss = mean over rows of -sum_k q_k log softmax(z)_k, and its gradient with
 * respect to the parameters laid out [W1 (D x H1) | b1 | W2 (H1 x H2) | b2 |
 * W3 (H2 x K) | b3] (weights row-major, input index first).  FP64 throughout;
 * xin is the already standardised input.  Returns the mean loss; grads may be
 * NULL. */
static double or_gelu (double a) { return 0.5 * a * (1.0 + erf(a / sqrt(2.0))) ; }
static double or_gelu_grad (double a)
{
  return 0.5 * (1.0 + erf(a / sqrt(2.0))) + a * exp(-0.5 * a * a) * 0.39894228040143267794 ; /* 1 / sqrt(2 pi) */
}

OR_API double
or_mlp_loss_grads (double const * params, size_t D, size_t H1, size_t H2, size_t K,
                   double const * xin, double const * q, size_t rows, double * grads)
{
  double const * W1 = params ;
  double const * b1 = W1 + D * H1 ;
  double const * W2 = b1 + H1 ;
  double const * b2 = W2 + H1 * H2 ;
  double const * W3 = b2 + H2 ;
  double const * b3 = W3 + H2 * K ;
  size_t nP = D * H1 + H1 + H1 * H2 + H2 + H2 * K + K ;
  double * a1 = malloc(sizeof(double) * H1) ;
  double * h1 = malloc(sizeof(double) * H1) ;
  double * a2 = malloc(sizeof(double) * H2) ;
  double * h2 = malloc(sizeof(double) * H2) ;
  double * z = malloc(sizeof(double) * K) ;
  double * dz = malloc(sizeof(double) * K) ;
  double * dh2 = malloc(sizeof(double) * H2) ;
  double * dh1 = malloc(sizeof(double) * H1) ;
  double loss = 0 ;
  size_t i, j, k, d ;
  if (grads) memset(grads, 0, sizeof(double) * nP) ;
  for (i = 0 ; i < rows ; ++i) {
    double const * x = xin + i * D ;
    double const * qi = q + i * K ;
    double m = -INFINITY, s = 0, ls ;
    for (j = 0 ; j < H1 ; ++j) {
      double t = b1[j] ;
      for (d = 0 ; d < D ; ++d) t += x[d] * W1[d * H1 + j] ;
      a1[j] = t ; h1[j] = or_gelu(t) ;
    }
    for (j = 0 ; j < H2 ; ++j) {
      double t = b2[j] ;
      for (d = 0 ; d < H1 ; ++d) t += h1[d] * W2[d * H2 + j] ;
      a2[j] = t ; h2[j] = or_gelu(t) ;
    }
    for (k = 0 ; k < K ; ++k) {
      double t = b3[k] ;
      for (d = 0 ; d < H2 ; ++d) t += h2[d] * W3[d * K + k] ;
      z[k] = t ; if (t > m) m = t ;
    }
    for (k = 0 ; k < K ; ++k) s += exp(z[k] - m) ;
    ls = log(s) ;
    for (k = 0 ; k < K ; ++k) {
      double lp = z[k] - m - ls ;
      if (qi[k] > 0) loss -= qi[k] * lp ;
      dz[k] = (exp(lp) - qi[k]) / (double)rows ;
    }
    if (!grads) continue ;
    {
      double * gW1 = grads ;
      double * gb1 = gW1 + D * H1 ;
      double * gW2 = gb1 + H1 ;
      double * gb2 = gW2 + H1 * H2 ;
      double * gW3 = gb2 + H2 ;
      double * gb3 = gW3 + H2 * K ;
      for (k = 0 ; k < K ; ++k) gb3[k] += dz[k] ;
      for (d = 0 ; d < H2 ; ++d) {
        double t = 0 ;
        for (k = 0 ; k < K ; ++k) { gW3[d * K + k] += h2[d] * dz[k] ; t += dz[k] * W3[d * K + k] ; }
        dh2[d] = t * or_gelu_grad(a2[d]) ;
      }
      for (j = 0 ; j < H2 ; ++j) gb2[j] += dh2[j] ;
      for (d = 0 ; d < H1 ; ++d) {
        double t = 0 ;
        for (j = 0 ; j < H2 ; ++j) { gW2[d * H2 + j] += h1[d] * dh2[j] ; t += dh2[j] * W2[d * H2 + j] ; }
        dh1[d] = t * or_gelu_grad(a1[d]) ;
      }
      for (j = 0 ; j < H1 ; ++j) gb1[j] += dh1[j] ;
      for (d = 0 ; d < D ; ++d)
        for (j = 0 ; j < H1 ; ++j) gW1[d * H1 + j] += x[d] * dh1[j] ;
    }
  }
  free(a1) ; free(h1) ; free(a2) ; free(h2) ; free(z) ; free(dz) ; free(dh2) ; free(dh1) ;
  return rows ? loss / (double)rows : 0.0 ;
}

/* ---- homogeneous kernel map (homkermap.c:239-426, 532-571) ----------------
 * or_homkermap maps n doubles into n x (2*order+1) doubles: the reference's
 * table of the windowed, periodised spectrum (kernel 0 intersection, 1 chi2,
 * 2 Jensen-Shannon; window 0 uniform, 1 rectangular; period < 0: default),
 * then linear interpolation between the rows bracketing each mantissa. */
/* separate sin and cos calls, as in the reference's build (a merged sincos
 * call can differ in the last bit) */
static double (*volatile or_sin) (double) = sin ;
static double (*volatile or_cos) (double) = cos ;

static double
or_hkm_spectrum (int kernel, double w)
{
  double sech = 2.0 / (exp(kPi * w) + exp(-kPi * w)) ;
  if (kernel == 0) return (2.0 / kPi) / (1 + 4 * w * w) ;
  if (kernel == 1) return sech ;
  return (2.0 / log(4.0)) * 2.0 / (exp(kPi * w) + exp(-kPi * w)) / (1 + 4 * w * w) ;
}

static double
or_hkm_window (int kernel, int window, double period, double w)
{
  double range, step, acc = 0, u ;
  if (window == 0) return or_hkm_spectrum(kernel, w) ;
  range = 2.0 / (period * 1e-2) ;
  step = 2 * range / (2 * 1024.0 + 1) ;
  for (u = -range ; u <= range ; u += step) {
    double a = (period / 2.0) * u ;
    double s = (a == 0.0) ? 1.0 : or_sin(a) / a ;
    s *= (period / (2.0 * kPi)) ;
    acc += s * or_hkm_spectrum(kernel, u + w) ;
  }
  acc *= step ;
  return acc > 0.0 ? acc : 0.0 ;
}

OR_API void
or_homkermap (double * out, double const * x, size_t n, int kernel, double gamma, size_t order,
              double period, int window)
{
  size_t const F = 2 * order + 1, S = 8 + 8 * order ;
  int const e0 = -20, e1 = 8 ;
  double const sub = 1.0 / S ;
  double * tab = (double*) malloc(sizeof(double) * F * S * (e1 - e0 + 1)) ;
  double * kap = (double*) malloc(sizeof(double) * (order + 1)) ;
  double * frq = (double*) malloc(sizeof(double) * (order + 1)) ;
  double L, m, *row ;
  size_t i, j, s ;
  int e ;
  if (period < 0) {
    if (window == 0) period = kernel == 1 ? 5.86 * sqrt(order + 0) + 3.65
                            : kernel == 2 ? 6.64 * sqrt(order + 0) + 7.24 : 2.38 * log(order + 0.8) + 5.6 ;
    else period = kernel == 1 ? 8.80 * sqrt(order + 4.44) - 12.6
                : kernel == 2 ? 9.63 * sqrt(order + 1.00) - 2.93 : 2.00 * log(order + 0.99) + 3.52 ;
    if (period < 1.0) period = 1.0 ;
  }
  L = 2.0 * kPi / period ;
  for (i = 0, j = 0 ; i <= order ; ) {
    frq[i] = (double) j ;
    kap[i] = or_hkm_window(kernel, window, period, j * L) ;
    ++ j ;
    if (kap[i] > 0 || j >= 3 * i) ++ i ;
  }
  row = tab ;
  for (e = e0 ; e <= e1 ; ++e) {
    for (s = 0, m = 1.0 ; s < S ; ++s, m += sub) {
      double xv = ldexp(m, e), lx = L * pow(xv, gamma), ll = L * log(xv) ;
      *row++ = sqrt(lx * kap[0]) ;
      for (j = 1 ; j <= order ; ++j) {
        double a = sqrt(2.0 * lx * kap[j]) ;
        *row++ = a * or_cos(frq[j] * ll) ;
        *row++ = a * or_sin(frq[j] * ll) ;
      }
    }
  }
  for (i = 0 ; i < n ; ++i) {
    double * o = out + i * F ;
    double const * lo, * hi ;
    double sg ;
    int ex ;
    m = frexp(x[i], &ex) ;
    sg = (m >= 0.0) ? 1.0 : -1.0 ;
    m *= 2 * sg ;
    ex -- ;
    if (m == 0 || ex <= e0 || ex >= e1) {
      for (j = 0 ; j < F ; ++j) o[j] = 0.0 ;
      continue ;
    }
    lo = tab + (size_t)(ex - e0) * S * F ;
    m -= 1.0 ;
    while (m >= sub) { m -= sub ; lo += F ; }
    hi = lo + F ;
    for (j = 0 ; j < F ; ++j) o[j] = sg * ((hi[j] - lo[j]) * (S * m) + lo[j]) ;
  }
  free(tab) ;
  free(kap) ;
  free(frq) ;
}
