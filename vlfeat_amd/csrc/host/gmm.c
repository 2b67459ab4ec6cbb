/* gmm.c -- diagonal-covariance GMM by EM: host driver of the GPU kernels.
 *
 * Drop-in for vl/gmm.c (VLFeat 0.9.21).  The object keeps means, covariances
 * and priors in host memory (the getters return these buffers, as the
 * reference does); the data being clustered and the N x K posterior matrix
 * live on the GPU for the duration of EM.  Per iteration:
 *   E-step        vlhip_gmm_posteriors  (gmm.c:711-798)
 *   stop test     gmm.c:1270-1286: iteration >= maxNumIterations, or
 *                 |(LL - previousLL) / LL| < 1e-5 after the first iteration
 *   restarts      gmm.c:815-999, from iteration 2: component masses come from
 *                 the moment kernel; the split runs on the GPU
 *   M-step        vlhip_gmm_moments (gmm.c:1064-1167, FP64 reduction) followed
 *                 by the finalisation of gmm.c:1169-1202 on the host.
 * The posterior matrix is copied to the host only when vl_gmm_get_posteriors
 * is called.  Initialisation (k-means, k-means++ seeding, data variance) runs
 * through vl/kmeans.h and the host random generator exactly as gmm.c:1315-1443
 * does, so identical seeds give identical starting points.
 */
#include "host_fatal.h"
#include "vl/gmm.h"
#include "vlhip.h"

#include <assert.h>
#include <math.h>
#include <stdio.h>
#include <string.h>

#define GMM_MIN_POSTERIOR 1e-2
#define GMM_MIN_PRIOR 1e-6

struct _VlGMM
{
  vl_type dataType ;
  vl_size dimension ;
  vl_size numClusters ;
  vl_size numData ;
  vl_size maxNumIterations ;
  vl_size numRepetitions ;
  int verbosity ;
  void * means ;
  void * covariances ;
  void * priors ;
  void * posteriors ;          /* host copy, allocated on first request */
  double * sigmaLowBound ;
  VlGMMInitialization initialization ;
  VlKMeans * kmeansInit ;
  double LL ;
  vl_bool kmeansInitIsOwner ;
  /* device state */
  void * dData ; vl_size dDataBytes ;
  void * dPost ; vl_size dPostBytes ;
  void * dParams ; vl_size dParamsBytes ;
  vl_size postRows ;           /* rows held by dPost (0 = none) */
  vl_bool hostPostValid ;      /* host posteriors mirror dPost */
  vl_bool devPostStale ;       /* dPost predates the last E-step (matrix-core EM) */
  int precision ;              /* VL_GMM_PRECISION_* of the EM iterations */
} ;

/* ---------------------------------------------------------------- */
/*                                          device buffer management */

static void *
dev_reserve (void * p, vl_size * cap, vl_size bytes)
{
  if (*cap >= bytes && p) return p ;
  if (p) vlhip_free(p) ;
  *cap = bytes ? bytes : 1 ;
  return vlhip_malloc(*cap) ;
}

static void
gmm_release_device (VlGMM * self)
{
  if (self->dData) vlhip_free(self->dData) ;
  if (self->dPost) vlhip_free(self->dPost) ;
  if (self->dParams) vlhip_free(self->dParams) ;
  self->dData = self->dPost = self->dParams = NULL ;
  self->dDataBytes = self->dPostBytes = self->dParamsBytes = 0 ;
  self->postRows = 0 ;
}

static void
gmm_stage_data (VlGMM * self, void const * data, vl_size numData)
{
  vl_size es = vl_get_type_size(self->dataType) ;
  vl_size bytes = es * self->dimension * numData ;
  self->dData = dev_reserve(self->dData, &self->dDataBytes, bytes) ;
  if (bytes) vlhip_memcpy_h2d(self->dData, data, bytes) ;
  self->dPost = dev_reserve(self->dPost, &self->dPostBytes, es * self->numClusters * numData) ;
  self->dParams = dev_reserve(self->dParams, &self->dParamsBytes,
                              es * (3 * self->numClusters * self->dimension + self->numClusters)) ;
}

/* device params layout: [means KD][covariances KD][priors K][oldMeans KD] */
static void
gmm_upload_params (VlGMM * self)
{
  vl_size es = vl_get_type_size(self->dataType) ;
  vl_size KD = self->numClusters * self->dimension ;
  unsigned char * p = (unsigned char*) self->dParams ;
  vlhip_memcpy_h2d(p, self->means, es * KD) ;
  vlhip_memcpy_h2d(p + es * KD, self->covariances, es * KD) ;
  vlhip_memcpy_h2d(p + 2 * es * KD, self->priors, es * self->numClusters) ;
}

static void
_vl_gmm_prepare_for_data (VlGMM * self, vl_size numData)
{
  if (self->numData < numData || !self->posteriors) {
    if (self->posteriors) vl_free(self->posteriors) ;
    self->posteriors = NULL ;
  }
  self->numData = numData ;
  self->hostPostValid = VL_FALSE ;
}

/* ---------------------------------------------------------------- */
/*                                                        Life-cycle */

VL_EXPORT VlGMM *
vl_gmm_new (vl_type dataType, vl_size dimension, vl_size numComponents)
{
  vl_index i ;
  vl_size size = vl_get_type_size(dataType) ;
  VlGMM * self = vl_calloc(1, sizeof(VlGMM)) ;
  self->dataType = dataType ;
  self->numClusters = numComponents ;
  self->numData = 0 ;
  self->dimension = dimension ;
  self->initialization = VlGMMRand ;
  self->verbosity = 0 ;
  self->maxNumIterations = 50 ;
  self->numRepetitions = 1 ;
  self->kmeansInit = NULL ;
  self->kmeansInitIsOwner = VL_FALSE ;
  self->priors = vl_calloc(numComponents, size) ;
  self->means = vl_calloc(numComponents * dimension, size) ;
  self->covariances = vl_calloc(numComponents * dimension, size) ;
  self->sigmaLowBound = vl_calloc(dimension, sizeof(double)) ;
  for (i = 0 ; i < (signed)dimension ; ++i) self->sigmaLowBound[i] = 1e-4 ;
  return self ;
}

VL_EXPORT void
vl_gmm_reset (VlGMM * self)
{
  if (self->posteriors) {
    vl_free(self->posteriors) ;
    self->posteriors = NULL ;
  }
  self->numData = 0 ;
  self->postRows = 0 ;
  self->hostPostValid = VL_FALSE ;
  if (self->kmeansInit && self->kmeansInitIsOwner) {
    vl_kmeans_delete(self->kmeansInit) ;
    self->kmeansInit = NULL ;
    self->kmeansInitIsOwner = VL_FALSE ;
  }
}

VL_EXPORT void
vl_gmm_delete (VlGMM * self)
{
  if (self->means) vl_free(self->means) ;
  if (self->covariances) vl_free(self->covariances) ;
  if (self->priors) vl_free(self->priors) ;
  if (self->posteriors) vl_free(self->posteriors) ;
  if (self->sigmaLowBound) vl_free(self->sigmaLowBound) ;
  if (self->kmeansInit && self->kmeansInitIsOwner) vl_kmeans_delete(self->kmeansInit) ;
  gmm_release_device(self) ;
  vl_free(self) ;
}

VL_EXPORT VlGMM *
vl_gmm_new_copy (VlGMM const * self)
{
  vl_size size = vl_get_type_size(self->dataType) ;
  VlGMM * gmm = vl_gmm_new(self->dataType, self->dimension, self->numClusters) ;
  gmm->initialization = self->initialization ;
  gmm->maxNumIterations = self->maxNumIterations ;
  gmm->numRepetitions = self->numRepetitions ;
  gmm->verbosity = self->verbosity ;
  gmm->precision = self->precision ;
  gmm->LL = self->LL ;
  memcpy(gmm->means, self->means, size * self->numClusters * self->dimension) ;
  memcpy(gmm->covariances, self->covariances, size * self->numClusters * self->dimension) ;
  memcpy(gmm->priors, self->priors, size * self->numClusters) ;
  return gmm ;
}

/* ---------------------------------------------------------------- */
/*                                               Getters and setters */

VL_EXPORT vl_type vl_gmm_get_data_type (VlGMM const * self) { return self->dataType ; }
VL_EXPORT vl_size vl_gmm_get_num_clusters (VlGMM const * self) { return self->numClusters ; }
VL_EXPORT vl_size vl_gmm_get_num_data (VlGMM const * self) { return self->numData ; }
VL_EXPORT double vl_gmm_get_loglikelihood (VlGMM const * self) { return self->LL ; }
VL_EXPORT int vl_gmm_get_verbosity (VlGMM const * self) { return self->verbosity ; }
VL_EXPORT void vl_gmm_set_verbosity (VlGMM * self, int verbosity) { self->verbosity = verbosity ; }
VL_EXPORT void const * vl_gmm_get_means (VlGMM const * self) { return self->means ; }
VL_EXPORT void const * vl_gmm_get_covariances (VlGMM const * self) { return self->covariances ; }
VL_EXPORT void const * vl_gmm_get_priors (VlGMM const * self) { return self->priors ; }
VL_EXPORT vl_size vl_gmm_get_max_num_iterations (VlGMM const * self) { return self->maxNumIterations ; }
VL_EXPORT void vl_gmm_set_max_num_iterations (VlGMM * self, vl_size n) { self->maxNumIterations = n ; }
VL_EXPORT vl_size vl_gmm_get_num_repetitions (VlGMM const * self) { return self->numRepetitions ; }
VL_EXPORT vl_size vl_gmm_get_dimension (VlGMM const * self) { return self->dimension ; }
VL_EXPORT VlGMMInitialization vl_gmm_get_initialization (VlGMM const * self) { return self->initialization ; }
VL_EXPORT void vl_gmm_set_initialization (VlGMM * self, VlGMMInitialization init) { self->initialization = init ; }
VL_EXPORT VlKMeans * vl_gmm_get_kmeans_init_object (VlGMM const * self) { return self->kmeansInit ; }
VL_EXPORT double const * vl_gmm_get_covariance_lower_bounds (VlGMM const * self) { return self->sigmaLowBound ; }

/* No counterpart in the reference: arithmetic of the EM iterations of float
 * mixtures (dimension <= 128 for the matrix-core modes; other cases, and
 * double data, always use the FP32 / FP64 parity path). */
VL_EXPORT void
vl_gmm_set_precision (VlGMM * self, int precision)
{
  if (precision != VL_GMM_PRECISION_FP32 && precision != VL_GMM_PRECISION_BF16X3 &&
      precision != VL_GMM_PRECISION_F16X3) {
    vl_host_fatal("vl_gmm_set_precision: unknown precision %d\n", precision) ;
  }
  self->precision = precision ;
}

VL_EXPORT int vl_gmm_get_precision (VlGMM const * self) { return self->precision ; }

VL_EXPORT void
vl_gmm_set_num_repetitions (VlGMM * self, vl_size numRepetitions)
{
  assert(numRepetitions >= 1) ;
  self->numRepetitions = numRepetitions ;
}

VL_EXPORT void
vl_gmm_set_kmeans_init_object (VlGMM * self, VlKMeans * kmeans)
{
  if (self->kmeansInit && self->kmeansInitIsOwner) vl_kmeans_delete(self->kmeansInit) ;
  self->kmeansInit = kmeans ;
  self->kmeansInitIsOwner = VL_FALSE ;
}

VL_EXPORT void
vl_gmm_set_covariance_lower_bounds (VlGMM * self, double const * bounds)
{
  memcpy(self->sigmaLowBound, bounds, sizeof(double) * self->dimension) ;
}

VL_EXPORT void
vl_gmm_set_covariance_lower_bound (VlGMM * self, double bound)
{
  vl_uindex i ;
  for (i = 0 ; i < self->dimension ; ++i) self->sigmaLowBound[i] = bound ;
}

VL_EXPORT void
vl_gmm_set_means (VlGMM * self, void const * means)
{
  memcpy(self->means, means, self->dimension * self->numClusters * vl_get_type_size(self->dataType)) ;
}

VL_EXPORT void
vl_gmm_set_covariances (VlGMM * self, void const * covariances)
{
  memcpy(self->covariances, covariances, self->dimension * self->numClusters * vl_get_type_size(self->dataType)) ;
}

VL_EXPORT void
vl_gmm_set_priors (VlGMM * self, void const * priors)
{
  memcpy(self->priors, priors, self->numClusters * vl_get_type_size(self->dataType)) ;
}

/* Posteriors of the last E-step (numData x numClusters), copied from the GPU
 * on first request after EM. */
VL_EXPORT void const *
vl_gmm_get_posteriors (VlGMM const * cself)
{
  VlGMM * self = (VlGMM*) cself ;
  vl_size es = vl_get_type_size(self->dataType) ;
  if (self->postRows == 0) return self->posteriors ;
  if (self->devPostStale) {
    /* the matrix-core EM keeps no posterior matrix: the posteriors of the
       last E-step are those of the final parameters, recomputed once here
       with the FP32 E-step */
    unsigned char * p = (unsigned char*) self->dParams ;
    vl_size KD = self->numClusters * self->dimension ;
    vlhip_lock() ;
    gmm_upload_params(self) ;
    vlhip_gmm_posteriors((int)self->dataType, self->dPost, self->numClusters, self->postRows,
                         p + 2 * es * KD, p, self->dimension, p + es * KD, self->dData) ;
    vlhip_unlock() ;
    self->devPostStale = VL_FALSE ;
    self->hostPostValid = VL_FALSE ;
  }
  if (!self->hostPostValid) {
    vl_size bytes = es * self->numClusters * self->postRows ;
    if (!self->posteriors) self->posteriors = vl_malloc(bytes) ;
    vlhip_memcpy_d2h(self->posteriors, self->dPost, bytes) ;
    self->hostPostValid = VL_TRUE ;
  }
  return self->posteriors ;
}

/* ---------------------------------------------------------------- */
/*                                         Posteriors (standalone)   */

static double
posteriors_any (vl_type t, void * posteriors, vl_size numClusters, vl_size numData,
                void const * priors, void const * means, vl_size dimension,
                void const * covariances, void const * data)
{
  vl_size es = vl_get_type_size(t) ;
  vl_size KD = numClusters * dimension ;
  unsigned char * dp ;
  void * dx, * dq ;
  double LL ;
  vlhip_lock() ;
  dp = (unsigned char*) vlhip_buffer(VLHIP_BUF_PARAMS, es * (2 * KD + numClusters)) ;
  dx = vlhip_buffer(VLHIP_BUF_DATA, es * dimension * (numData ? numData : 1)) ;
  dq = vlhip_buffer(VLHIP_BUF_POST, es * numClusters * (numData ? numData : 1)) ;
  vlhip_memcpy_h2d(dp, means, es * KD) ;
  vlhip_memcpy_h2d(dp + es * KD, covariances, es * KD) ;
  vlhip_memcpy_h2d(dp + 2 * es * KD, priors, es * numClusters) ;
  if (numData) vlhip_memcpy_h2d(dx, data, es * dimension * numData) ;
  LL = vlhip_gmm_posteriors((int)t, dq, numClusters, numData, dp + 2 * es * KD, dp, dimension,
                            dp + es * KD, dx) ;
  if (numData) vlhip_memcpy_d2h(posteriors, dq, es * numClusters * numData) ;
  vlhip_unlock() ;
  return LL ;
}

VL_EXPORT double
vl_get_gmm_data_posteriors_f (float * posteriors, vl_size numClusters, vl_size numData,
                              float const * priors, float const * means, vl_size dimension,
                              float const * covariances, float const * data)
{
  return posteriors_any(VL_TYPE_FLOAT, posteriors, numClusters, numData, priors, means, dimension,
                        covariances, data) ;
}

VL_EXPORT double
vl_get_gmm_data_posteriors_d (double * posteriors, vl_size numClusters, vl_size numData,
                              double const * priors, double const * means, vl_size dimension,
                              double const * covariances, double const * data)
{
  return posteriors_any(VL_TYPE_DOUBLE, posteriors, numClusters, numData, priors, means, dimension,
                        covariances, data) ;
}

/* ---------------------------------------------------------------- */
/*                 Type-generic pieces (instantiated for float/double) */

#define GMM_TYPE float
#define GMM_SFX f
#include "gmm_impl.h"
#undef GMM_TYPE
#undef GMM_SFX

#define GMM_TYPE double
#define GMM_SFX d
#include "gmm_impl.h"
#undef GMM_TYPE
#undef GMM_SFX

/* ---------------------------------------------------------------- */
/*                                                   Public drivers  */

VL_EXPORT void
vl_gmm_init_with_rand_data (VlGMM * self, void const * data, vl_size numData)
{
  vl_gmm_reset(self) ;
  switch (self->dataType) {
    case VL_TYPE_FLOAT : gmm_init_rand_f(self, (float const*)data, numData) ; break ;
    case VL_TYPE_DOUBLE : gmm_init_rand_d(self, (double const*)data, numData) ; break ;
    default : abort() ;
  }
}

VL_EXPORT void
vl_gmm_init_with_kmeans (VlGMM * self, void const * data, vl_size numData, VlKMeans * kmeansInit)
{
  vl_gmm_reset(self) ;
  switch (self->dataType) {
    case VL_TYPE_FLOAT : gmm_init_kmeans_f(self, (float const*)data, numData, kmeansInit) ; break ;
    case VL_TYPE_DOUBLE : gmm_init_kmeans_d(self, (double const*)data, numData, kmeansInit) ; break ;
    default : abort() ;
  }
}

VL_EXPORT double
vl_gmm_em (VlGMM * self, void const * data, vl_size numData)
{
  switch (self->dataType) {
    case VL_TYPE_FLOAT : return gmm_em_f(self, (float const*)data, numData) ;
    case VL_TYPE_DOUBLE : return gmm_em_d(self, (double const*)data, numData) ;
    default : abort() ;
  }
  return 0 ;
}

/* gmm.c:1557-1648: repetitions keep the highest log-likelihood solution
 * (the first one as well, so a NaN likelihood still yields a model). */
VL_EXPORT double
vl_gmm_cluster (VlGMM * self, void const * data, vl_size numData)
{
  vl_size size = vl_get_type_size(self->dataType) ;
  vl_size KD = self->numClusters * self->dimension ;
  void * bestPriors = vl_malloc(size * self->numClusters) ;
  void * bestMeans = vl_malloc(size * KD) ;
  void * bestCovariances = vl_malloc(size * KD) ;
  void * bestPost = NULL ;
  vl_size bestPostBytes = 0 ;
  double bestLL = -VL_INFINITY_D ;
  vl_uindex repetition ;

  assert(self->numRepetitions >= 1) ;
  for (repetition = 0 ; repetition < self->numRepetitions ; ++repetition) {
    double LL ;
    double timeRef ;
    if (self->verbosity) {
      VL_PRINTF("gmm: fit %d of %d\n", (int)repetition + 1, (int)self->numRepetitions) ;
    }
    timeRef = vl_get_cpu_time() ;
    switch (self->initialization) {
      case VlGMMKMeans : vl_gmm_init_with_kmeans(self, data, numData, NULL) ; break ;
      case VlGMMRand : vl_gmm_init_with_rand_data(self, data, numData) ; break ;
      case VlGMMCustom : break ;
      default : abort() ;
    }
    if (self->verbosity) {
      VL_PRINTF("gmm: initial model ready after %.2f s\n", vl_get_cpu_time() - timeRef) ;
    }
    timeRef = vl_get_cpu_time() ;
    LL = vl_gmm_em(self, data, numData) ;
    if (self->verbosity) {
      VL_PRINTF("gmm: EM done after %.2f s, log-likelihood %f\n",
                vl_get_cpu_time() - timeRef, LL) ;
    }
    if (LL > bestLL || repetition == 0) {
      void * t ;
      vl_size tb ;
      t = bestPriors ; bestPriors = self->priors ; self->priors = t ;
      t = bestMeans ; bestMeans = self->means ; self->means = t ;
      t = bestCovariances ; bestCovariances = self->covariances ; self->covariances = t ;
      /* keep the device posteriors of the best repetition */
      t = bestPost ; bestPost = self->dPost ; self->dPost = t ;
      tb = bestPostBytes ; bestPostBytes = self->dPostBytes ; self->dPostBytes = tb ;
      bestLL = LL ;
    }
  }
  vl_free(self->priors) ;
  vl_free(self->means) ;
  vl_free(self->covariances) ;
  if (self->dPost) vlhip_free(self->dPost) ;
  self->priors = bestPriors ;
  self->means = bestMeans ;
  self->covariances = bestCovariances ;
  self->dPost = bestPost ;
  self->dPostBytes = bestPostBytes ;
  self->postRows = numData ;
  self->hostPostValid = VL_FALSE ;
  if (self->posteriors) { vl_free(self->posteriors) ; self->posteriors = NULL ; }
  self->LL = bestLL ;
  if (self->verbosity) {
    VL_PRINTF("gmm: best log-likelihood over the fits %f\n", self->LL) ;
  }
  return bestLL ;
}
