/* gmm_impl.h -- type-generic part of gmm.c, included once with
 * GMM_TYPE = float / GMM_SFX = f and once with double / d. */

#define FN(name) VL_XCAT(name##_, GMM_SFX)
#define TYPE GMM_TYPE
#define TYPE_CODE ((sizeof(TYPE) == sizeof(float)) ? VL_TYPE_FLOAT : VL_TYPE_DOUBLE)

/* gmm.c:1005-1030 */
static void
FN(gmm_apply_bounds) (VlGMM * self)
{
  vl_uindex k, d ;
  vl_size numAdjusted = 0 ;
  TYPE * cov = (TYPE*) self->covariances ;
  double const * lbs = self->sigmaLowBound ;
  for (k = 0 ; k < self->numClusters ; ++k) {
    vl_bool adjusted = VL_FALSE ;
    for (d = 0 ; d < self->dimension ; ++d) {
      if (cov[k * self->dimension + d] < lbs[d]) {
        cov[k * self->dimension + d] = (TYPE) lbs[d] ;
        adjusted = VL_TRUE ;
      }
    }
    if (adjusted) numAdjusted ++ ;
  }
  if (numAdjusted > 0 && self->verbosity > 0) {
    VL_PRINT("gmm: detected %d of %d modes with at least one dimension "
             "with covariance too small (set to lower bound)\n",
             (int)numAdjusted, (int)self->numClusters) ;
  }
}

/* Moments of the current device posteriors against the current means. */
static void
FN(gmm_moments) (VlGMM * self, double * out, vl_size numData)
{
  vl_size KD = self->numClusters * self->dimension ;
  unsigned char * p = (unsigned char*) self->dParams ;
  TYPE * dOld = (TYPE*)(p + sizeof(TYPE) * (2 * KD + self->numClusters)) ;
  vlhip_memcpy_h2d(dOld, self->means, sizeof(TYPE) * KD) ;
  vlhip_gmm_moments(TYPE_CODE, out, self->dPost, self->dData, dOld, numData,
                    self->dimension, self->numClusters, GMM_MIN_POSTERIOR / self->numClusters) ;
}

/* M-step finalisation, gmm.c:1169-1202, from moments computed with
 * oldMeans = the current means.  out = [S0 K | mass K | S1 KD | S2 KD]. */
static void
FN(gmm_finalize) (VlGMM * self, double const * out)
{
  vl_size K = self->numClusters, D = self->dimension ;
  TYPE * priors = (TYPE*) self->priors ;
  TYPE * means = (TYPE*) self->means ;
  TYPE * covs = (TYPE*) self->covariances ;
  double const * S0 = out ;
  double const * S1 = out + 2 * K ;
  double const * S2 = S1 + K * D ;
  vl_uindex k, d ;
  TYPE sum = 0 ;
  for (k = 0 ; k < K ; ++k) {
    TYPE mass = (TYPE) S0[k] ;
    vl_bool update = (mass >= 1e-6 / K) ;
    priors[k] = mass ;
    for (d = 0 ; d < D ; ++d) {
      TYPE oldMu = means[k * D + d] ;
      if (update) {
        TYPE mu = (TYPE)(S1[k * D + d] / S0[k]) ;
        TYPE cv = (TYPE)(S2[k * D + d] / S0[k]) ;
        TYPE diff = mu - oldMu ;
        means[k * D + d] = mu ;
        covs[k * D + d] = cv - diff * diff ;
      } else {
        means[k * D + d] = (TYPE) S1[k * D + d] ;
        covs[k * D + d] = (TYPE) S2[k * D + d] ;
      }
    }
  }
  FN(gmm_apply_bounds)(self) ;
  for (k = 0 ; k < K ; ++k) sum += priors[k] ;
  sum = (TYPE) VL_MAX((double)sum, 1e-12) ;
  for (k = 0 ; k < K ; ++k) priors[k] /= sum ;
}

/* Empty-component restarts, the rule of gmm.c:815-999.  A component whose
 * total posterior mass (massIn, no threshold) is below
 * GMM_MIN_POSTERIOR * max(1, N / K) is refilled by splitting the component
 * with the largest term in the mixture's entropy,
 *     pi_k (D/2 (1 + log 2 pi) + 1/2 sum_d log sigma_kd^2 - log pi_k)
 * (summed in that order, so that the same donor wins as in the reference),
 * along the donor's dimension of largest variance at its mean: the GPU moves
 * the donor's posteriors on one side to the empty component, and the M-step
 * is redone after every split.  Returns the number of splits. */
static vl_size
FN(gmm_restart) (VlGMM * self, vl_size numData, double const * massIn)
{
  vl_size const K = self->numClusters, D = self->dimension ;
  TYPE const * priors = (TYPE const*) self->priors ;
  TYPE const * means = (TYPE const*) self->means ;
  TYPE const * covs = (TYPE const*) self->covariances ;
  double const threshold = GMM_MIN_POSTERIOR * VL_MAX(1.0, (double)numData / K) ;
  double * mass ;
  double * moments ;
  vl_index target, k, d ;
  vl_size splits = 0 ;
  if (K <= 1) return 0 ;
  mass = (double*) vl_malloc(sizeof(double) * K) ;
  memcpy(mass, massIn, sizeof(double) * K) ;
  moments = (double*) vl_malloc(sizeof(double) * (2 * K + 2 * K * D)) ;

  for (target = 0 ; target < (signed)K ; ++target) {
    double donorTerm = -VL_INFINITY_D, widest = -VL_INFINITY_D ;
    vl_index donor = -1, axis = -1 ;
    if (mass[target] >= threshold) continue ;
    if (self->verbosity) {
      VL_PRINTF("gmm: component %d carries almost no posterior mass (%f)\n", (int)target, mass[target]) ;
    }
    for (k = 0 ; k < (signed)K ; ++k) {
      double term ;
      if (priors[k] < GMM_MIN_PRIOR) continue ;
      term = 0.5 * D * (1.0 + log(2 * VL_PI)) ;
      for (d = 0 ; d < (signed)D ; ++d) term += 0.5 * log((double)covs[k * D + d]) ;
      term = priors[k] * (term - log((double)priors[k])) ;
      if (term > donorTerm) { donorTerm = term ; donor = k ; }
    }
    if (donor == target || donor < 0) {
      if (self->verbosity) {
        VL_PRINTF("gmm: no other component can be split to refill component %d\n", (int)target) ;
      }
      continue ;
    }
    if (self->verbosity) {
      VL_PRINTF("gmm: refilling component %d by splitting component %d (prior %f, mass %f, entropy term %f)\n",
                (int)target, (int)donor, (double)priors[donor], mass[donor], donorTerm) ;
    }
    for (d = 0 ; d < (signed)D ; ++d) {
      double const v = covs[donor * D + d] ;
      if (v > widest) { widest = v ; axis = d ; }
    }
    vlhip_gmm_split(TYPE_CODE, self->dPost, self->dData, numData, D, K, (size_t)target, (size_t)donor,
                    (size_t)axis, (double)means[axis + donor * D]) ;
    FN(gmm_moments)(self, moments, numData) ;
    FN(gmm_finalize)(self, moments) ;
    splits ++ ;
  }
  vl_free(moments) ;
  vl_free(mass) ;
  return splits ;
}

/* Whether some component would be restarted (the test of gmm.c:897-899). */
static vl_bool
FN(gmm_needs_restart) (VlGMM * self, vl_size numData, double const * mass)
{
  vl_uindex k ;
  if (self->numClusters <= 1) return VL_FALSE ;
  for (k = 0 ; k < self->numClusters ; ++k) {
    /* the complement of the skip test of gmm_restart, so that no component
       it would restart is missed */
    if (!(mass[k] >= GMM_MIN_POSTERIOR * VL_MAX(1.0, (double)numData / self->numClusters))) return VL_TRUE ;
  }
  return VL_FALSE ;
}

/* EM iterations, gmm.c:1220-1308.  With a matrix-core precision (float data,
 * dimension <= 128) each iteration is one vlhip_gmm_em_step: log-likelihood
 * and M-step moments without a posterior matrix.  An iteration that may
 * restart an empty component recomputes its posteriors and moments with the
 * FP32 parity path first, so that the split of gmm.c:968-990 works on the
 * posterior matrix exactly as in the FP32 mode. */
static double
FN(gmm_em) (VlGMM * self, TYPE const * data, vl_size numData)
{
  vl_size K = self->numClusters, D = self->dimension ;
  vl_size KD = K * D ;
  vl_size iteration ;
  double previousLL = -VL_INFINITY_D ;
  double LL = -VL_INFINITY_D ;
  double * out = (double*) vl_malloc(sizeof(double) * (2 * K + 2 * KD)) ;
  unsigned char * p ;
  int mfma = (TYPE_CODE == VL_TYPE_FLOAT && self->precision != VL_GMM_PRECISION_FP32 &&
              D <= 128 && numData > 0) ;
  int mode = (self->precision == VL_GMM_PRECISION_F16X3) ? VLHIP_MODE_F16X3 : VLHIP_MODE_BF16X3 ;

  vlhip_lock() ;
  _vl_gmm_prepare_for_data(self, numData) ;
  gmm_stage_data(self, data, numData) ;
  self->postRows = numData ;
  self->devPostStale = VL_FALSE ;
  p = (unsigned char*) self->dParams ;
  FN(gmm_apply_bounds)(self) ;

  for (iteration = 0 ; 1 ; ++iteration) {
    double eps ;
    double time = vl_get_cpu_time() ;
    gmm_upload_params(self) ;
    if (mfma) {
      /* the last iteration (iteration == maxNumIterations) needs no moments */
      LL = vlhip_gmm_em_step(mode, (iteration < self->maxNumIterations) ? out : NULL, p + 2 * sizeof(TYPE) * KD, p,
                             p + sizeof(TYPE) * KD, D, K, self->dData, numData) ;
      self->devPostStale = VL_TRUE ;
    } else {
      LL = vlhip_gmm_posteriors(TYPE_CODE, self->dPost, K, numData,
                                p + 2 * sizeof(TYPE) * KD, p, D, p + sizeof(TYPE) * KD, self->dData) ;
    }
    if (self->verbosity > 1) {
      VL_PRINTF("gmm: EM: posteriors of all data in %.2f s\n", vl_get_cpu_time() - time) ;
    }
    if (self->verbosity) {
      VL_PRINTF("gmm: EM step %d: log-likelihood %f, change %f\n",
                (int)iteration, LL, LL - previousLL) ;
    }
    if (iteration >= self->maxNumIterations) {
      if (self->verbosity) {
        VL_PRINTF("gmm: EM stops after the iteration limit (%d)\n",
                  (int)self->maxNumIterations) ;
      }
      break ;
    }
    eps = fabs((LL - previousLL) / LL) ;
    if ((iteration > 0) && (eps < 0.00001)) {
      if (self->verbosity) {
        VL_PRINTF("gmm: EM stops: relative log-likelihood change %f below the tolerance\n", eps) ;
      }
      break ;
    }
    previousLL = LL ;

    if (!mfma) FN(gmm_moments)(self, out, numData) ;
    if (iteration > 1) {
      if (mfma && FN(gmm_needs_restart)(self, numData, out + K)) {
        /* this iteration's posteriors and moments on the parity path */
        vlhip_gmm_posteriors(TYPE_CODE, self->dPost, K, numData,
                             p + 2 * sizeof(TYPE) * KD, p, D, p + sizeof(TYPE) * KD, self->dData) ;
        self->devPostStale = VL_FALSE ;
        FN(gmm_moments)(self, out, numData) ;
      }
      if (FN(gmm_restart)(self, numData, out + K) > 0) {
        FN(gmm_moments)(self, out, numData) ;
      }
    }
    FN(gmm_finalize)(self, out) ;
  }
  self->hostPostValid = VL_FALSE ;
  vlhip_unlock() ;
  vl_free(out) ;
  return LL ;
}

/* gmm.c:1315-1370 */
static void
FN(gmm_init_kmeans) (VlGMM * self, TYPE const * data, vl_size numData, VlKMeans * kmeansInit)
{
  vl_size K = self->numClusters, D = self->dimension ;
  vl_uint32 * assignments = (vl_uint32*) vl_malloc(sizeof(vl_uint32) * (numData ? numData : 1)) ;
  double * out = (double*) vl_malloc(sizeof(double) * (2 * K + 2 * K * D)) ;
  vl_uint32 * dAssign ;

  _vl_gmm_prepare_for_data(self, numData) ;
  memset(self->means, 0, sizeof(TYPE) * K * D) ;
  memset(self->priors, 0, sizeof(TYPE) * K) ;
  memset(self->covariances, 0, sizeof(TYPE) * K * D) ;

  if (kmeansInit) vl_gmm_set_kmeans_init_object(self, kmeansInit) ;
  if (self->kmeansInit == NULL) {
    /* the reference's default initialiser, gmm.c:1336-1355 */
    vl_size ncomparisons = VL_MAX(numData / 4, 10) ;
    VlKMeans * km = vl_kmeans_new(self->dataType, VlDistanceL2) ;
    vl_kmeans_set_initialization(km, VlKMeansRandomSelection) ;
    vl_kmeans_set_max_num_iterations(km, 5) ;
    vl_kmeans_set_max_num_comparisons(km, ncomparisons) ;
    vl_kmeans_set_num_trees(km, 1) ;
    vl_kmeans_set_algorithm(km, VlKMeansANN) ;
    vl_kmeans_set_num_repetitions(km, 1) ;
    vl_kmeans_set_verbosity(km, self->verbosity) ;
    self->kmeansInit = km ;
    self->kmeansInitIsOwner = VL_TRUE ;
  }

  vl_kmeans_cluster(self->kmeansInit, data, D, numData, K) ;
  vl_kmeans_quantize(self->kmeansInit, assignments, NULL, data, numData) ;

  vlhip_lock() ;
  gmm_stage_data(self, data, numData) ;
  dAssign = (vl_uint32*) vlhip_buffer(VLHIP_BUF_AUX, sizeof(vl_uint32) * (numData ? numData : 1)) ;
  if (numData) vlhip_memcpy_h2d(dAssign, assignments, sizeof(vl_uint32) * numData) ;
  vlhip_gmm_onehot(TYPE_CODE, self->dPost, dAssign, numData, K) ;
  self->postRows = numData ;
  self->hostPostValid = VL_FALSE ;
  FN(gmm_moments)(self, out, numData) ;
  FN(gmm_finalize)(self, out) ;
  vlhip_unlock() ;
  vl_free(out) ;
  vl_free(assignments) ;
}

/* gmm.c:1376-1413: per-dimension unbiased variance, summed in TYPE in data order. */
static void
FN(gmm_init_sigma) (TYPE const * data, TYPE * initSigma, vl_size D, vl_size numData)
{
  vl_uindex i, d ;
  TYPE * dataMean ;
  memset(initSigma, 0, sizeof(TYPE) * D) ;
  if (numData <= 1) return ;
  dataMean = (TYPE*) vl_calloc(D, sizeof(TYPE)) ;
  for (d = 0 ; d < D ; ++d) {
    for (i = 0 ; i < numData ; ++i) dataMean[d] += data[i * D + d] ;
    dataMean[d] /= numData ;
  }
  for (d = 0 ; d < D ; ++d) {
    for (i = 0 ; i < numData ; ++i) {
      TYPE diff = data[i * D + d] - dataMean[d] ;
      initSigma[d] += diff * diff ;
    }
    initSigma[d] /= numData - 1 ;
  }
  vl_free(dataMean) ;
}

/* gmm.c:1415-1443 */
static void
FN(gmm_init_rand) (VlGMM * self, TYPE const * data, vl_size numData)
{
  vl_size K = self->numClusters, D = self->dimension ;
  vl_uindex i, k, d ;
  VlKMeans * kmeans ;
  _vl_gmm_prepare_for_data(self, numData) ;
  for (i = 0 ; i < K ; ++i) ((TYPE*)self->priors)[i] = (TYPE)(1.0 / K) ;
  FN(gmm_init_sigma)(data, (TYPE*)self->covariances, D, numData) ;
  for (k = 1 ; k < K ; ++k)
    for (d = 0 ; d < D ; ++d)
      ((TYPE*)self->covariances)[k * D + d] = ((TYPE*)self->covariances)[d] ;
  kmeans = vl_kmeans_new(self->dataType, VlDistanceL2) ;
  vl_kmeans_init_centers_plus_plus(kmeans, data, D, numData, K) ;
  memcpy(self->means, vl_kmeans_get_centers(kmeans), sizeof(TYPE) * D * K) ;
  vl_kmeans_delete(kmeans) ;
}

#undef FN
#undef TYPE
#undef TYPE_CODE
