/* kmeans_impl.h -- type-generic part of kmeans.c, included with
 * KM_TYPE = float / KM_SFX = f and with double / d. */

#define FN(name) VL_XCAT(name##_, KM_SFX)
#define TYPE KM_TYPE
#define TYPE_CODE ((sizeof(TYPE) == sizeof(float)) ? VL_TYPE_FLOAT : VL_TYPE_DOUBLE)
#define DISTFN(t) ((sizeof(TYPE) == sizeof(float)) ? (void*)vl_get_vector_comparison_function_f(t) \
                                                   : (void*)vl_get_vector_comparison_function_d(t))

typedef TYPE (*FN(km_dist_t)) (vl_size, TYPE const *, TYPE const *) ;

/* Knuth shuffle with the calling thread's generator (shuffle-def.h:82-93). */
static void
FN(km_shuffle) (vl_uindex * array, vl_size size, VlRand * rand)
{
  vl_uindex n = size ;
  while (n > 1) {
    vl_uindex k = vl_rand_uindex(rand, n) ;
    vl_uindex t ;
    n -- ;
    t = array[n] ; array[n] = array[k] ; array[k] = t ;
  }
}

/* kmeans.c:527-582: distinct random data points as centers. */
static void
FN(km_init_rand) (VlKMeans * self, TYPE const * data, vl_size dimension, vl_size numData, vl_size numCenters)
{
  vl_uindex i, j, k ;
  VlRand * rand = vl_get_rand() ;
  FN(km_dist_t) dist = (FN(km_dist_t)) DISTFN(self->distance) ;
  vl_uindex * perm = (vl_uindex*) vl_malloc(sizeof(vl_uindex) * (numData ? numData : 1)) ;
  self->dimension = dimension ;
  self->numCenters = numCenters ;
  self->centers = vl_malloc(sizeof(TYPE) * dimension * numCenters) ;
  for (i = 0 ; i < numData ; ++i) perm[i] = i ;
  FN(km_shuffle)(perm, numData, rand) ;
  for (k = 0, i = 0 ; k < numCenters ; ++i) {
    if (numCenters - k < numData - i) {
      vl_bool duplicate = VL_FALSE ;
      for (j = 0 ; j < k ; ++j) {
        TYPE dj = dist(dimension, data + dimension * perm[i], (TYPE*)self->centers + dimension * j) ;
        duplicate |= (dj == 0) ;
      }
      if (duplicate) continue ;
    }
    memcpy((TYPE*)self->centers + dimension * k, data + dimension * perm[i], sizeof(TYPE) * dimension) ;
    k ++ ;
  }
  vl_free(perm) ;
}

/* kmeans.c:588-649: k-means++ seeding.  The distances of each new center to
 * all data come from the GPU (host lane order, bitwise the host function);
 * the running minima, energy and sampling walk run here in data order. */
static void
FN(km_init_pp) (VlKMeans * self, TYPE const * data, vl_size dimension, vl_size numData, vl_size numCenters)
{
  vl_uindex x, c ;
  VlRand * rand = vl_get_rand() ;
  TYPE * distances = (TYPE*) vl_malloc(sizeof(TYPE) * (numData ? numData : 1)) ;
  TYPE * minDistances = (TYPE*) vl_malloc(sizeof(TYPE) * (numData ? numData : 1)) ;
  void * dData, * dPoint, * dOut ;
  self->dimension = dimension ;
  self->numCenters = numCenters ;
  self->centers = vl_malloc(sizeof(TYPE) * dimension * numCenters) ;
  for (x = 0 ; x < numData ; ++x) minDistances[x] = (TYPE) VL_INFINITY_D ;

  vlhip_lock() ;
  dData = vlhip_buffer(VLHIP_BUF_DATA, sizeof(TYPE) * dimension * (numData ? numData : 1)) ;
  dPoint = vlhip_buffer(VLHIP_BUF_PARAMS, sizeof(TYPE) * dimension) ;
  dOut = vlhip_buffer(VLHIP_BUF_AUX, sizeof(TYPE) * (numData ? numData : 1)) ;
  if (numData) vlhip_memcpy_h2d(dData, data, sizeof(TYPE) * dimension * numData) ;

  x = vl_rand_uindex(rand, numData) ;
  c = 0 ;
  while (1) {
    TYPE energy = 0 ;
    TYPE acc = 0 ;
    TYPE thresh = (TYPE) vl_rand_real1(rand) ;
    memcpy((TYPE*)self->centers + c * dimension, data + x * dimension, sizeof(TYPE) * dimension) ;
    c ++ ;
    if (c == numCenters) break ;
    vlhip_memcpy_h2d(dPoint, (TYPE*)self->centers + (c - 1) * dimension, sizeof(TYPE) * dimension) ;
    vlhip_l2_to_point(TYPE_CODE, dOut, dData, dPoint, numData, dimension, host_lanes(self->dataType)) ;
    vlhip_memcpy_d2h(distances, dOut, sizeof(TYPE) * numData) ;
    for (x = 0 ; x < numData ; ++x) {
      minDistances[x] = VL_MIN(minDistances[x], distances[x]) ;
      energy += minDistances[x] ;
    }
    for (x = 0 ; x < numData - 1 ; ++x) {
      acc += minDistances[x] ;
      if (acc >= thresh * energy) break ;
    }
  }
  vlhip_unlock() ;
  vl_free(distances) ;
  vl_free(minDistances) ;
}

/* kmeans.c:708-771: assignments from a freshly built kd-forest over the
 * centers.  With update, a datum keeps its previous center unless the forest
 * found a strictly closer one.  When the comparison budget covers every
 * center (or is 0, exact search) the branch-and-bound search is exhaustive
 * and returns the exact nearest center, so the assignment runs on the GPU
 * quantizer; the forest is still built so the generator advances exactly as
 * in the reference. */
static void
FN(km_quantize_ann) (VlKMeans * self, vl_uint32 * assignments, TYPE * distances,
                     TYPE const * data, vl_size numData, vl_bool update)
{
  VlKDForest * forest = vl_kdforest_new(self->dataType, self->dimension, self->numTrees, self->distance) ;
  vl_bool exhaustive = (self->maxNumComparisons == 0 || self->maxNumComparisons >= self->numCenters) ;
  FN(km_dist_t) dist = (FN(km_dist_t)) DISTFN(self->distance) ;
  vl_uindex x ;
  vl_kdforest_set_max_num_comparisons(forest, self->maxNumComparisons) ;
  vl_kdforest_set_thresholding_method(forest, VL_KDTREE_MEDIAN) ;
  vl_kdforest_build(forest, self->numCenters, self->centers) ;

  if (exhaustive) {
    vl_uint32 * na = (vl_uint32*) vl_malloc(sizeof(vl_uint32) * (numData ? numData : 1)) ;
    TYPE * nd = (TYPE*) vl_malloc(sizeof(TYPE) * (numData ? numData : 1)) ;
    void * dData ;
    vlhip_lock() ;
    dData = vlhip_buffer(VLHIP_BUF_DATA, sizeof(TYPE) * self->dimension * (numData ? numData : 1)) ;
    if (numData) vlhip_memcpy_h2d(dData, data, sizeof(TYPE) * self->dimension * numData) ;
    quantize_staged(self, dData, numData, na, nd) ;
    vlhip_unlock() ;
    for (x = 0 ; x < numData ; ++x) {
      if (distances) {
        if (!update) {
          distances[x] = nd[x] ;
          assignments[x] = na[x] ;
        } else {
          TYPE prev = dist(self->dimension, data + self->dimension * x,
                           (TYPE*)self->centers + self->dimension * assignments[x]) ;
          if (prev > nd[x]) { distances[x] = nd[x] ; assignments[x] = na[x] ; }
          else distances[x] = prev ;
        }
      } else {
        assignments[x] = na[x] ;
      }
    }
    vl_free(na) ;
    vl_free(nd) ;
  } else {
    VlKDForestSearcher * searcher = vl_kdforest_new_searcher(forest) ;
    VlKDForestNeighbor nb ;
    for (x = 0 ; x < numData ; ++x) {
      vl_kdforestsearcher_query(searcher, &nb, 1, data + x * self->dimension) ;
      if (distances) {
        if (!update) {
          distances[x] = (TYPE) nb.distance ;
          assignments[x] = (vl_uint32) nb.index ;
        } else {
          TYPE prev = dist(self->dimension, data + self->dimension * x,
                           (TYPE*)self->centers + self->dimension * assignments[x]) ;
          if (prev > (TYPE) nb.distance) {
            distances[x] = (TYPE) nb.distance ;
            assignments[x] = (vl_uint32) nb.index ;
          } else {
            distances[x] = prev ;
          }
        }
      } else {
        assignments[x] = (vl_uint32) nb.index ;
      }
    }
  }
  vl_kdforest_delete(forest) ;
}

/* The kd-forest build of an ANN quantization (kmeans.c:723-726) draws from
 * the calling thread's generator; an exhaustive search does not need the
 * forest, but it is built so that the generator advances as in the reference. */
static void
FN(km_forest_draws) (VlKMeans * self)
{
  VlKDForest * forest = vl_kdforest_new(self->dataType, self->dimension, self->numTrees, self->distance) ;
  vl_kdforest_set_max_num_comparisons(forest, self->maxNumComparisons) ;
  vl_kdforest_set_thresholding_method(forest, VL_KDTREE_MEDIAN) ;
  vl_kdforest_build(forest, self->numCenters, self->centers) ;
  vl_kdforest_delete(forest) ;
}

/* Lloyd (kmeans.c:828-978) and ANN (kmeans.c:1003-1156) refinement; the two
 * differ only in how assignments are obtained.  The data are staged on the
 * GPU once.  Per iteration the GPU computes the assignments (exact quantizer;
 * for an exhaustive ANN search also the keep-previous rule of kmeans.c:750-763)
 * and the new centers; the host sums the energy in data order in double
 * (kmeans.c:862-864), counts the cluster masses, lists each cluster's data in
 * data order and draws the restart datum of every empty cluster in cluster
 * order (kmeans.c:919-925), so the generator and every center agree with the
 * reference bit for bit.  A non-exhaustive ANN search runs the kd-forest on
 * the host. */
static double
FN(km_refine) (VlKMeans * self, TYPE const * data, vl_size numData)
{
  vl_size const K = self->numCenters, D = self->dimension ;
  vl_size const nd = numData ? numData : 1 ;
  vl_size const es = sizeof(TYPE) ;
  vl_size c, x, iteration ;
  double previousEnergy = VL_INFINITY_D ;
  double initialEnergy = VL_INFINITY_D ;
  double energy ;
  vl_bool const ann = (self->algorithm == VlKMeansANN) ;
  vl_bool const annExhaustive = ann && (self->maxNumComparisons == 0 || self->maxNumComparisons >= K) ;
  int const lanes = host_lanes(self->dataType) ;
  TYPE * distances = (TYPE*) vl_malloc(es * nd) ;
  vl_uint32 * assignments = (vl_uint32*) vl_malloc(sizeof(vl_uint32) * nd) ;
  vl_size * clusterMasses = (vl_size*) vl_malloc(sizeof(vl_size) * (K ? K : 1)) ;
  vl_size * cursor = (vl_size*) vl_malloc(sizeof(vl_size) * (K ? K : 1)) ;
  /* one upload per iteration: members (numData indices), then the cluster
   * starts (K + 1) and the restart data (K) as 64-bit values */
  vl_size const memBytes = (sizeof(vl_uint32) * nd + 7) & ~(vl_size)7 ;
  vl_size const updBytes = memBytes + sizeof(vl_uint64) * (2 * K + 1) ;
  unsigned char * upd = (unsigned char*) vl_malloc(updBytes) ;
  vl_uint32 * members = (vl_uint32*) upd ;
  vl_uint64 * starts = (vl_uint64*) (upd + memBytes) ;
  vl_uint64 * restartRow = starts + K + 1 ;
  VlRand * rand = vl_get_rand() ;
  void * dData, * dCenters, * dDist ;
  unsigned char * dUpd ;
  vl_uint32 * dAssign, * dPrev ;

  if (numData > 0xffffffffu) {
    VL_PRINTF("vl_kmeans_refine_centers: more than 2^32 data are not supported\n") ;
    abort() ;
  }
  vlhip_lock() ;
  dData = vlhip_buffer(VLHIP_BUF_DATA, es * D * nd) ;
  dCenters = vlhip_buffer(VLHIP_BUF_PARAMS, es * D * (K ? K : 1)) ;
  dAssign = (vl_uint32*) vlhip_buffer(VLHIP_BUF_AUX, sizeof(vl_uint32) * nd) ;
  dDist = vlhip_buffer(VLHIP_BUF_AUX2, es * nd) ;
  dPrev = (vl_uint32*) vlhip_buffer(VLHIP_BUF_POST, sizeof(vl_uint32) * nd) ;
  dUpd = (unsigned char*) vlhip_buffer(VLHIP_BUF_ENC, updBytes) ;
  if (numData) vlhip_memcpy_h2d(dData, data, es * D * numData) ;

  for (energy = VL_INFINITY_D, iteration = 0 ; 1 ; ++iteration) {
    vl_size numRestarted = 0 ;
    if (ann && !annExhaustive) {
      FN(km_quantize_ann)(self, assignments, distances, data, numData, iteration > 0) ;
    } else {
      if (annExhaustive) FN(km_forest_draws)(self) ;
      vlhip_memcpy_h2d(dCenters, self->centers, es * D * K) ;
      if (annExhaustive && iteration > 0 && numData) {
        vlhip_memcpy_d2d(dPrev, dAssign, sizeof(vl_uint32) * numData) ;
      }
      vlhip_kmeans_quantize(TYPE_CODE, dAssign, dDist, dData, dCenters, numData, D, K, lanes) ;
      if (annExhaustive && iteration > 0) {
        vlhip_kmeans_ann_merge(TYPE_CODE, dAssign, dDist, dPrev, dData, dCenters, numData, D, lanes) ;
      }
      if (numData) {
        vlhip_memcpy_d2h(assignments, dAssign, sizeof(vl_uint32) * numData) ;
        vlhip_memcpy_d2h(distances, dDist, es * numData) ;
      }
    }
    energy = 0 ;
    for (x = 0 ; x < numData ; ++x) energy += distances[x] ;
    if (self->verbosity) {
      VL_PRINTF("kmeans: %s step %d, energy %g\n", ann ? "ANN" : "Lloyd", (int)iteration, energy) ;
    }
    if (iteration >= self->maxNumIterations) break ;
    if (energy == previousEnergy) break ;
    if (iteration == 0) {
      initialEnergy = energy ;
    } else {
      double eps = (previousEnergy - energy) / (initialEnergy - energy) ;
      if (eps < self->minEnergyVariation) break ;
    }
    previousEnergy = energy ;
    /* the reference would draw restart data from an empty set here */
    if (numData == 0) break ;

    memset(clusterMasses, 0, sizeof(vl_size) * K) ;
    for (x = 0 ; x < numData ; ++x) clusterMasses[assignments[x]] ++ ;
    starts[0] = 0 ;
    for (c = 0 ; c < K ; ++c) {
      cursor[c] = starts[c] ;
      starts[c + 1] = starts[c] + clusterMasses[c] ;
    }
    for (x = 0 ; x < numData ; ++x) members[cursor[assignments[x]] ++] = (vl_uint32) x ;
    for (c = 0 ; c < K ; ++c) {
      restartRow[c] = 0 ;
      if (clusterMasses[c] == 0) {
        restartRow[c] = vl_rand_uindex(rand, numData) ;
        numRestarted ++ ;
      }
    }
    vlhip_memcpy_h2d(dUpd, upd, updBytes) ;
    vlhip_kmeans_update_centers(TYPE_CODE, dCenters, dData, (vl_uint32 const*) dUpd,
                                (uint64_t const*) (dUpd + memBytes), (uint64_t const*) (dUpd + memBytes) + K + 1,
                                K, D) ;
    vlhip_memcpy_d2h(self->centers, dCenters, es * D * K) ;
    if (self->verbosity && numRestarted) {
      VL_PRINTF("kmeans: step %d moved %d empty centers to random data\n", (int)iteration, (int)numRestarted) ;
    }
  }
  vlhip_unlock() ;
  vl_free(distances) ;
  vl_free(assignments) ;
  vl_free(clusterMasses) ;
  vl_free(cursor) ;
  vl_free(upd) ;
  return energy ;
}

/* Elkan refinement (kmeans.c:1162-1745).  The per-datum steps (initial
 * assignment, moving the bounds, the bound-guided reassignment, the final
 * distances) run on the GPU (vlhip_kmeans_elkan); the host does what the
 * reference does once per iteration for all data: cluster masses, the new
 * centers' data lists and restart draws (the same center update as Lloyd's,
 * kmeans.c:1313-1345), each center's move, the center-center distances
 * (kept in self->centerDistances, as the reference keeps them) and each
 * center's nearest-center distance, all with the host's lane-order distance.
 * The loop stops after maxNumIterations iterations or when no datum changed
 * center; the returned energy is the sum of the final exact distances in data
 * order. */
static double
FN(km_refine_elkan) (VlKMeans * self, TYPE const * data, vl_size numData)
{
  vl_size const K = self->numCenters, D = self->dimension, es = sizeof(TYPE) ;
  vl_size c, j, x, iteration ;
  int const lanes = host_lanes(self->dataType) ;
  FN(km_dist_t) distFn = (FN(km_dist_t)) DISTFN(VlDistanceL2) ;
  VlRand * rand = vl_get_rand() ;
  vl_uint32 * assignments ;
  vl_size * clusterMasses, * cursor ;
  TYPE * oldCenters, * move, * nearest, * cd, * dist ;
  vl_size memBytes, updBytes ;
  unsigned char * upd ;
  vl_uint32 * members ;
  vl_uint64 * starts, * restartRow ;
  void * dData, * dCenters, * dTable, * dMove, * dNearest, * dUB, * dLB, * dDist ;
  unsigned char * dExact, * dUpd ;
  vl_uint32 * dAssign ;
  unsigned * dMoved ;
  double energy = 0 ;

  if (numData == 0) return 0 ;
  if (numData > 0xffffffffu) {
    VL_PRINTF("vl_kmeans_refine_centers: more than 2^32 data are not supported\n") ;
    abort() ;
  }
  if (!self->centerDistances) self->centerDistances = vl_malloc(es * K * K) ;
  cd = (TYPE*) self->centerDistances ;
  assignments = (vl_uint32*) vl_malloc(sizeof(vl_uint32) * numData) ;
  clusterMasses = (vl_size*) vl_malloc(sizeof(vl_size) * K) ;
  cursor = (vl_size*) vl_malloc(sizeof(vl_size) * K) ;
  oldCenters = (TYPE*) vl_malloc(es * D * K) ;
  move = (TYPE*) vl_malloc(es * K) ;
  nearest = (TYPE*) vl_malloc(es * K) ;
  dist = (TYPE*) vl_malloc(es * numData) ;
  memBytes = (sizeof(vl_uint32) * numData + 7) & ~(vl_size)7 ;
  updBytes = memBytes + sizeof(vl_uint64) * (2 * K + 1) ;
  upd = (unsigned char*) vl_malloc(updBytes) ;
  members = (vl_uint32*) upd ;
  starts = (vl_uint64*) (upd + memBytes) ;
  restartRow = starts + K + 1 ;

  dData = vlhip_malloc(es * D * numData) ;
  dCenters = vlhip_malloc(es * D * K) ;
  dTable = vlhip_malloc(es * K * K) ;
  dMove = vlhip_malloc(es * K) ;
  dNearest = vlhip_malloc(es * K) ;
  dUB = vlhip_malloc(es * numData) ;
  dLB = vlhip_malloc(es * numData * K) ;
  dDist = vlhip_malloc(es * numData) ;
  dExact = (unsigned char*) vlhip_malloc(numData) ;
  dAssign = (vl_uint32*) vlhip_malloc(sizeof(vl_uint32) * numData) ;
  dMoved = (unsigned*) vlhip_malloc(sizeof(unsigned)) ;
  dUpd = (unsigned char*) vlhip_malloc(updBytes) ;
  vlhip_memcpy_h2d(dData, data, es * D * numData) ;
  vlhip_memcpy_h2d(dCenters, self->centers, es * D * K) ;

#define KM_CENTER_DISTANCES()                                                        \
  for (j = 0 ; j < K ; ++j) {                                                        \
    for (c = 0 ; c <= j ; ++c) {                                                     \
      TYPE z = distFn(D, (TYPE*) self->centers + c * D, (TYPE*) self->centers + j * D) ; \
      cd[c + j * K] = z ;                                                            \
      cd[j + c * K] = z ;                                                            \
    }                                                                                \
  }

  /* iteration 0: assignment to the initial centers */
  KM_CENTER_DISTANCES()
  vlhip_memcpy_h2d(dTable, cd, es * K * K) ;
  vlhip_kmeans_elkan(TYPE_CODE, 0, dData, dCenters, dTable, NULL, numData, D, K, lanes, dAssign, dUB, dExact, dLB,
                     NULL, NULL) ;
  if (self->verbosity) {
    vlhip_memcpy_d2h(dist, dUB, es * numData) ;
    for (energy = 0, x = 0 ; x < numData ; ++x) energy += dist[x] ;
    VL_PRINTF("kmeans: Elkan iteration 0, energy bound %g\n", energy) ;
  }

  for (iteration = 1 ; ; ++iteration) {
    unsigned anyMoved = 0 ;
    vl_size numRestarted = 0 ;
    /* new centers: sums in data order, divided by the mass; empty clusters
     * restart at a random datum, drawn in cluster order */
    vlhip_memcpy_d2h(assignments, dAssign, sizeof(vl_uint32) * numData) ;
    memset(clusterMasses, 0, sizeof(vl_size) * K) ;
    for (x = 0 ; x < numData ; ++x) clusterMasses[assignments[x]] ++ ;
    starts[0] = 0 ;
    for (c = 0 ; c < K ; ++c) {
      cursor[c] = starts[c] ;
      starts[c + 1] = starts[c] + clusterMasses[c] ;
    }
    for (x = 0 ; x < numData ; ++x) members[cursor[assignments[x]] ++] = (vl_uint32) x ;
    for (c = 0 ; c < K ; ++c) {
      restartRow[c] = 0 ;
      if (clusterMasses[c] == 0) {
        restartRow[c] = vl_rand_uindex(rand, numData) ;
        numRestarted ++ ;
      }
    }
    memcpy(oldCenters, self->centers, es * D * K) ;
    vlhip_memcpy_h2d(dUpd, upd, updBytes) ;
    vlhip_kmeans_update_centers(TYPE_CODE, dCenters, dData, (vl_uint32 const*) dUpd,
                                (uint64_t const*) (dUpd + memBytes), (uint64_t const*) (dUpd + memBytes) + K + 1,
                                K, D) ;
    vlhip_memcpy_d2h(self->centers, dCenters, es * D * K) ;
    /* each center's move, the new center distances and nearest-center distances */
    for (c = 0 ; c < K ; ++c) move[c] = distFn(D, (TYPE*) self->centers + c * D, oldCenters + c * D) ;
    KM_CENTER_DISTANCES()
    for (c = 0 ; c < K ; ++c) {
      nearest[c] = (TYPE) VL_INFINITY_D ;
      for (j = 0 ; j < K ; ++j) {
        if (j == c) continue ;
        nearest[c] = VL_MIN(nearest[c], cd[j + c * K]) ;
      }
    }
    vlhip_memcpy_h2d(dMove, move, es * K) ;
    vlhip_memcpy_h2d(dTable, cd, es * K * K) ;
    vlhip_memcpy_h2d(dNearest, nearest, es * K) ;
    vlhip_kmeans_elkan(TYPE_CODE, 1, dData, dCenters, NULL, dMove, numData, D, K, lanes, dAssign, dUB, dExact, dLB,
                       NULL, NULL) ;
    vlhip_memset(dMoved, 0, sizeof(unsigned)) ;
    vlhip_kmeans_elkan(TYPE_CODE, 2, dData, dCenters, dTable, dNearest, numData, D, K, lanes, dAssign, dUB, dExact,
                       dLB, dMoved, NULL) ;
    vlhip_memcpy_d2h(&anyMoved, dMoved, sizeof(unsigned)) ;
    if (self->verbosity) {
      vlhip_memcpy_d2h(dist, dUB, es * numData) ;
      for (energy = 0, x = 0 ; x < numData ; ++x) energy += dist[x] ;
      VL_PRINTF("kmeans: Elkan iteration %d, energy bound %g%s\n", (int) iteration, energy,
                numRestarted ? ", empty centers restarted" : "") ;
    }
    if (iteration >= self->maxNumIterations) break ;
    if (!anyMoved) break ;
  }
#undef KM_CENTER_DISTANCES

  vlhip_kmeans_elkan(TYPE_CODE, 3, dData, dCenters, NULL, NULL, numData, D, K, lanes, dAssign, dUB, dExact, dLB,
                     NULL, dDist) ;
  vlhip_memcpy_d2h(dist, dDist, es * numData) ;
  for (energy = 0, x = 0 ; x < numData ; ++x) energy += dist[x] ;

  {
    void * bufs [] = {dData, dCenters, dTable, dMove, dNearest, dUB, dLB, dDist, dExact, dAssign, dMoved, dUpd} ;
    for (j = 0 ; j < sizeof(bufs) / sizeof(bufs[0]) ; ++j) vlhip_free(bufs[j]) ;
  }
  vl_free(assignments) ;
  vl_free(clusterMasses) ;
  vl_free(cursor) ;
  vl_free(oldCenters) ;
  vl_free(move) ;
  vl_free(nearest) ;
  vl_free(dist) ;
  vl_free(upd) ;
  return energy ;
}

#undef FN
#undef TYPE
#undef TYPE_CODE
#undef DISTFN
