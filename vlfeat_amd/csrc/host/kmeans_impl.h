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

/* Lloyd (kmeans.c:828-978) and ANN (kmeans.c:1003-1156) refinement: the
 * two differ only in how assignments are obtained. */
static double
FN(km_refine) (VlKMeans * self, TYPE const * data, vl_size numData)
{
  vl_size c, d, x, iteration ;
  double previousEnergy = VL_INFINITY_D ;
  double initialEnergy = VL_INFINITY_D ;
  double energy ;
  vl_bool ann = (self->algorithm == VlKMeansANN) ;
  TYPE * distances = (TYPE*) vl_malloc(sizeof(TYPE) * (numData ? numData : 1)) ;
  vl_uint32 * assignments = (vl_uint32*) vl_malloc(sizeof(vl_uint32) * (numData ? numData : 1)) ;
  vl_size * clusterMasses = (vl_size*) vl_malloc(sizeof(vl_size) * (numData > self->numCenters ? numData : self->numCenters)) ;
  VlRand * rand = vl_get_rand() ;
  void * dData = NULL ;

  if (!ann) {
    vlhip_lock() ;
    dData = vlhip_buffer(VLHIP_BUF_DATA, sizeof(TYPE) * self->dimension * (numData ? numData : 1)) ;
    if (numData) vlhip_memcpy_h2d(dData, data, sizeof(TYPE) * self->dimension * numData) ;
  }

  for (energy = VL_INFINITY_D, iteration = 0 ; 1 ; ++iteration) {
    vl_size numRestarted = 0 ;
    if (ann) {
      FN(km_quantize_ann)(self, assignments, distances, data, numData, iteration > 0) ;
    } else {
      quantize_staged(self, dData, numData, assignments, distances) ;
    }
    energy = 0 ;
    for (x = 0 ; x < numData ; ++x) energy += distances[x] ;
    if (self->verbosity) {
      VL_PRINTF("kmeans: %s iter %d: energy = %g\n", ann ? "ANN" : "Lloyd", (int)iteration, energy) ;
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

    memset(clusterMasses, 0, sizeof(vl_size) * (numData > self->numCenters ? numData : self->numCenters)) ;
    for (x = 0 ; x < numData ; ++x) clusterMasses[assignments[x]] ++ ;
    memset(self->centers, 0, sizeof(TYPE) * self->dimension * self->numCenters) ;
    for (x = 0 ; x < numData ; ++x) {
      TYPE * cpt = (TYPE*)self->centers + assignments[x] * self->dimension ;
      TYPE const * xpt = data + x * self->dimension ;
      for (d = 0 ; d < self->dimension ; ++d) cpt[d] += xpt[d] ;
    }
    for (c = 0 ; c < self->numCenters ; ++c) {
      TYPE * cpt = (TYPE*)self->centers + c * self->dimension ;
      if (clusterMasses[c] > 0) {
        TYPE mass = (TYPE) clusterMasses[c] ;
        for (d = 0 ; d < self->dimension ; ++d) cpt[d] /= mass ;
      } else {
        vl_uindex r = vl_rand_uindex(rand, numData) ;
        numRestarted ++ ;
        for (d = 0 ; d < self->dimension ; ++d) cpt[d] = data[r * self->dimension + d] ;
      }
    }
    if (self->verbosity && numRestarted) {
      VL_PRINTF("kmeans: iter %d: restarted %d centers\n", (int)iteration, (int)numRestarted) ;
    }
  }
  if (!ann) vlhip_unlock() ;
  vl_free(distances) ;
  vl_free(assignments) ;
  vl_free(clusterMasses) ;
  return energy ;
}

#undef FN
#undef TYPE
#undef TYPE_CODE
#undef DISTFN
