/* kmeans.c -- k-means quantizer: host driver of the GPU exact quantizer.
 *
 * Drop-in for the L2 path of vl/kmeans.c (VLFeat 0.9.21):
 *   object life cycle           kmeans.c:385-482
 *   random-selection seeding    kmeans.c:527-582 (host RNG + host L2 duplicate test)
 *   k-means++ seeding           kmeans.c:588-649 (host RNG; distances on the GPU in
 *                               the host's lane order, sums on the host in data order)
 *   exact quantization          kmeans.c:655-702 (GPU, bitwise-identical distances)
 *   Lloyd refinement            kmeans.c:828-978 (GPU assignment and center update,
 *                               sums in data order; energy, masses and restart
 *                               draws on the host)
 *   ANN refinement              kmeans.c:1003-1156 (exhaustive search: as Lloyd plus
 *                               the keep-previous rule on the GPU; otherwise the
 *                               randomised kd-forest of kdtree.c on the host)
 *   Elkan refinement            kmeans.c:1162-1745 (per-datum bound updates,
 *                               reassignment and distances on the GPU, the
 *                               reference's float bounds bit for bit; center
 *                               update as Lloyd's)
 *   repetitions                 kmeans.c:2023-2093
 * Only the L2 distance is provided on this path.
 */
#include "host_fatal.h"
#include "vl/kmeans.h"
#include "vl/kdtree.h"
#include "vlhip.h"

#include <assert.h>
#include <stdio.h>
#include <string.h>

static int
host_lanes (vl_type t)
{
  int isDouble = (t == VL_TYPE_DOUBLE) ;
  if (vl_get_simd_enabled() && vl_cpu_has_avx()) return isDouble ? 4 : 8 ;
  if (vl_get_simd_enabled() && vl_cpu_has_sse2()) return isDouble ? 2 : 4 ;
  return 1 ;
}

static void
require_l2 (VlKMeans const * self, char const * who)
{
  if (self->distance != VlDistanceL2) {
    vl_host_fatal("%s: only the L2 distance is provided by this library\n", who) ;
  }
}

VL_EXPORT void
vl_kmeans_reset (VlKMeans * self)
{
  self->numCenters = 0 ;
  self->dimension = 0 ;
  if (self->centers) vl_free(self->centers) ;
  if (self->centerDistances) vl_free(self->centerDistances) ;
  self->centers = NULL ;
  self->centerDistances = NULL ;
}

VL_EXPORT VlKMeans *
vl_kmeans_new (vl_type dataType, VlVectorComparisonType distance)
{
  VlKMeans * self = vl_calloc(1, sizeof(VlKMeans)) ;
  self->algorithm = VlKMeansLloyd ;
  self->distance = distance ;
  self->dataType = dataType ;
  self->verbosity = 0 ;
  self->maxNumIterations = 100 ;
  self->minEnergyVariation = 1e-4 ;
  self->numRepetitions = 1 ;
  self->centers = NULL ;
  self->centerDistances = NULL ;
  self->numTrees = 3 ;
  self->maxNumComparisons = 100 ;
  vl_kmeans_reset(self) ;
  return self ;
}

VL_EXPORT VlKMeans *
vl_kmeans_new_copy (VlKMeans const * kmeans)
{
  VlKMeans * self = vl_calloc(1, sizeof(VlKMeans)) ;
  self->algorithm = kmeans->algorithm ;
  self->distance = kmeans->distance ;
  self->dataType = kmeans->dataType ;
  self->verbosity = kmeans->verbosity ;
  self->maxNumIterations = kmeans->maxNumIterations ;
  self->numRepetitions = kmeans->numRepetitions ;
  self->initialization = kmeans->initialization ;
  self->minEnergyVariation = kmeans->minEnergyVariation ;
  self->dimension = kmeans->dimension ;
  self->numCenters = kmeans->numCenters ;
  self->numTrees = kmeans->numTrees ;
  self->maxNumComparisons = kmeans->maxNumComparisons ;
  if (kmeans->centers) {
    vl_size n = vl_get_type_size(self->dataType) * self->dimension * self->numCenters ;
    self->centers = vl_malloc(n) ;
    memcpy(self->centers, kmeans->centers, n) ;
  }
  if (kmeans->centerDistances) {
    vl_size n = vl_get_type_size(self->dataType) * self->numCenters * self->numCenters ;
    self->centerDistances = vl_malloc(n) ;
    memcpy(self->centerDistances, kmeans->centerDistances, n) ;
  }
  return self ;
}

VL_EXPORT void
vl_kmeans_delete (VlKMeans * self)
{
  vl_kmeans_reset(self) ;
  vl_free(self) ;
}

/* Stage data (and the current centers) on the device; caller holds the lock. */
static void
stage (VlKMeans * self, void const * data, vl_size numData, void ** dData, void ** dCenters)
{
  vl_size es = vl_get_type_size(self->dataType) ;
  *dData = vlhip_buffer(VLHIP_BUF_DATA, es * self->dimension * (numData ? numData : 1)) ;
  if (numData) vlhip_memcpy_h2d(*dData, data, es * self->dimension * numData) ;
  if (dCenters) {
    *dCenters = vlhip_buffer(VLHIP_BUF_PARAMS, es * self->dimension * (self->numCenters ? self->numCenters : 1)) ;
    vlhip_memcpy_h2d(*dCenters, self->centers, es * self->dimension * self->numCenters) ;
  }
}

/* Exact assignments and distances of staged data to the current centers. */
static void
quantize_staged (VlKMeans * self, void * dData, vl_size numData,
                 vl_uint32 * assignments, void * distances)
{
  vl_size es = vl_get_type_size(self->dataType) ;
  void * dCenters = vlhip_buffer(VLHIP_BUF_PARAMS, es * self->dimension * self->numCenters) ;
  vl_uint32 * dAssign = (vl_uint32*) vlhip_buffer(VLHIP_BUF_AUX, sizeof(vl_uint32) * (numData ? numData : 1)) ;
  void * dDist = vlhip_buffer(VLHIP_BUF_AUX2, es * (numData ? numData : 1)) ;
  vlhip_memcpy_h2d(dCenters, self->centers, es * self->dimension * self->numCenters) ;
  vlhip_kmeans_quantize((int)self->dataType, dAssign, distances ? dDist : NULL, dData, dCenters,
                        numData, self->dimension, self->numCenters, host_lanes(self->dataType)) ;
  if (numData) {
    vlhip_memcpy_d2h(assignments, dAssign, sizeof(vl_uint32) * numData) ;
    if (distances) vlhip_memcpy_d2h(distances, dDist, es * numData) ;
  }
}

VL_EXPORT void
vl_kmeans_quantize (VlKMeans * self, vl_uint32 * assignments, void * distances,
                    void const * data, vl_size numData)
{
  void * dData ;
  if (self->dataType != VL_TYPE_FLOAT && self->dataType != VL_TYPE_DOUBLE) abort() ;
  require_l2(self, "vl_kmeans_quantize") ;
  vlhip_lock() ;
  stage(self, data, numData, &dData, NULL) ;
  quantize_staged(self, dData, numData, assignments, distances) ;
  vlhip_unlock() ;
}

/* ---------------------------------------------------------------- */
/*                         type-generic seeding and refinement        */

#define KM_TYPE float
#define KM_SFX f
#include "kmeans_impl.h"
#undef KM_TYPE
#undef KM_SFX

#define KM_TYPE double
#define KM_SFX d
#include "kmeans_impl.h"
#undef KM_TYPE
#undef KM_SFX

VL_EXPORT void
vl_kmeans_set_centers (VlKMeans * self, void const * centers, vl_size dimension, vl_size numCenters)
{
  vl_size n ;
  vl_kmeans_reset(self) ;
  if (self->dataType != VL_TYPE_FLOAT && self->dataType != VL_TYPE_DOUBLE) abort() ;
  n = vl_get_type_size(self->dataType) * dimension * numCenters ;
  self->dimension = dimension ;
  self->numCenters = numCenters ;
  self->centers = vl_malloc(n) ;
  memcpy(self->centers, centers, n) ;
}

VL_EXPORT void
vl_kmeans_init_centers_with_rand_data (VlKMeans * self, void const * data, vl_size dimension,
                                       vl_size numData, vl_size numCenters)
{
  vl_kmeans_reset(self) ;
  switch (self->dataType) {
    case VL_TYPE_FLOAT : km_init_rand_f(self, (float const*)data, dimension, numData, numCenters) ; break ;
    case VL_TYPE_DOUBLE : km_init_rand_d(self, (double const*)data, dimension, numData, numCenters) ; break ;
    default : abort() ;
  }
}

VL_EXPORT void
vl_kmeans_init_centers_plus_plus (VlKMeans * self, void const * data, vl_size dimension,
                                  vl_size numData, vl_size numCenters)
{
  vl_kmeans_reset(self) ;
  require_l2(self, "vl_kmeans_init_centers_plus_plus") ;
  switch (self->dataType) {
    case VL_TYPE_FLOAT : km_init_pp_f(self, (float const*)data, dimension, numData, numCenters) ; break ;
    case VL_TYPE_DOUBLE : km_init_pp_d(self, (double const*)data, dimension, numData, numCenters) ; break ;
    default : abort() ;
  }
}

VL_EXPORT void
vl_kmeans_quantize_ann (VlKMeans * self, vl_uint32 * assignments, void * distances,
                        void const * data, vl_size numData, vl_size update)
{
  require_l2(self, "vl_kmeans_quantize_ann") ;
  switch (self->dataType) {
    case VL_TYPE_FLOAT : km_quantize_ann_f(self, assignments, (float*)distances, (float const*)data, numData, update != 0) ; break ;
    case VL_TYPE_DOUBLE : km_quantize_ann_d(self, assignments, (double*)distances, (double const*)data, numData, update != 0) ; break ;
    default : abort() ;
  }
}

/* The reference header declares this spelling (kmeans.h:95) while its
 * library defines vl_kmeans_quantize_ann (kmeans.c:1947); both are exported. */
VL_EXPORT void
vl_kmeans_quantize_ANN (VlKMeans * self, vl_uint32 * assignments, void * distances,
                        void const * data, vl_size numData, vl_size iteration)
{
  vl_kmeans_quantize_ann(self, assignments, distances, data, numData, iteration) ;
}

VL_EXPORT double
vl_kmeans_refine_centers (VlKMeans * self, void const * data, vl_size numData)
{
  assert(self->centers) ;
  require_l2(self, "vl_kmeans_refine_centers") ;
  switch (self->dataType) {
    case VL_TYPE_FLOAT :
      return self->algorithm == VlKMeansElkan ? km_refine_elkan_f(self, (float const*)data, numData)
                                              : km_refine_f(self, (float const*)data, numData) ;
    case VL_TYPE_DOUBLE :
      return self->algorithm == VlKMeansElkan ? km_refine_elkan_d(self, (double const*)data, numData)
                                              : km_refine_d(self, (double const*)data, numData) ;
    default : abort() ;
  }
  return 0 ;
}

VL_EXPORT double
vl_kmeans_cluster (VlKMeans * self, void const * data, vl_size dimension,
                   vl_size numData, vl_size numCenters)
{
  vl_uindex repetition ;
  double bestEnergy = VL_INFINITY_D ;
  void * bestCenters = NULL ;

  for (repetition = 0 ; repetition < self->numRepetitions ; ++repetition) {
    double energy ;
    double timeRef ;
    if (self->verbosity) {
      VL_PRINTF("kmeans: run %d of %d\n", (int)repetition + 1, (int)self->numRepetitions) ;
    }
    timeRef = vl_get_cpu_time() ;
    switch (self->initialization) {
      case VlKMeansRandomSelection :
        vl_kmeans_init_centers_with_rand_data(self, data, dimension, numData, numCenters) ;
        break ;
      case VlKMeansPlusPlus :
        vl_kmeans_init_centers_plus_plus(self, data, dimension, numData, numCenters) ;
        break ;
      default :
        abort() ;
    }
    if (self->verbosity) {
      VL_PRINTF("kmeans: seeds chosen after %.2f s\n", vl_get_cpu_time() - timeRef) ;
    }
    timeRef = vl_get_cpu_time() ;
    energy = vl_kmeans_refine_centers(self, data, numData) ;
    if (self->verbosity) {
      VL_PRINTF("kmeans: refinement done after %.2f s, energy %g\n", vl_get_cpu_time() - timeRef, energy) ;
    }
    /* repetition 0 is kept even when its energy is NaN (kmeans.c:2070-2073) */
    if (energy < bestEnergy || repetition == 0) {
      void * temp ;
      bestEnergy = energy ;
      if (bestCenters == NULL) {
        bestCenters = vl_malloc(vl_get_type_size(self->dataType) * self->dimension * self->numCenters) ;
      }
      temp = bestCenters ;
      bestCenters = self->centers ;
      self->centers = temp ;
    }
  }
  vl_free(self->centers) ;
  self->centers = bestCenters ;
  self->energy = bestEnergy ;
  return bestEnergy ;
}
