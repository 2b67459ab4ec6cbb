/* kmeans.h -- k-means quantizer (drop-in for VLFeat 0.9.21 vl/kmeans.h).
 *
 * VlKMeans is a public structure in the reference (kmeans.h:45-68) whose
 * fields are read by inline accessors compiled into clients, so its layout
 * is reproduced field for field.  Entry points replace kmeans.h:73-127:
 *   vl_kmeans_new/new_copy/delete/reset      kmeans.c:385-482
 *   vl_kmeans_cluster                        kmeans.c:2023-2093
 *   vl_kmeans_quantize (exact, GPU)          kmeans.c:1906-1925, 655-702
 *   vl_kmeans_quantize_ANN / _ann            kmeans.h:95 declares _ANN, kmeans.c:1947 defines _ann
 *   vl_kmeans_set_centers                    kmeans.c:1804-1830
 *   vl_kmeans_init_centers_with_rand_data    kmeans.c:1841-1870, 527-582
 *   vl_kmeans_init_centers_plus_plus         kmeans.c:1876-1900, 588-649
 *   vl_kmeans_refine_centers                 kmeans.c:1984-2011
 */
#ifndef VL_KMEANS_H
#define VL_KMEANS_H

#include "generic.h"
#include "random.h"
#include "mathop.h"

typedef enum _VlKMeansAlgorithm {
  VlKMeansLloyd,
  VlKMeansElkan,
  VlKMeansANN
} VlKMeansAlgorithm ;

typedef enum _VlKMeansInitialization {
  VlKMeansRandomSelection,
  VlKMeansPlusPlus
} VlKMeansInitialization ;

typedef struct _VlKMeans
{
  vl_type dataType ;
  vl_size dimension ;
  vl_size numCenters ;
  vl_size numTrees ;
  vl_size maxNumComparisons ;
  VlKMeansInitialization initialization ;
  VlKMeansAlgorithm algorithm ;
  VlVectorComparisonType distance ;
  vl_size maxNumIterations ;
  double minEnergyVariation ;
  vl_size numRepetitions ;
  int verbosity ;
  void * centers ;
  void * centerDistances ;
  double energy ;
  VlFloatVectorComparisonFunction floatVectorComparisonFn ;
  VlDoubleVectorComparisonFunction doubleVectorComparisonFn ;
} VlKMeans ;

VL_EXPORT VlKMeans * vl_kmeans_new (vl_type dataType, VlVectorComparisonType distance) ;
VL_EXPORT VlKMeans * vl_kmeans_new_copy (VlKMeans const * kmeans) ;
VL_EXPORT void vl_kmeans_delete (VlKMeans * self) ;
VL_EXPORT void vl_kmeans_reset (VlKMeans * self) ;

VL_EXPORT double vl_kmeans_cluster (VlKMeans * self, void const * data,
                                    vl_size dimension, vl_size numData, vl_size numCenters) ;
VL_EXPORT void vl_kmeans_quantize (VlKMeans * self, vl_uint32 * assignments, void * distances,
                                   void const * data, vl_size numData) ;
VL_EXPORT void vl_kmeans_quantize_ANN (VlKMeans * self, vl_uint32 * assignments, void * distances,
                                       void const * data, vl_size numData, vl_size iteration) ;
VL_EXPORT void vl_kmeans_quantize_ann (VlKMeans * self, vl_uint32 * assignments, void * distances,
                                       void const * data, vl_size numData, vl_size iteration) ;

VL_EXPORT void vl_kmeans_set_centers (VlKMeans * self, void const * centers,
                                      vl_size dimension, vl_size numCenters) ;
VL_EXPORT void vl_kmeans_init_centers_with_rand_data (VlKMeans * self, void const * data,
                                                      vl_size dimensions, vl_size numData, vl_size numCenters) ;
VL_EXPORT void vl_kmeans_init_centers_plus_plus (VlKMeans * self, void const * data,
                                                 vl_size dimensions, vl_size numData, vl_size numCenters) ;
VL_EXPORT double vl_kmeans_refine_centers (VlKMeans * self, void const * data, vl_size numData) ;

/* Field accessors (inline in the reference too, kmeans.h:172-431). */
VL_INLINE vl_type vl_kmeans_get_data_type (VlKMeans const * self) { return self->dataType ; }
VL_INLINE VlVectorComparisonType vl_kmeans_get_distance (VlKMeans const * self) { return self->distance ; }
VL_INLINE VlKMeansAlgorithm vl_kmeans_get_algorithm (VlKMeans const * self) { return self->algorithm ; }
VL_INLINE VlKMeansInitialization vl_kmeans_get_initialization (VlKMeans const * self) { return self->initialization ; }
VL_INLINE vl_size vl_kmeans_get_num_repetitions (VlKMeans const * self) { return self->numRepetitions ; }
VL_INLINE vl_size vl_kmeans_get_dimension (VlKMeans const * self) { return self->dimension ; }
VL_INLINE vl_size vl_kmeans_get_num_centers (VlKMeans const * self) { return self->numCenters ; }
VL_INLINE int vl_kmeans_get_verbosity (VlKMeans const * self) { return self->verbosity ; }
VL_INLINE vl_size vl_kmeans_get_max_num_iterations (VlKMeans const * self) { return self->maxNumIterations ; }
VL_INLINE double vl_kmeans_get_min_energy_variation (VlKMeans const * self) { return self->minEnergyVariation ; }
VL_INLINE vl_size vl_kmeans_get_max_num_comparisons (VlKMeans const * self) { return self->maxNumComparisons ; }
VL_INLINE vl_size vl_kmeans_get_num_trees (VlKMeans const * self) { return self->numTrees ; }
VL_INLINE double vl_kmeans_get_energy (VlKMeans const * self) { return self->energy ; }
VL_INLINE void const * vl_kmeans_get_centers (VlKMeans const * self) { return self->centers ; }

VL_INLINE void vl_kmeans_set_algorithm (VlKMeans * self, VlKMeansAlgorithm a) { self->algorithm = a ; }
VL_INLINE void vl_kmeans_set_initialization (VlKMeans * self, VlKMeansInitialization i) { self->initialization = i ; }
VL_INLINE void vl_kmeans_set_num_repetitions (VlKMeans * self, vl_size n) { /* reference asserts n >= 1 */ self->numRepetitions = n ; }
VL_INLINE void vl_kmeans_set_max_num_iterations (VlKMeans * self, vl_size n) { self->maxNumIterations = n ; }
VL_INLINE void vl_kmeans_set_min_energy_variation (VlKMeans * self, double v) { self->minEnergyVariation = v ; }
VL_INLINE void vl_kmeans_set_verbosity (VlKMeans * self, int v) { self->verbosity = v ; }
VL_INLINE void vl_kmeans_set_max_num_comparisons (VlKMeans * self, vl_size n) { self->maxNumComparisons = n ; }
VL_INLINE void vl_kmeans_set_num_trees (VlKMeans * self, vl_size n) { self->numTrees = n ; }

#endif /* VL_KMEANS_H */
