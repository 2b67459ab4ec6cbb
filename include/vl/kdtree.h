/* kdtree.h -- randomised kd-forest used by the approximate k-means
 * refinement (VLFeat 0.9.21 vl/kdtree.h).
 *
 * Replaces vl_kdforest_new / build / query / query_with_array / delete and
 * the parameter getters (kdtree.c:332-380, 530-580, 774-1075; kdtree.h:
 * 136-180) for float and double data and the L2 distance.  The build
 * draws from the calling thread's generator exactly as the reference does
 * (variance sampling and split-dimension choice, kdtree.c:206 and 260), so
 * k-means runs that interleave tree builds with other random draws stay in
 * step with the reference.  The forest object is opaque.
 */
#ifndef VL_KDTREE_H
#define VL_KDTREE_H

#include "generic.h"
#include "mathop.h"

#define VL_KDTREE_SPLIT_HEAP_SIZE 5
#define VL_KDTREE_VARIANCE_EST_NUM_SAMPLES 1024

typedef enum _VlKDTreeThresholdingMethod
{
  VL_KDTREE_MEDIAN,
  VL_KDTREE_MEAN
} VlKDTreeThresholdingMethod ;

typedef struct _VlKDForestNeighbor {
  double distance ;
  vl_uindex index ;
} VlKDForestNeighbor ;

typedef struct _VlKDForest VlKDForest ;
typedef struct _VlKDForestSearcher VlKDForestSearcher ;

VL_EXPORT VlKDForest * vl_kdforest_new (vl_type dataType, vl_size dimension, vl_size numTrees,
                                        VlVectorComparisonType normType) ;
VL_EXPORT void vl_kdforest_delete (VlKDForest * self) ;
VL_EXPORT void vl_kdforest_build (VlKDForest * self, vl_size numData, void const * data) ;
VL_EXPORT void vl_kdforest_set_max_num_comparisons (VlKDForest * self, vl_size n) ;
VL_EXPORT vl_size vl_kdforest_get_max_num_comparisons (VlKDForest * self) ;
VL_EXPORT void vl_kdforest_set_thresholding_method (VlKDForest * self, VlKDTreeThresholdingMethod method) ;
VL_EXPORT vl_size vl_kdforest_get_num_nodes_of_tree (VlKDForest const * self, vl_uindex treeIndex) ;
VL_EXPORT vl_size vl_kdforest_get_depth_of_tree (VlKDForest const * self, vl_uindex treeIndex) ;
VL_EXPORT VlKDForestSearcher * vl_kdforest_new_searcher (VlKDForest * kdforest) ;
VL_EXPORT void vl_kdforestsearcher_delete (VlKDForestSearcher * searcher) ;
VL_EXPORT vl_size vl_kdforestsearcher_query (VlKDForestSearcher * self, VlKDForestNeighbor * neighbors,
                                             vl_size numNeighbors, void const * query) ;
VL_EXPORT vl_size vl_kdforest_query (VlKDForest * self, VlKDForestNeighbor * neighbors,
                                     vl_size numNeighbors, void const * query) ;
VL_EXPORT vl_size vl_kdforest_query_with_array (VlKDForest * self, vl_uint32 * index, vl_size numNeighbors,
                                                vl_size numQueries, void * distance, void const * queries) ;
VL_EXPORT vl_size vl_kdforest_get_num_trees (VlKDForest const * self) ;
VL_EXPORT vl_size vl_kdforest_get_data_dimension (VlKDForest const * self) ;
VL_EXPORT vl_type vl_kdforest_get_data_type (VlKDForest const * self) ;
VL_EXPORT VlKDTreeThresholdingMethod vl_kdforest_get_thresholding_method (VlKDForest const * self) ;
VL_EXPORT VlKDForest * vl_kdforest_searcher_get_forest (VlKDForestSearcher const * self) ;
VL_EXPORT VlKDForestSearcher * vl_kdforest_get_searcher (VlKDForest const * self, vl_uindex pos) ;

#endif /* VL_KDTREE_H */
