/* vlad.h -- VLAD encoding (drop-in for VLFeat 0.9.21 vl/vlad.h).
 *
 * vl_vlad_encode replaces vl/vlad.h:45-50 (implementation vlad.c:292-315 and
 * 163-242): same arguments, flags and abort() on an unknown type.  assignments
 * is the dense numData x numClusters soft-assignment matrix the reference
 * takes; enc is a caller-allocated host buffer of dimension*numClusters.
 */
#ifndef VL_VLAD_H
#define VL_VLAD_H

#include "generic.h"

#define VL_VLAD_FLAG_NORMALIZE_COMPONENTS (0x1 << 0)
#define VL_VLAD_FLAG_SQUARE_ROOT          (0x1 << 1)
#define VL_VLAD_FLAG_UNNORMALIZED         (0x1 << 2)
#define VL_VLAD_FLAG_NORMALIZE_MASS       (0x1 << 3)

VL_EXPORT void vl_vlad_encode
(void * enc, vl_type dataType,
 void const * means, vl_size dimension, vl_size numClusters,
 void const * data, vl_size numData,
 void const * assignments,
 int flags) ;

/* Hard-assignment form (no counterpart in the reference): assignment[i] is
 * the cluster index of datum i, as produced by vl_kmeans_quantize.  The result
 * equals vl_vlad_encode with the corresponding one-hot assignment matrix
 * (vlad.c:179-189 only visits entries > 0), without materialising it. */
VL_EXPORT void vl_vlad_encode_hard
(void * enc, vl_type dataType,
 void const * means, vl_size dimension, vl_size numClusters,
 void const * data, vl_size numData,
 vl_uint32 const * assignment,
 int flags) ;

#endif /* VL_VLAD_H */
