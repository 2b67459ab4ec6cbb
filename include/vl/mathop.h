/* mathop.h -- constants and vector comparison functions used by the GMM,
 * k-means and encoder interfaces.
 *
 * Replaces the part of vl/mathop.h (VLFeat 0.9.21) those interfaces need:
 * VL_PI / epsilons (mathop.h:28-45), the comparison-type enum (627-643), the
 * comparison-function getters (665-676) and the all-pairs evaluator (678-688).
 * The host-side functions here are plain C; the device kernels use the same
 * accumulation order for L2 where bit-exact argmin parity is required.
 */
#ifndef VL_MATHOP_H
#define VL_MATHOP_H

#include <math.h>
#include "generic.h"

#define VL_PI 3.141592653589793
#define VL_EPSILON_F 1.19209290E-07F
#define VL_EPSILON_D 2.220446049250313e-16
#define VL_INFINITY_D ((double)INFINITY)
#define VL_INFINITY_F ((float)INFINITY)
#define VL_NAN_D ((double)NAN)
#define VL_NAN_F ((float)NAN)

typedef float (*VlFloatVectorComparisonFunction)(vl_size dimension, float const * X, float const * Y) ;
typedef double (*VlDoubleVectorComparisonFunction)(vl_size dimension, double const * X, double const * Y) ;
typedef float (*VlFloatVector3ComparisonFunction)(vl_size dimension, float const * X, float const * Y, float const * Z) ;
typedef double (*VlDoubleVector3ComparisonFunction)(vl_size dimension, double const * X, double const * Y, double const * Z) ;

enum _VlVectorComparisonType {
  VlDistanceL1,
  VlDistanceL2,
  VlDistanceChi2,
  VlDistanceHellinger,
  VlDistanceJS,
  VlDistanceMahalanobis,
  VlKernelL1,
  VlKernelL2,
  VlKernelChi2,
  VlKernelHellinger,
  VlKernelJS
} ;
typedef enum _VlVectorComparisonType VlVectorComparisonType ;

VL_INLINE char const *
vl_get_vector_comparison_type_name (int type)
{
  switch (type) {
    case VlDistanceL1 : return "l1" ;
    case VlDistanceL2 : return "l2" ;
    case VlDistanceChi2 : return "chi2" ;
    case VlDistanceMahalanobis : return "mahalanobis" ;
    case VlKernelL1 : return "kl1" ;
    case VlKernelL2 : return "kl2" ;
    case VlKernelChi2 : return "kchi2" ;
    default : return NULL ;
  }
}

VL_INLINE float vl_abs_f (float x) { return fabsf(x) ; }
VL_INLINE double vl_abs_d (double x) { return fabs(x) ; }
VL_INLINE float vl_sqrt_f (float x) { return sqrtf(x) ; }
VL_INLINE double vl_sqrt_d (double x) { return sqrt(x) ; }

VL_EXPORT VlFloatVectorComparisonFunction vl_get_vector_comparison_function_f (VlVectorComparisonType type) ;
VL_EXPORT VlDoubleVectorComparisonFunction vl_get_vector_comparison_function_d (VlVectorComparisonType type) ;
VL_EXPORT VlFloatVector3ComparisonFunction vl_get_vector_3_comparison_function_f (VlVectorComparisonType type) ;
VL_EXPORT VlDoubleVector3ComparisonFunction vl_get_vector_3_comparison_function_d (VlVectorComparisonType type) ;

VL_EXPORT void
vl_eval_vector_comparison_on_all_pairs_f (float * result, vl_size dimension,
                                          float const * X, vl_size numDataX,
                                          float const * Y, vl_size numDataY,
                                          VlFloatVectorComparisonFunction function) ;
VL_EXPORT void
vl_eval_vector_comparison_on_all_pairs_d (double * result, vl_size dimension,
                                          double const * X, vl_size numDataX,
                                          double const * Y, vl_size numDataY,
                                          VlDoubleVectorComparisonFunction function) ;

#endif /* VL_MATHOP_H */
