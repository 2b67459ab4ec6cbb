/* mathop.c -- host vector comparison functions.
 *
 * The squared L2 and Mahalanobis distances are evaluated in the reduction
 * order of the reference's SIMD kernels so that host-side consumers (k-means
 * seeding duplicate checks, k-means++ energies) see bitwise the same values:
 * LANES partial sums over the vectorised prefix, the horizontal-add tree of
 * mathop_avx.c:43-66 / mathop_sse2.c:43-70, then the scalar tail.  The lane
 * count follows the reference dispatch (mathop.c:472-550): AVX when the CPU
 * has it and SIMD is enabled (8 floats / 4 doubles), else SSE2 (4 / 2), else
 * scalar.  This file is compiled with FP contraction disabled.
 */
#include "vl/mathop.h"

#include <assert.h>
#include <stdio.h>

#define HSUM8(a) ((((a)[0] + (a)[1]) + ((a)[4] + (a)[5])) + (((a)[2] + (a)[3]) + ((a)[6] + (a)[7])))
#define HSUM4(a) (((a)[0] + (a)[2]) + ((a)[1] + (a)[3]))
#define HSUM2(a) ((a)[0] + (a)[1])

#define DEFINE_L2(NAME, T, LANES, HSUM)                                         \
  static T NAME (vl_size dimension, T const * X, T const * Y)                  \
  {                                                                            \
    T acc[LANES] ; vl_size d, l, dv = (dimension / LANES) * LANES ; T s ;      \
    for (l = 0 ; l < LANES ; ++l) acc[l] = 0 ;                                 \
    for (d = 0 ; d < dv ; d += LANES) {                                        \
      for (l = 0 ; l < LANES ; ++l) {                                          \
        T delta = X[d + l] - Y[d + l] ; T d2 = delta * delta ;                 \
        acc[l] = acc[l] + d2 ;                                                 \
      }                                                                        \
    }                                                                          \
    s = HSUM(acc) ;                                                            \
    for (d = dv ; d < dimension ; ++d) {                                       \
      T delta = X[d] - Y[d] ; T d2 = delta * delta ; s = s + d2 ;              \
    }                                                                          \
    return s ;                                                                 \
  }

#define DEFINE_MAHA(NAME, T, LANES, HSUM)                                       \
  static T NAME (vl_size dimension, T const * X, T const * MU, T const * S)    \
  {                                                                            \
    T acc[LANES] ; vl_size d, l, dv = (dimension / LANES) * LANES ; T s ;      \
    for (l = 0 ; l < LANES ; ++l) acc[l] = 0 ;                                 \
    for (d = 0 ; d < dv ; d += LANES) {                                        \
      for (l = 0 ; l < LANES ; ++l) {                                          \
        T delta = X[d + l] - MU[d + l] ; T d2 = delta * delta ;                \
        T d2s = d2 * S[d + l] ; acc[l] = acc[l] + d2s ;                        \
      }                                                                        \
    }                                                                          \
    s = HSUM(acc) ;                                                            \
    for (d = dv ; d < dimension ; ++d) {                                       \
      T delta = X[d] - MU[d] ; T d2 = delta * delta ; s = s + d2 * S[d] ;      \
    }                                                                          \
    return s ;                                                                 \
  }

DEFINE_L2(l2_f8, float, 8, HSUM8)
DEFINE_L2(l2_f4, float, 4, HSUM4)
DEFINE_L2(l2_d4, double, 4, HSUM4)
DEFINE_L2(l2_d2, double, 2, HSUM2)
DEFINE_MAHA(maha_f8, float, 8, HSUM8)
DEFINE_MAHA(maha_f4, float, 4, HSUM4)
DEFINE_MAHA(maha_d4, double, 4, HSUM4)
DEFINE_MAHA(maha_d2, double, 2, HSUM2)

static float l2_f1 (vl_size n, float const * X, float const * Y)
{ float s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) { float t = X[d] - Y[d] ; float t2 = t * t ; s = s + t2 ; } return s ; }
static double l2_d1 (vl_size n, double const * X, double const * Y)
{ double s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) { double t = X[d] - Y[d] ; double t2 = t * t ; s = s + t2 ; } return s ; }
static float maha_f1 (vl_size n, float const * X, float const * M, float const * S)
{ float s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) { float t = X[d] - M[d] ; float t2 = t * t ; s = s + t2 * S[d] ; } return s ; }
static double maha_d1 (vl_size n, double const * X, double const * M, double const * S)
{ double s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) { double t = X[d] - M[d] ; double t2 = t * t ; s = s + t2 * S[d] ; } return s ; }

/* Non-L2 comparisons are scalar (their SIMD variants in the reference only
 * change rounding); they are outside the GMM / encoder path. */
static float l1_f (vl_size n, float const * X, float const * Y)
{ float s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) s += fabsf(X[d] - Y[d]) ; return s ; }
static double l1_d (vl_size n, double const * X, double const * Y)
{ double s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) s += fabs(X[d] - Y[d]) ; return s ; }
static float chi2_f (vl_size n, float const * X, float const * Y)
{ float s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) { float a = X[d] - Y[d], b = X[d] + Y[d] ; if (b) s += a * a / b ; } return s ; }
static double chi2_d (vl_size n, double const * X, double const * Y)
{ double s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) { double a = X[d] - Y[d], b = X[d] + Y[d] ; if (b) s += a * a / b ; } return s ; }
static float kl2_f (vl_size n, float const * X, float const * Y)
{ float s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) s += X[d] * Y[d] ; return s ; }
static double kl2_d (vl_size n, double const * X, double const * Y)
{ double s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) s += X[d] * Y[d] ; return s ; }
static float kl1_f (vl_size n, float const * X, float const * Y)
{ float s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) { float a = fabsf(X[d]), b = fabsf(Y[d]) ; s += a + b - fabsf(X[d] - Y[d]) ; } return s / 2 ; }
static double kl1_d (vl_size n, double const * X, double const * Y)
{ double s = 0 ; vl_size d ; for (d = 0 ; d < n ; ++d) { double a = fabs(X[d]), b = fabs(Y[d]) ; s += a + b - fabs(X[d] - Y[d]) ; } return s / 2 ; }

static int
host_lanes (int isDouble)
{
  if (vl_get_simd_enabled() && vl_cpu_has_avx()) return isDouble ? 4 : 8 ;
  if (vl_get_simd_enabled() && vl_cpu_has_sse2()) return isDouble ? 2 : 4 ;
  return 1 ;
}

VL_EXPORT VlFloatVectorComparisonFunction
vl_get_vector_comparison_function_f (VlVectorComparisonType type)
{
  switch (type) {
    case VlDistanceL2 : {
      int l = host_lanes(0) ;
      return l == 8 ? l2_f8 : (l == 4 ? l2_f4 : l2_f1) ;
    }
    case VlDistanceL1 : return l1_f ;
    case VlDistanceChi2 : return chi2_f ;
    case VlKernelL2 : return kl2_f ;
    case VlKernelL1 : return kl1_f ;
    default :
      VL_PRINTF("vl_get_vector_comparison_function_f: comparison %d is not provided by this library\n", (int)type) ;
      abort() ;
  }
  return 0 ;
}

VL_EXPORT VlDoubleVectorComparisonFunction
vl_get_vector_comparison_function_d (VlVectorComparisonType type)
{
  switch (type) {
    case VlDistanceL2 : {
      int l = host_lanes(1) ;
      return l == 4 ? l2_d4 : (l == 2 ? l2_d2 : l2_d1) ;
    }
    case VlDistanceL1 : return l1_d ;
    case VlDistanceChi2 : return chi2_d ;
    case VlKernelL2 : return kl2_d ;
    case VlKernelL1 : return kl1_d ;
    default :
      VL_PRINTF("vl_get_vector_comparison_function_d: comparison %d is not provided by this library\n", (int)type) ;
      abort() ;
  }
  return 0 ;
}

VL_EXPORT VlFloatVector3ComparisonFunction
vl_get_vector_3_comparison_function_f (VlVectorComparisonType type)
{
  int l ;
  if (type != VlDistanceMahalanobis) abort() ;
  l = host_lanes(0) ;
  return l == 8 ? maha_f8 : (l == 4 ? maha_f4 : maha_f1) ;
}

VL_EXPORT VlDoubleVector3ComparisonFunction
vl_get_vector_3_comparison_function_d (VlVectorComparisonType type)
{
  int l ;
  if (type != VlDistanceMahalanobis) abort() ;
  l = host_lanes(1) ;
  return l == 4 ? maha_d4 : (l == 2 ? maha_d2 : maha_d1) ;
}

/* result[yi * numDataX + xi] = f(X_xi, Y_yi); with Y == NULL the symmetric
 * matrix of X against itself (mathop.c:554-596). */
#define DEFINE_ALL_PAIRS(NAME, T, FN)                                           \
  VL_EXPORT void NAME (T * result, vl_size dimension, T const * X, vl_size numDataX, \
                       T const * Y, vl_size numDataY, FN function)             \
  {                                                                            \
    vl_uindex xi, yi ;                                                         \
    if (dimension == 0 || numDataX == 0) return ;                              \
    assert(X) ;                                                                \
    if (Y) {                                                                   \
      if (numDataY == 0) return ;                                              \
      for (yi = 0 ; yi < numDataY ; ++yi)                                      \
        for (xi = 0 ; xi < numDataX ; ++xi)                                    \
          *result++ = function(dimension, X + xi * dimension, Y + yi * dimension) ; \
    } else {                                                                   \
      for (yi = 0 ; yi < numDataX ; ++yi)                                      \
        for (xi = 0 ; xi <= yi ; ++xi) {                                       \
          T z = function(dimension, X + xi * dimension, X + yi * dimension) ;  \
          result[yi * numDataX + xi] = z ;                                     \
          result[xi * numDataX + yi] = z ;                                     \
        }                                                                      \
    }                                                                          \
  }

DEFINE_ALL_PAIRS(vl_eval_vector_comparison_on_all_pairs_f, float, VlFloatVectorComparisonFunction)
DEFINE_ALL_PAIRS(vl_eval_vector_comparison_on_all_pairs_d, double, VlDoubleVectorComparisonFunction)
