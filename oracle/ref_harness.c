/* ref_harness.c -- array loops over the reference's scalar entry points, for
 * timing and checking the reference as its MATLAB front-ends drive it (test
 * infrastructure, linked only into _ref/libvlref.so).
 *
 * ref_homkermap_array_{f,d}: one vl_homogeneouskernelmap_evaluate call per
 * value, outputs of value i at out[i*(2*order+1)], the loop of
 * toolbox/misc/vl_homkermap.c:170-184. */
#include <stddef.h>

typedef struct _VlHomogeneousKernelMap VlHomogeneousKernelMap ;
void vl_homogeneouskernelmap_evaluate_f (VlHomogeneousKernelMap const *, float *, unsigned long long, double) ;
void vl_homogeneouskernelmap_evaluate_d (VlHomogeneousKernelMap const *, double *, unsigned long long, double) ;
unsigned long long vl_homogeneouskernelmap_get_dimension (VlHomogeneousKernelMap const *) ;

__attribute__((visibility("default"))) void
ref_homkermap_array_f (VlHomogeneousKernelMap const * map, float * out, float const * x, size_t n)
{
  size_t const F = (size_t) vl_homogeneouskernelmap_get_dimension(map) ;
  size_t i ;
  for (i = 0 ; i < n ; ++i) vl_homogeneouskernelmap_evaluate_f(map, out + i * F, 1, x[i]) ;
}

__attribute__((visibility("default"))) void
ref_homkermap_array_d (VlHomogeneousKernelMap const * map, double * out, double const * x, size_t n)
{
  size_t const F = (size_t) vl_homogeneouskernelmap_get_dimension(map) ;
  size_t i ;
  for (i = 0 ; i < n ; ++i) vl_homogeneouskernelmap_evaluate_d(map, out + i * F, 1, x[i]) ;
}
