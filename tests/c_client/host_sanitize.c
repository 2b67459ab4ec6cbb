/* host_sanitize.c -- exercises the host-only parts of the library (the
 * MT19937 generator, the randomised kd-forest build and queries, the
 * homogeneous kernel map table and its scalar evaluation, the SIMD-order
 * distance functions, the error state) and prints checksums of every
 * result.  tests/test_host_sanitize.py builds it twice -- against host
 * objects compiled with AddressSanitizer and UndefinedBehaviorSanitizer,
 * and against the product library -- and requires both runs to be clean
 * and to print the same checksums.  No GPU work is started. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vl/generic.h"
#include "vl/random.h"
#include "vl/mathop.h"
#include "vl/kdtree.h"
#include "vl/homkermap.h"

static unsigned long long fnv (unsigned long long h, void const * p, size_t n)
{
  unsigned char const * b = (unsigned char const *) p ;
  size_t i ;
  for (i = 0 ; i < n ; ++i) { h ^= b[i] ; h *= 1099511628211ull ; }
  return h ;
}

int main (void)
{
  unsigned long long h = 1469598103934665603ull ;
  VlRand * r = vl_get_rand() ;
  vl_uint32 key [4] = {0x123, 0x234, 0x345, 0x456} ;
  vl_size i, t, n = 3000, D = 13, Q = 257 ;
  float * pts = malloc(sizeof(float) * n * D) ;
  float * qs = malloc(sizeof(float) * Q * D) ;
  double * dpts = malloc(sizeof(double) * n * D) ;
  vl_uint32 * idx = malloc(sizeof(vl_uint32) * Q * 3) ;
  float * dist = malloc(sizeof(float) * Q * 3) ;
  vl_index perm [97] ;

  /* generator: array and scalar seeding, permutations, 64-bit and real draws */
  vl_rand_seed_by_array(r, key, 4) ;
  for (i = 0 ; i < 2000 ; ++i) { vl_uint32 v = vl_rand_uint32(r) ; h = fnv(h, &v, sizeof(v)) ; }
  vl_rand_seed(r, 5489) ;
  for (i = 0 ; i < 97 ; ++i) perm[i] = (vl_index) i ;
  vl_rand_permute_indexes(r, perm, 97) ;
  h = fnv(h, perm, sizeof(perm)) ;
  for (i = 0 ; i < 100 ; ++i) {
    vl_uint64 a = vl_rand_uint64(r) ; double b = vl_rand_res53(r) ; vl_uindex c = vl_rand_uindex(r, 1000) ;
    h = fnv(h, &a, sizeof(a)) ; h = fnv(h, &b, sizeof(b)) ; h = fnv(h, &c, sizeof(c)) ;
  }
  printf("rand %016llx\n", h) ;

  /* kd-forest: float and double data, exact and bounded searches */
  for (i = 0 ; i < n * D ; ++i) { pts[i] = (float) (vl_rand_res53(r) - 0.5) ; dpts[i] = pts[i] ; }
  for (i = 0 ; i < 40 * D ; ++i) pts[i + 40 * D] = pts[i] ;  /* duplicate points */
  for (i = 0 ; i < Q * D ; ++i) qs[i] = (float) (vl_rand_res53(r) - 0.5) ;
  for (t = 1 ; t <= 4 ; t += 3) {
    VlKDForest * f = vl_kdforest_new(VL_TYPE_FLOAT, D, t, VlDistanceL2) ;
    VlKDForestNeighbor nb [3] ;
    vl_rand_seed(r, 17 + (vl_uint32) t) ;
    vl_kdforest_build(f, n, pts) ;
    vl_kdforest_set_max_num_comparisons(f, t == 1 ? 0 : 25) ;
    vl_set_num_threads(3) ;
    h = fnv(h, (vl_size[]){vl_kdforest_query_with_array(f, idx, 3, Q, dist, qs)}, sizeof(vl_size)) ;
    vl_set_num_threads(0) ;
    h = fnv(h, idx, sizeof(vl_uint32) * Q * 3) ;
    h = fnv(h, dist, sizeof(float) * Q * 3) ;
    for (i = 0 ; i < 20 ; ++i) {
      vl_kdforest_query(f, nb, 3, qs + i * D) ;
      h = fnv(h, nb, sizeof(nb)) ;
    }
    h = fnv(h, (vl_size[]){vl_kdforest_get_depth_of_tree(f, 0)}, sizeof(vl_size)) ;
    vl_kdforest_delete(f) ;
  }
  {
    VlKDForest * f = vl_kdforest_new(VL_TYPE_DOUBLE, D, 2, VlDistanceL2) ;
    VlKDForestSearcher * s ;
    VlKDForestNeighbor nb [5] ;
    vl_kdforest_set_thresholding_method(f, VL_KDTREE_MEAN) ;
    vl_kdforest_build(f, n, dpts) ;
    s = vl_kdforest_new_searcher(f) ;
    for (i = 0 ; i < 20 ; ++i) {
      double q [13] ;
      vl_size d ;
      for (d = 0 ; d < D ; ++d) q[d] = qs[i * D + d] ;
      vl_kdforestsearcher_query(s, nb, 5, q) ;
      h = fnv(h, nb, sizeof(nb)) ;
    }
    vl_kdforestsearcher_delete(s) ;
    vl_kdforest_delete(f) ;
  }
  printf("kdforest %016llx\n", h) ;

  /* homogeneous kernel map: three kernels, two windows, several orders */
  {
    int k, w ;
    for (k = 0 ; k < 3 ; ++k) for (w = 0 ; w < 2 ; ++w) {
      vl_size order = (vl_size) (1 + 2 * k) ;
      VlHomogeneousKernelMap * m =
        vl_homogeneouskernelmap_new((VlHomogeneousKernelType) k, 0.5 + 0.5 * w, order, -1,
                                    (VlHomogeneousKernelMapWindowType) w) ;
      double out [16] ;
      float outf [16] ;
      double x ;
      for (x = -3.0 ; x < 3.0 ; x += 0.0137) {
        vl_homogeneouskernelmap_evaluate_d(m, out, 1, x * x * x) ;
        vl_homogeneouskernelmap_evaluate_f(m, outf, 1, x) ;
        h = fnv(h, out, sizeof(double) * (2 * order + 1)) ;
        h = fnv(h, outf, sizeof(float) * (2 * order + 1)) ;
      }
      vl_homogeneouskernelmap_delete(m) ;
    }
  }
  printf("homkermap %016llx\n", h) ;

  /* distance kernels in the host's lane order */
  {
    VlFloatVectorComparisonFunction ff = vl_get_vector_comparison_function_f(VlDistanceL2) ;
    VlDoubleVectorComparisonFunction fd = vl_get_vector_comparison_function_d(VlDistanceL2) ;
    vl_size d ;
    for (d = 1 ; d <= 2 * D ; ++d) {
      float a = ff(d, pts, pts + 2 * D) ;
      double b = fd(d, dpts, dpts + 2 * D) ;
      h = fnv(h, &a, sizeof(a)) ;
      h = fnv(h, &b, sizeof(b)) ;
    }
  }
  printf("mathop %016llx\n", h) ;

  /* error state */
  vl_set_last_error(VL_ERR_BAD_ARG, "checked %d %s", 7, "values") ;
  if (vl_get_last_error() != VL_ERR_BAD_ARG || strcmp(vl_get_last_error_message(), "checked 7 values")) return 3 ;
  printf("error ok\n") ;

  free(pts) ; free(qs) ; free(dpts) ; free(idx) ; free(dist) ;
  return 0 ;
}
