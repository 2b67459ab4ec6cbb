/* vl_client.c -- a C caller of the VLFeat API, built against the headers in include/vl
 * and linked with -lvl: the sequence an application of the reference runs
 * (apps/recognition: GMM vocabulary, Fisher vectors, k-means words, VLAD).
 *
 *   vl_client <dir> <N> <D> <K> <n>
 *
 * reads <dir>/data.f32 (N x D float) and <dir>/init.f32 (K x D means, K x D
 * variances, K priors), then writes into <dir>:
 *   gmm.f32     means, covariances, priors after 10 EM iterations (and ll.f64)
 *   fisher.f32  improved Fisher vector of the first n descriptors
 *   assign.u32  nearest means of the first n descriptors (vl_kmeans_quantize)
 *   vlad.f32    VLAD of the first n descriptors with those one-hot assignments
 *   words.u32   nearest means by an exact kd-forest query (vl_kdforest_query_with_array)
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "vl/generic.h"
#include "vl/gmm.h"
#include "vl/fisher.h"
#include "vl/vlad.h"
#include "vl/kmeans.h"
#include "vl/kdtree.h"

static void * read_file (char const * dir, char const * name, size_t bytes)
{
  char path [4096] ;
  FILE * f ;
  void * buf = malloc(bytes) ;
  snprintf(path, sizeof(path), "%s/%s", dir, name) ;
  f = fopen(path, "rb") ;
  if (!f || fread(buf, 1, bytes, f) != bytes) { fprintf(stderr, "cannot read %s\n", path) ; exit(2) ; }
  fclose(f) ;
  return buf ;
}

static void write_file (char const * dir, char const * name, void const * buf, size_t bytes)
{
  char path [4096] ;
  FILE * f ;
  snprintf(path, sizeof(path), "%s/%s", dir, name) ;
  f = fopen(path, "wb") ;
  if (!f || fwrite(buf, 1, bytes, f) != bytes) { fprintf(stderr, "cannot write %s\n", path) ; exit(2) ; }
  fclose(f) ;
}

int main (int argc, char ** argv)
{
  char const * dir ;
  vl_size N, D, K, n, i ;
  float * data, * init, * onehot, * fisher, * vlad ;
  vl_uint32 * assign, * words ;
  VlGMM * gmm ;
  VlKMeans * km ;
  VlKDForest * forest ;
  double ll ;
  float const * means, * covs, * priors ;

  if (argc != 6) { fprintf(stderr, "usage: vl_client dir N D K n\n") ; return 2 ; }
  dir = argv[1] ;
  N = strtoull(argv[2], 0, 10) ; D = strtoull(argv[3], 0, 10) ;
  K = strtoull(argv[4], 0, 10) ; n = strtoull(argv[5], 0, 10) ;
  data = read_file(dir, "data.f32", sizeof(float) * N * D) ;
  init = read_file(dir, "init.f32", sizeof(float) * (2 * K * D + K)) ;

  /* vocabulary: a diagonal GMM by EM from a given start */
  gmm = vl_gmm_new(VL_TYPE_FLOAT, D, K) ;
  vl_gmm_set_max_num_iterations(gmm, 10) ;
  vl_gmm_set_initialization(gmm, VlGMMCustom) ;
  vl_gmm_set_means(gmm, init) ;
  vl_gmm_set_covariances(gmm, init + K * D) ;
  vl_gmm_set_priors(gmm, init + 2 * K * D) ;
  ll = vl_gmm_cluster(gmm, data, N) ;
  means = vl_gmm_get_means(gmm) ;
  covs = vl_gmm_get_covariances(gmm) ;
  priors = vl_gmm_get_priors(gmm) ;
  write_file(dir, "gmm.f32", means, sizeof(float) * K * D) ;
  write_file(dir, "gmm_covs.f32", covs, sizeof(float) * K * D) ;
  write_file(dir, "gmm_priors.f32", priors, sizeof(float) * K) ;
  write_file(dir, "ll.f64", &ll, sizeof(double)) ;

  /* Fisher vector of the first n descriptors */
  fisher = vl_malloc(sizeof(float) * 2 * K * D) ;
  vl_fisher_encode(fisher, VL_TYPE_FLOAT, means, D, K, covs, priors, data, n, VL_FISHER_FLAG_IMPROVED) ;
  write_file(dir, "fisher.f32", fisher, sizeof(float) * 2 * K * D) ;

  /* hard assignments to the means, then VLAD */
  km = vl_kmeans_new(VL_TYPE_FLOAT, VlDistanceL2) ;
  vl_kmeans_set_centers(km, means, D, K) ;
  assign = vl_malloc(sizeof(vl_uint32) * n) ;
  vl_kmeans_quantize(km, assign, NULL, data, n) ;
  write_file(dir, "assign.u32", assign, sizeof(vl_uint32) * n) ;
  onehot = vl_calloc(n * K, sizeof(float)) ;
  for (i = 0 ; i < n ; ++i) onehot[i * K + assign[i]] = 1.f ;
  vlad = vl_malloc(sizeof(float) * K * D) ;
  vl_vlad_encode(vlad, VL_TYPE_FLOAT, means, D, K, data, n, onehot, VL_VLAD_FLAG_SQUARE_ROOT) ;
  write_file(dir, "vlad.f32", vlad, sizeof(float) * K * D) ;

  /* the same words by an exhaustive kd-forest query */
  vl_rand_seed(vl_get_rand(), 7) ;
  forest = vl_kdforest_new(VL_TYPE_FLOAT, D, 1, VlDistanceL2) ;
  vl_kdforest_build(forest, K, means) ;
  vl_kdforest_set_max_num_comparisons(forest, 0) ;
  words = vl_malloc(sizeof(vl_uint32) * n) ;
  vl_kdforest_query_with_array(forest, words, 1, n, NULL, data) ;
  write_file(dir, "words.u32", words, sizeof(vl_uint32) * n) ;

  vl_kdforest_delete(forest) ;
  vl_kmeans_delete(km) ;
  vl_gmm_delete(gmm) ;
  vl_free(fisher) ; vl_free(onehot) ; vl_free(vlad) ; vl_free(assign) ; vl_free(words) ;
  free(data) ; free(init) ;
  printf("vl_client: done (LL %.6g)\n", ll) ;
  return 0 ;
}
