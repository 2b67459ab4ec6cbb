/* fisher_vlad.c -- host entry points of the Fisher and VLAD encoders.
 *
 * vl_fisher_encode  replaces vl/fisher.c:548-579 (algorithm 363-488)
 * vl_vlad_encode    replaces vl/vlad.c:292-315   (algorithm 163-242)
 *
 * Both stage the caller's host arrays into library-owned device buffers, run
 * the encoder kernels (vlhip_fisher_encode_batch / vlhip_vlad_encode_batch)
 * and copy the encoding back into the caller's buffer.  An unknown data type
 * aborts, as in the reference.
 */
#include "host_fatal.h"
#include "vl/fisher.h"
#include "vl/vlad.h"
#include "vlhip.h"

#include <assert.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static __thread int g_precision = VL_FISHER_PRECISION_FP32 ;

VL_EXPORT void
vl_fisher_set_precision (int precision)
{
  if (precision != VL_FISHER_PRECISION_FP32 && precision != VL_FISHER_PRECISION_BF16X3 &&
      precision != VL_FISHER_PRECISION_F16X3 && precision != VL_FISHER_PRECISION_GATED &&
      precision != VL_FISHER_PRECISION_FP32_MFMA) {
    vl_host_fatal("vl_fisher_set_precision: unknown precision %d\n", precision) ;
  }
  g_precision = precision ;
}

VL_EXPORT int vl_fisher_get_precision (void) { return g_precision ; }

VL_EXPORT void
vl_fisher_set_gate (void * gate, int candidates, float need)
{
  vlhip_fisher_set_gate(gate, candidates, need) ;
}

static void
check_type (vl_type t, char const * who)
{
  if (t != VL_TYPE_FLOAT && t != VL_TYPE_DOUBLE) {
    vl_host_fatal("%s: unsupported data type %u\n", who, (unsigned)t) ;
  }
}

static vl_size
fisher_run (void * enc, vl_type dataType, void const * means, vl_size dimension, vl_size numClusters,
            void const * covariances, void const * priors, void const * data,
            vl_uint64 const * offsets, vl_size numImages, int flags, vl_size * numTermsOut)
{
  vl_size es = vl_get_type_size(dataType) ;
  vl_size KD = dimension * numClusters ;
  vl_size numData = offsets[numImages] ;
  vl_size total = 0 ;
  unsigned char * params ;
  void * dData, * dEnc ;
  vl_uint64 * dOff ;
  unsigned long long * dTerms ;
  vl_uindex i ;

  vlhip_lock() ;
  params = (unsigned char*) vlhip_buffer(VLHIP_BUF_PARAMS, es * (2 * KD + numClusters)) ;
  dData = vlhip_buffer(VLHIP_BUF_DATA, es * dimension * (numData ? numData : 1)) ;
  dEnc = vlhip_buffer(VLHIP_BUF_ENC, es * 2 * KD * numImages) ;
  dOff = (vl_uint64*) vlhip_buffer(VLHIP_BUF_AUX, sizeof(vl_uint64) * (numImages + 1)) ;
  dTerms = (unsigned long long*) vlhip_buffer(VLHIP_BUF_AUX2, sizeof(unsigned long long) * numImages) ;
  /* stream-ordered copies; the caller's arrays stay valid for the whole call,
   * and the last download waits for everything before it */
  vlhip_memcpy_h2d_async(params, means, es * KD) ;
  vlhip_memcpy_h2d_async(params + es * KD, covariances, es * KD) ;
  vlhip_memcpy_h2d_async(params + 2 * es * KD, priors, es * numClusters) ;
  if (numData) vlhip_memcpy_h2d_async(dData, data, es * dimension * numData) ;
  vlhip_memcpy_h2d_async(dOff, offsets, sizeof(vl_uint64) * (numImages + 1)) ;
  vlhip_fisher_encode_batch((int)dataType, dEnc, params, params + es * KD, params + 2 * es * KD,
                            dimension, numClusters, dData, (uint64_t const*)dOff, numImages, flags,
                            dataType == VL_TYPE_FLOAT ? g_precision : VLHIP_MODE_FP32, dTerms) ;
  if (numTermsOut) {
    unsigned long long * h = (unsigned long long*) malloc(sizeof(unsigned long long) * numImages) ;
    vlhip_memcpy_d2h_async(enc, dEnc, es * 2 * KD * numImages) ;
    vlhip_memcpy_d2h(h, dTerms, sizeof(unsigned long long) * numImages) ;
    for (i = 0 ; i < numImages ; ++i) { numTermsOut[i] = h[i] ; total += h[i] ; }
    free(h) ;
  } else {
    vlhip_memcpy_d2h(enc, dEnc, es * 2 * KD * numImages) ;
  }
  vlhip_unlock() ;
  return total ;
}

VL_EXPORT vl_size
vl_fisher_encode (void * enc, vl_type dataType,
                  void const * means, vl_size dimension, vl_size numClusters,
                  void const * covariances, void const * priors,
                  void const * data, vl_size numData, int flags)
{
  vl_uint64 offsets [2] ;
  vl_size terms = 0 ;
  check_type(dataType, "vl_fisher_encode") ;
  assert(numClusters >= 1) ;
  assert(dimension >= 1) ;
  offsets[0] = 0 ;
  offsets[1] = numData ;
  fisher_run(enc, dataType, means, dimension, numClusters, covariances, priors, data,
             offsets, 1, flags, &terms) ;
  return terms ;
}

VL_EXPORT void
vl_fisher_encode_batch (void * enc, vl_type dataType,
                        void const * means, vl_size dimension, vl_size numClusters,
                        void const * covariances, void const * priors,
                        void const * data, vl_uint64 const * offsets, vl_size numImages,
                        int flags, vl_size * numTerms)
{
  check_type(dataType, "vl_fisher_encode_batch") ;
  assert(numClusters >= 1) ;
  assert(dimension >= 1) ;
  if (numImages == 0) return ;
  fisher_run(enc, dataType, means, dimension, numClusters, covariances, priors, data,
             offsets, numImages, flags, numTerms) ;
}

/* A dense assignment matrix whose rows each hold at most one positive entry,
 * equal to 1, is a hard assignment: only positive entries contribute
 * (vlad.c:181-188) and a weight of 1 adds the descriptor unscaled, so the
 * encoding from the cluster indices (0xffffffff for an empty row) is the same
 * number for number and the caller's N x K matrix need not be copied to the
 * GPU.  One branch-free pass per row counts the positive entries, sums their
 * indices and flags any positive entry other than 1 (vectorised; an AVX2
 * clone is picked at run time where the host has it).  Returns 0 when some
 * row is not of that form. */
#define VLAD_ONEHOT_ROW(NAME, TYPE)                                                  \
  __attribute__((optimize("O3", "tree-vectorize"), target_clones("avx2", "default"))) \
  static int NAME (TYPE const * row, vl_size numClusters, vl_uint32 * out)           \
  {                                                                                  \
    unsigned positives = 0, other = 0 ;                                              \
    vl_uint64 at = 0 ;                                                               \
    vl_uindex k ;                                                                    \
    for (k = 0 ; k < numClusters ; ++k) {                                            \
      unsigned pos = row[k] > 0 ;                                                    \
      positives += pos ;                                                             \
      at += pos ? k : 0 ;                                                            \
      other |= pos & (row[k] != 1) ;                                                 \
    }                                                                                \
    if (positives == 0) { *out = 0xffffffffu ; return 1 ; }                          \
    if (positives > 1 || other) return 0 ;                                           \
    *out = (vl_uint32) at ;                                                          \
    return 1 ;                                                                       \
  }
VLAD_ONEHOT_ROW(vlad_onehot_row_f, float)
VLAD_ONEHOT_ROW(vlad_onehot_row_d, double)

/* Rows are checked in parallel host threads (vl_get_max_threads(), as the
 * reference sizes its OpenMP loops) when the matrix is large; the outcome does
 * not depend on the thread count.  A thread stops checking after its first
 * row that is not one-hot. */
static int
vlad_onehot_indices (vl_uint32 * out, vl_type dataType, void const * soft, vl_size numData, vl_size numClusters)
{
  long long i ;
  int ok = 1 ;
  int nt = (int) vl_get_max_threads() ;
  if (numClusters >= 0xffffffffu) return 0 ;
  if (nt < 1) nt = 1 ;
#pragma omp parallel for schedule(static) reduction(&&:ok) num_threads(nt) if (numData * numClusters >= (1u << 18))
  for (i = 0 ; i < (long long) numData ; ++i) {
    if (ok) {
      ok = (dataType == VL_TYPE_FLOAT)
        ? vlad_onehot_row_f((float const *) soft + i * numClusters, numClusters, out + i)
        : vlad_onehot_row_d((double const *) soft + i * numClusters, numClusters, out + i) ;
    }
  }
  return ok ;
}

static void
vlad_run (void * enc, vl_type dataType, void const * means, vl_size dimension, vl_size numClusters,
          void const * data, vl_size numData, void const * soft, vl_uint32 const * hard, int flags)
{
  vl_size es = vl_get_type_size(dataType) ;
  vl_size KD = dimension * numClusters ;
  vl_uint64 offsets [2] ;
  void * dMeans, * dData, * dEnc, * dAssign ;
  vl_uint64 * dOff ;
  vl_uint32 * indices = NULL ;

  offsets[0] = 0 ;
  offsets[1] = numData ;
  vlhip_lock() ;
  dMeans = vlhip_buffer(VLHIP_BUF_PARAMS, es * KD) ;
  dData = vlhip_buffer(VLHIP_BUF_DATA, es * dimension * (numData ? numData : 1)) ;
  dEnc = vlhip_buffer(VLHIP_BUF_ENC, es * KD) ;
  dOff = (vl_uint64*) vlhip_buffer(VLHIP_BUF_AUX, sizeof(offsets)) ;
  /* stream-ordered copies (the caller's arrays, offsets and indices stay
   * valid until the final download, which waits for everything before it);
   * the means and data travel while the assignment matrix is examined */
  vlhip_memcpy_h2d_async(dMeans, means, es * KD) ;
  if (numData) vlhip_memcpy_h2d_async(dData, data, es * dimension * numData) ;
  vlhip_memcpy_h2d_async(dOff, offsets, sizeof(offsets)) ;
  if (soft && numData) {
    indices = (vl_uint32*) malloc(sizeof(vl_uint32) * numData) ;
    if (indices && vlad_onehot_indices(indices, dataType, soft, numData, numClusters)) {
      hard = indices ;
      soft = NULL ;
    }
  }
  dAssign = vlhip_buffer(VLHIP_BUF_POST, (soft ? es * numClusters : sizeof(vl_uint32)) * (numData ? numData : 1)) ;
  if (numData) {
    if (soft) vlhip_memcpy_h2d_async(dAssign, soft, es * numClusters * numData) ;
    else vlhip_memcpy_h2d_async(dAssign, hard, sizeof(vl_uint32) * numData) ;
  }
  vlhip_vlad_encode_batch((int)dataType, dEnc, dMeans, dimension, numClusters, dData, (uint64_t const*)dOff, 1,
                          soft ? dAssign : NULL, soft ? NULL : (uint32_t const*)dAssign, flags) ;
  vlhip_memcpy_d2h(enc, dEnc, es * KD) ;
  vlhip_unlock() ;
  free(indices) ;
}

VL_EXPORT void
vl_vlad_encode (void * enc, vl_type dataType,
                void const * means, vl_size dimension, vl_size numClusters,
                void const * data, vl_size numData,
                void const * assignments, int flags)
{
  check_type(dataType, "vl_vlad_encode") ;
  vlad_run(enc, dataType, means, dimension, numClusters, data, numData, assignments, NULL, flags) ;
}

VL_EXPORT void
vl_vlad_encode_hard (void * enc, vl_type dataType,
                     void const * means, vl_size dimension, vl_size numClusters,
                     void const * data, vl_size numData,
                     vl_uint32 const * assignment, int flags)
{
  check_type(dataType, "vl_vlad_encode_hard") ;
  vlad_run(enc, dataType, means, dimension, numClusters, data, numData, NULL, assignment, flags) ;
}
