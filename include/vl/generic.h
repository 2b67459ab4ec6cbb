/* generic.h -- runtime services of the MI355X VLFeat-compatible library.
 *
 * Replaces the subset of vl/generic.h (VLFeat 0.9.21) that the GMM, Fisher,
 * VLAD and k-means interfaces and their callers use: type codes
 * (generic.h:35-44), type sizes (84-99), allocator and printf hooks (138-156),
 * thread count (114-116), SIMD switch (104-108), the per-thread random
 * generator (110), the per-thread error state (120-132) and timers (204-206).
 */
#ifndef VL_GENERIC_H
#define VL_GENERIC_H

#include <stdlib.h>
#include "host.h"
#include "random.h"

#define VL_VERSION_STRING "0.9.21"

#define VL_TYPE_FLOAT   1
#define VL_TYPE_DOUBLE  2
#define VL_TYPE_INT8    3
#define VL_TYPE_UINT8   4
#define VL_TYPE_INT16   5
#define VL_TYPE_UINT16  6
#define VL_TYPE_INT32   7
#define VL_TYPE_UINT32  8
#define VL_TYPE_INT64   9
#define VL_TYPE_UINT64  10

typedef vl_uint32 vl_type ;

/* Size in bytes of one element of the given type; abort() on an unknown code. */
VL_INLINE vl_size
vl_get_type_size (vl_type type)
{
  switch (type) {
    case VL_TYPE_DOUBLE : return sizeof(double) ;
    case VL_TYPE_FLOAT  : return sizeof(float) ;
    case VL_TYPE_INT64  : case VL_TYPE_UINT64 : return 8 ;
    case VL_TYPE_INT32  : case VL_TYPE_UINT32 : return 4 ;
    case VL_TYPE_INT16  : case VL_TYPE_UINT16 : return 2 ;
    case VL_TYPE_INT8   : case VL_TYPE_UINT8  : return 1 ;
    default : abort() ;
  }
  return 0 ;
}

VL_EXPORT char const * vl_get_version_string (void) ;
VL_EXPORT void vl_set_simd_enabled (vl_bool x) ;
VL_EXPORT vl_bool vl_get_simd_enabled (void) ;
VL_EXPORT vl_bool vl_cpu_has_avx (void) ;
VL_EXPORT vl_bool vl_cpu_has_sse3 (void) ;
VL_EXPORT vl_bool vl_cpu_has_sse2 (void) ;
VL_EXPORT vl_size vl_get_num_cpus (void) ;
VL_EXPORT VlRand * vl_get_rand (void) ;

/* Error state of the calling thread (generic.h:120-132, generic.c:1238-1275):
 * the code and message of the last error set with vl_set_last_error. */
#define VL_ERR_MSG_LEN 1024
#define VL_ERR_OK       0
#define VL_ERR_OVERFLOW 1
#define VL_ERR_ALLOC    2
#define VL_ERR_BAD_ARG  3
#define VL_ERR_IO       4
#define VL_ERR_EOF      5
#define VL_ERR_NO_MORE  5
VL_EXPORT int vl_get_last_error (void) ;
VL_EXPORT char const * vl_get_last_error_message (void) ;
VL_EXPORT int vl_set_last_error (int error, char const * errorMessage, ...) ;

VL_EXPORT vl_size vl_get_max_threads (void) ;
VL_EXPORT void vl_set_num_threads (vl_size n) ;
VL_EXPORT vl_size vl_get_thread_limit (void) ;

VL_EXPORT void vl_set_alloc_func (void *(*malloc_func) (size_t),
                                  void *(*realloc_func) (void*,size_t),
                                  void *(*calloc_func) (size_t, size_t),
                                  void  (*free_func) (void*)) ;
VL_EXPORT void *vl_malloc (size_t n) ;
VL_EXPORT void *vl_realloc (void *ptr, size_t n) ;
VL_EXPORT void *vl_calloc (size_t n, size_t size) ;
VL_EXPORT void *vl_memalign (size_t n, size_t size) ;
VL_EXPORT void  vl_free (void* ptr) ;

typedef int (*printf_func_t) (char const *format, ...) ;
VL_EXPORT void vl_set_printf_func (printf_func_t printf_func) ;
VL_EXPORT printf_func_t vl_get_printf_func (void) ;

/* Messages go through the printf hook, which a MATLAB-style client may
 * redirect; a NULL hook silences them. */
#define VL_PRINTF (*vl_get_printf_func())
#define VL_PRINT  (*vl_get_printf_func())

VL_EXPORT void vl_tic (void) ;
VL_EXPORT double vl_toc (void) ;
VL_EXPORT double vl_get_cpu_time (void) ;

#endif /* VL_GENERIC_H */
