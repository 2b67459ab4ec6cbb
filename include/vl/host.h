/* host.h -- basic types and export macros of the MI355X VLFeat-compatible library.
 *
 * ABI contract: the integer typedefs below are the LP64 ones VLFeat 0.9.21 uses
 * (vl/host.h:376-394), so a client compiled against the upstream headers links
 * against this library unchanged.  Only LP64 Linux (x86_64 host of an MI355X)
 * is supported.
 */
#ifndef VL_HOST_H
#define VL_HOST_H

#include <stddef.h>

#ifdef __cplusplus
#  define VL_EXTERN_C extern "C"
#else
#  define VL_EXTERN_C extern
#endif

#define VL_EXPORT VL_EXTERN_C __attribute__((visibility("default")))
#define VL_INLINE static __inline__

#define VL_TRUE 1
#define VL_FALSE 0

typedef long long            vl_int64 ;
typedef int                  vl_int32 ;
typedef short                vl_int16 ;
typedef char                 vl_int8 ;
typedef unsigned long long   vl_uint64 ;
typedef unsigned int         vl_uint32 ;
typedef unsigned short       vl_uint16 ;
typedef unsigned char        vl_uint8 ;
typedef int                  vl_int ;
typedef unsigned int         vl_uint ;
typedef int                  vl_bool ;
typedef vl_int64             vl_intptr ;
typedef vl_uint64            vl_uintptr ;
typedef vl_uint64            vl_size ;
typedef vl_int64             vl_index ;
typedef vl_uint64            vl_uindex ;

#define VL_MIN(x,y) (((x)<(y))?(x):(y))
#define VL_MAX(x,y) (((x)>(y))?(x):(y))

#define VL_STRINGIFY(x) #x
#define VL_XSTRINGIFY(x) VL_STRINGIFY(x)
#define VL_CAT(x,y) x ## y
#define VL_XCAT(x,y) VL_CAT(x,y)

#endif /* VL_HOST_H */
