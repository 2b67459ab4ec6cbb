/* random.h -- MT19937 generator with the VLFeat 0.9.21 interface.
 *
 * Replaces vl/random.h:21-159.  The state layout (624 words + index) is part of
 * the ABI because the inline helpers below read it through the exported
 * vl_rand_uint32 only; it is kept identical so a VlRand embedded in a client
 * structure has the upstream size.
 */
#ifndef VL_RANDOM_H
#define VL_RANDOM_H

#include "host.h"

typedef struct _VlRand {
  vl_uint32 mt [624] ;
  vl_uint32 mti ;
} VlRand ;

VL_EXPORT void vl_rand_init (VlRand * self) ;
VL_EXPORT void vl_rand_seed (VlRand * self, vl_uint32 s) ;
VL_EXPORT void vl_rand_seed_by_array (VlRand * self, vl_uint32 const key [], vl_size keySize) ;
VL_EXPORT vl_uint32 vl_rand_uint32 (VlRand * self) ;
VL_EXPORT void vl_rand_permute_indexes (VlRand * self, vl_index * array, vl_size size) ;

/* 64 random bits: high word drawn first (vl/random.h:84-90). */
VL_INLINE vl_uint64
vl_rand_uint64 (VlRand * self)
{
  vl_uint64 hi = vl_rand_uint32(self) ;
  vl_uint64 lo = vl_rand_uint32(self) ;
  return (hi << 32) | lo ;
}

VL_INLINE vl_int64  vl_rand_int63 (VlRand * self) { return (vl_int64)(vl_rand_uint64(self) >> 1) ; }
VL_INLINE vl_int32  vl_rand_int31 (VlRand * self) { return (vl_int32)(vl_rand_uint32(self) >> 1) ; }

/* [0,1] with 32-bit resolution (divide by 2^32-1). */
VL_INLINE double vl_rand_real1 (VlRand * self) { return vl_rand_uint32(self) * (1.0/4294967295.0) ; }
/* [0,1) (divide by 2^32). */
VL_INLINE double vl_rand_real2 (VlRand * self) { return vl_rand_uint32(self) * (1.0/4294967296.0) ; }
/* (0,1): offset by one half before dividing by 2^32. */
VL_INLINE double vl_rand_real3 (VlRand * self) { return (((double)vl_rand_uint32(self)) + 0.5) * (1.0/4294967296.0) ; }
/* [0,1) with 53-bit resolution. */
VL_INLINE double
vl_rand_res53 (VlRand * self)
{
  vl_uint32 a = vl_rand_uint32(self) >> 5, b = vl_rand_uint32(self) >> 6 ;
  return (a * 67108864.0 + b) * (1.0/9007199254740992.0) ;
}

/* Index in [0, range-1] by modulo reduction (biased for non powers of two,
 * exactly as vl/random.h:65-75). */
VL_INLINE vl_uindex
vl_rand_uindex (VlRand * self, vl_uindex range)
{
  if (range <= 0xffffffff) {
    return (vl_rand_uint32(self) % (vl_uint32)range) ;
  }
  return (vl_rand_uint64(self) % range) ;
}

#endif /* VL_RANDOM_H */
