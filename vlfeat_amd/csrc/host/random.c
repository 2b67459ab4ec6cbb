/* random.c -- Mersenne Twister MT19937 (Matsumoto & Nishimura 1998, the
 * published mt19937ar algorithm) behind the VLFeat 0.9.21 VlRand interface
 * (vl/random.c:132-260, vl/random.h:21-159).
 *
 * Seeding: vl_rand_seed is the standard init_genrand (Knuth multiplier
 * 1812433253), vl_rand_seed_by_array the standard init_by_array from seed
 * 19650218.  An unseeded generator (after vl_rand_init) seeds itself with
 * 5489 on its first draw, as mt19937ar does.  vl_rand_permute_indexes is the
 * descending Fisher-Yates shuffle of vl/random.c:214-223.
 */
#include "vl/random.h"

#include <string.h>

enum { MT_N = 624, MT_M = 397 } ;
#define MT_MATRIX_A 0x9908b0dfU
#define MT_UPPER 0x80000000U
#define MT_LOWER 0x7fffffffU

VL_EXPORT void
vl_rand_init (VlRand * self)
{
  memset(self->mt, 0, sizeof(self->mt)) ;
  self->mti = MT_N + 1 ;
}

VL_EXPORT void
vl_rand_seed (VlRand * self, vl_uint32 s)
{
  vl_uint32 i ;
  self->mt[0] = s ;
  for (i = 1 ; i < MT_N ; ++i) {
    vl_uint32 prev = self->mt[i - 1] ;
    self->mt[i] = 1812433253U * (prev ^ (prev >> 30)) + i ;
  }
  self->mti = MT_N ;
}

VL_EXPORT void
vl_rand_seed_by_array (VlRand * self, vl_uint32 const key [], vl_size keySize)
{
  vl_uint32 i = 1, j = 0 ;
  vl_size k ;
  vl_rand_seed(self, 19650218U) ;
  for (k = (MT_N > keySize ? MT_N : keySize) ; k > 0 ; --k) {
    vl_uint32 prev = self->mt[i - 1] ;
    self->mt[i] = (self->mt[i] ^ ((prev ^ (prev >> 30)) * 1664525U)) + key[j] + j ;
    ++i ; ++j ;
    if (i >= MT_N) { self->mt[0] = self->mt[MT_N - 1] ; i = 1 ; }
    if (j >= keySize) j = 0 ;
  }
  for (k = MT_N - 1 ; k > 0 ; --k) {
    vl_uint32 prev = self->mt[i - 1] ;
    self->mt[i] = (self->mt[i] ^ ((prev ^ (prev >> 30)) * 1566083941U)) - i ;
    ++i ;
    if (i >= MT_N) { self->mt[0] = self->mt[MT_N - 1] ; i = 1 ; }
  }
  self->mt[0] = 0x80000000U ;
}

static void
mt_regenerate (VlRand * self)
{
  vl_uint32 * mt = self->mt ;
  int kk ;
  for (kk = 0 ; kk < MT_N ; ++kk) {
    vl_uint32 y = (mt[kk] & MT_UPPER) | (mt[(kk + 1) % MT_N] & MT_LOWER) ;
    mt[kk] = mt[(kk + MT_M) % MT_N] ^ (y >> 1) ^ ((y & 1U) ? MT_MATRIX_A : 0U) ;
  }
  self->mti = 0 ;
}

VL_EXPORT vl_uint32
vl_rand_uint32 (VlRand * self)
{
  vl_uint32 y ;
  if (self->mti >= MT_N) {
    if (self->mti == MT_N + 1) vl_rand_seed(self, 5489U) ;
    mt_regenerate(self) ;
  }
  y = self->mt[self->mti++] ;
  y ^= (y >> 11) ;
  y ^= (y << 7) & 0x9d2c5680U ;
  y ^= (y << 15) & 0xefc60000U ;
  y ^= (y >> 18) ;
  return y ;
}

VL_EXPORT void
vl_rand_permute_indexes (VlRand * self, vl_index * array, vl_size size)
{
  vl_index i ;
  for (i = (vl_index)size - 1 ; i > 0 ; --i) {
    vl_index j = (vl_int) vl_rand_uindex(self, (vl_uindex)i + 1) ;
    vl_index tmp = array[i] ; array[i] = array[j] ; array[j] = tmp ;
  }
}
