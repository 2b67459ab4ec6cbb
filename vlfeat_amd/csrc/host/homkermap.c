/* homkermap.c -- homogeneous kernel map object (VLFeat 0.9.21 vl/homkermap.c).
 *
 * vl_homogeneouskernelmap_new builds the map's interpolation table on the
 * host in double precision, as the reference does at construction
 * (homkermap.c:326-426): the kernel's spectrum (239-255), optionally smoothed
 * by a rectangular window (270-297), sampled at the first order+1 useful
 * frequencies, and the feature values of x = mantissa * 2^exponent for
 * numSubdivisions mantissas in [1, 2) and exponents -20 .. 8.  Evaluation
 * interpolates linearly between two table rows (homkermap.c:532-571):
 * one value on the host (the reference's scalar entry points), whole arrays
 * on the GPU through vlhip_homkermap_evaluate.  The same double operations in
 * the same order are used everywhere, so every path returns the reference's
 * values bit for bit.
 */
#include "vl/homkermap.h"
#include "vlhip.h"
#include "svm_private.h"

#include <assert.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define HKM_PI 3.141592653589793

/* sin and cos are called one by one, as in the reference's own build: a
 * compiler that merges a sin and a cos of the same argument into one sincos
 * call changes the last bit of some table entries */
static double (*volatile hkm_sin) (double) = sin ;
static double (*volatile hkm_cos) (double) = cos ;

struct _VlHomogeneousKernelMap {
  VlHomogeneousKernelType kernelType ;
  VlHomogeneousKernelMapWindowType windowType ;
  double gamma ;
  vl_size order ;
  double period ;
  vl_size numSubdivisions ;
  double subdivision ;
  vl_index minExponent ;
  vl_index maxExponent ;
  double * table ;        /* rows of 2*order+1 values, then kappa and freq */
  void * deviceTable ;    /* device copy of the rows, made on first array use */
} ;

/* Spectrum kappa(omega) of the kernel (homkermap.c:239-255). */
static double
hkm_spectrum (VlHomogeneousKernelType type, double omega)
{
  switch (type) {
    case VlHomogeneousKernelIntersection:
      return (2.0 / HKM_PI) / (1 + 4 * omega * omega) ;
    case VlHomogeneousKernelChi2:
      return 2.0 / (exp(HKM_PI * omega) + exp(-HKM_PI * omega)) ;
    case VlHomogeneousKernelJS:
      return (2.0 / log(4.0)) * 2.0 / (exp(HKM_PI * omega) + exp(-HKM_PI * omega)) / (1 + 4 * omega * omega) ;
    default:
      abort() ;
  }
}

static double
hkm_sinc (double x)
{
  return (x == 0.0) ? 1.0 : hkm_sin(x) / x ;
}

/* Spectrum seen through the window (homkermap.c:270-297): the rectangular
 * window of width `period` convolves it with a sinc, summed over 2049
 * frequencies by the running abscissa the reference uses, then clipped at 0. */
static double
hkm_smooth_spectrum (VlHomogeneousKernelMap const * self, double omega)
{
  if (self->windowType == VlHomogeneousKernelMapWindowUniform) {
    return hkm_spectrum(self->kernelType, omega) ;
  } else if (self->windowType == VlHomogeneousKernelMapWindowRectangular) {
    double const range = 2.0 / (self->period * 1e-2) ;
    double const step = 2 * range / (2 * 1024.0 + 1) ;
    double acc = 0 ;
    double w ;
    for (w = -range ; w <= range ; w += step) {
      double win = hkm_sinc((self->period / 2.0) * w) ;
      win *= (self->period / (2.0 * HKM_PI)) ;
      acc += win * hkm_spectrum(self->kernelType, w + omega) ;
    }
    acc *= step ;
    return (acc > 0.0) ? acc : 0.0 ;
  }
  abort() ;
}

/* Default period of the reference (homkermap.c:345-363): empirical fits per
 * kernel and window, at least 1. */
static double
hkm_default_period (VlHomogeneousKernelType type, VlHomogeneousKernelMapWindowType window, vl_size order)
{
  double p = 1.0 ;
  if (window == VlHomogeneousKernelMapWindowUniform) {
    switch (type) {
      case VlHomogeneousKernelChi2:         p = 5.86 * sqrt(order + 0) + 3.65 ; break ;
      case VlHomogeneousKernelJS:           p = 6.64 * sqrt(order + 0) + 7.24 ; break ;
      case VlHomogeneousKernelIntersection: p = 2.38 * log(order + 0.8) + 5.6 ; break ;
    }
  } else {
    switch (type) {
      case VlHomogeneousKernelChi2:         p = 8.80 * sqrt(order + 4.44) - 12.6 ; break ;
      case VlHomogeneousKernelJS:           p = 9.63 * sqrt(order + 1.00) - 2.93 ; break ;
      case VlHomogeneousKernelIntersection: p = 2.00 * log(order + 0.99) + 3.52 ; break ;
    }
  }
  return (p > 1.0) ? p : 1.0 ;
}

VL_EXPORT VlHomogeneousKernelMap *
vl_homogeneouskernelmap_new (VlHomogeneousKernelType kernelType, double gamma, vl_size order, double period,
                             VlHomogeneousKernelMapWindowType windowType)
{
  VlHomogeneousKernelMap * self ;
  vl_size rowLength, numRows ;
  double * kappa, * freq, * row ;
  double L, mantissa ;
  vl_index exponent ;
  vl_uindex i, j, s ;

  assert(gamma > 0) ;
  assert(kernelType == VlHomogeneousKernelIntersection || kernelType == VlHomogeneousKernelChi2 ||
         kernelType == VlHomogeneousKernelJS) ;
  assert(windowType == VlHomogeneousKernelMapWindowUniform || windowType == VlHomogeneousKernelMapWindowRectangular) ;
  self = (VlHomogeneousKernelMap*) vl_malloc(sizeof(VlHomogeneousKernelMap)) ;
  if (!self) return NULL ;
  if (period < 0) period = hkm_default_period(kernelType, windowType, order) ;

  self->kernelType = kernelType ;
  self->windowType = windowType ;
  self->gamma = gamma ;
  self->order = order ;
  self->period = period ;
  self->numSubdivisions = 8 + 8 * order ;
  self->subdivision = 1.0 / self->numSubdivisions ;
  self->minExponent = -20 ;
  self->maxExponent = 8 ;
  self->deviceTable = NULL ;

  rowLength = 2 * order + 1 ;
  numRows = self->numSubdivisions * (vl_size)(self->maxExponent - self->minExponent + 1) ;
  self->table = (double*) vl_malloc(sizeof(double) * (rowLength * numRows + 2 * (order + 1))) ;
  if (!self->table) {
    vl_free(self) ;
    return NULL ;
  }
  kappa = self->table + rowLength * numRows ;
  freq = kappa + (order + 1) ;
  L = 2.0 * HKM_PI / self->period ;

  /* the first order+1 frequencies (multiples of L) whose smoothed spectrum is
   * positive, giving up on a frequency after 3i multiples (homkermap.c:392-400) */
  for (i = 0, j = 0 ; i <= order ; ) {
    freq[i] = (double) j ;
    kappa[i] = hkm_smooth_spectrum(self, j * L) ;
    ++ j ;
    if (kappa[i] > 0 || j >= 3 * i) ++ i ;
  }

  /* one row of the map per (exponent, mantissa); the mantissa advances by
   * repeated addition of the subdivision, as in the reference (402-424) */
  row = self->table ;
  for (exponent = self->minExponent ; exponent <= self->maxExponent ; ++ exponent) {
    mantissa = 1.0 ;
    for (s = 0 ; s < self->numSubdivisions ; ++ s, mantissa += self->subdivision) {
      double const x = ldexp(mantissa, (int) exponent) ;
      double const Lxgamma = L * pow(x, self->gamma) ;
      double const Llogx = L * log(x) ;
      *row++ = sqrt(Lxgamma * kappa[0]) ;
      for (j = 1 ; j <= order ; ++ j) {
        double const amp = sqrt(2.0 * Lxgamma * kappa[j]) ;
        *row++ = amp * hkm_cos(freq[j] * Llogx) ;
        *row++ = amp * hkm_sin(freq[j] * Llogx) ;
      }
    }
  }
  return self ;
}

VL_EXPORT void
vl_homogeneouskernelmap_delete (VlHomogeneousKernelMap * self)
{
  if (!self) return ;
  if (self->deviceTable) vlhip_free(self->deviceTable) ;
  vl_free(self->table) ;
  vl_free(self) ;
}

VL_EXPORT vl_size vl_homogeneouskernelmap_get_order (VlHomogeneousKernelMap const * self) { return self->order ; }
VL_EXPORT vl_size vl_homogeneouskernelmap_get_dimension (VlHomogeneousKernelMap const * self) { return 2 * self->order + 1 ; }
VL_EXPORT VlHomogeneousKernelType
vl_homogeneouskernelmap_get_kernel_type (VlHomogeneousKernelMap const * self) { return self->kernelType ; }
VL_EXPORT VlHomogeneousKernelMapWindowType
vl_homogeneouskernelmap_get_window_type (VlHomogeneousKernelMap const * self) { return self->windowType ; }

/* One value (homkermap.c:532-571): x = sign * mantissa * 2^exponent with the
 * mantissa in [1, 2); zero outside the table's exponents; the subdivision is
 * found by repeated subtraction and the two bracketing rows are interpolated. */
#define HKM_EVALUATE(NAME, T)                                                                  \
  VL_EXPORT void NAME (VlHomogeneousKernelMap const * self, T * destination, vl_size stride, double x) \
  {                                                                                            \
    int exponent ;                                                                             \
    vl_uindex j ;                                                                              \
    vl_size const F = 2 * self->order + 1 ;                                                    \
    double mantissa = frexp(x, &exponent) ;                                                    \
    double const sign = (mantissa >= 0.0) ? +1.0 : -1.0 ;                                      \
    double const * lo, * hi ;                                                                  \
    mantissa *= 2 * sign ;                                                                     \
    exponent -- ;                                                                              \
    if (mantissa == 0 || exponent <= self->minExponent || exponent >= self->maxExponent) {     \
      for (j = 0 ; j < F ; ++j) destination[j * stride] = (T) 0.0 ;                            \
      return ;                                                                                 \
    }                                                                                          \
    lo = self->table + (exponent - self->minExponent) * self->numSubdivisions * F ;            \
    mantissa -= 1.0 ;                                                                          \
    while (mantissa >= self->subdivision) {                                                    \
      mantissa -= self->subdivision ;                                                          \
      lo += F ;                                                                                \
    }                                                                                          \
    hi = lo + F ;                                                                              \
    for (j = 0 ; j < F ; ++j)                                                                  \
      destination[j * stride] = (T) sign * ((hi[j] - lo[j]) * (self->numSubdivisions * mantissa) + lo[j]) ; \
  }
HKM_EVALUATE(vl_homogeneouskernelmap_evaluate_d, double)
HKM_EVALUATE(vl_homogeneouskernelmap_evaluate_f, float)

/* the map's table on the device, uploaded on first use (caller holds the lock) */
static double const *
homkermap_device_table (VlHomogeneousKernelMap const * self)
{
  VlHomogeneousKernelMap * mut = (VlHomogeneousKernelMap*) self ;
  vl_size const F = 2 * self->order + 1 ;
  vl_size const rows = self->numSubdivisions * (vl_size)(self->maxExponent - self->minExponent + 1) ;
  if (!mut->deviceTable) {
    mut->deviceTable = vlhip_malloc(sizeof(double) * F * rows) ;
    vlhip_memcpy_h2d(mut->deviceTable, self->table, sizeof(double) * F * rows) ;
  }
  return (double const*) mut->deviceTable ;
}

void
homkermap_evaluate_device (VlHomogeneousKernelMap const * self, vl_type dataType, void * dOut,
                           void const * dIn, vl_size n)
{
  if (n == 0) return ;
  vlhip_lock() ;
  vlhip_homkermap_evaluate((int) dataType, dOut, dIn, n, homkermap_device_table(self), (int) self->order,
                           (int) self->numSubdivisions, (int) self->minExponent, (int) self->maxExponent) ;
  vlhip_unlock() ;
}

VL_EXPORT void
vl_homogeneouskernelmap_evaluate_array (VlHomogeneousKernelMap const * self, vl_type dataType,
                                        void * destination, void const * data, vl_size numData)
{
  vl_size const F = 2 * self->order + 1 ;
  vl_size es ;
  void * dIn, * dOut ;
  if (dataType != VL_TYPE_FLOAT && dataType != VL_TYPE_DOUBLE) {
    VL_PRINTF("vl_homogeneouskernelmap_evaluate_array: unsupported data type %u\n", (unsigned) dataType) ;
    abort() ;
  }
  if (numData == 0) return ;
  es = vl_get_type_size(dataType) ;
  vlhip_lock() ;
  dIn = vlhip_buffer(VLHIP_BUF_DATA, es * numData) ;
  dOut = vlhip_buffer(VLHIP_BUF_ENC, es * F * numData) ;
  vlhip_memcpy_h2d_async(dIn, data, es * numData) ;
  homkermap_evaluate_device(self, dataType, dOut, dIn, numData) ;
  vlhip_memcpy_d2h(destination, dOut, es * F * numData) ;
  vlhip_unlock() ;
}
