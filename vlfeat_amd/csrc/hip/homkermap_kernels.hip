// homkermap_kernels.hip -- homogeneous kernel map of whole arrays.
//
// Every value x of the input becomes the 2*order+1 values of its map
// (vl/homkermap.c:532-571): x = sign * mantissa * 2^exponent with the mantissa
// in [1, 2), zero when x is 0 or its exponent is outside (minExponent,
// maxExponent), otherwise linear interpolation between the two table rows that
// bracket the mantissa, the row found by repeated subtraction of the
// subdivision as the reference finds it.  Double arithmetic without
// contraction, so each value is the reference's bit for bit.  One thread per
// input value; the table (at most a few tens of KB) stays in the caches.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip.h"

namespace vlhip {

template <typename T>
__global__ __launch_bounds__(256) void homkermap_kernel(const T* __restrict__ in, long long n,
                                                        const double* __restrict__ table, int order,
                                                        int numSub, int minExp, int maxExp, T* __restrict__ out) {
#pragma clang fp contract(off)
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int F = 2 * order + 1;
  T* dst = out + i * F;
  int exponent;
  double mantissa = frexp((double)in[i], &exponent);
  const double sign = (mantissa >= 0.0) ? +1.0 : -1.0;
  mantissa *= 2 * sign;
  exponent--;
  if (mantissa == 0 || exponent <= minExp || exponent >= maxExp) {
    for (int j = 0; j < F; ++j) dst[j] = T(0);
    return;
  }
  const double sub = 1.0 / (double)numSub;
  const double* lo = table + (size_t)(exponent - minExp) * (size_t)numSub * F;
  mantissa -= 1.0;
  while (mantissa >= sub) {
    mantissa -= sub;
    lo += F;
  }
  const double* hi = lo + F;
  const double scale = (double)numSub * mantissa;
  for (int j = 0; j < F; ++j) dst[j] = (T)((T)sign * ((hi[j] - lo[j]) * scale + lo[j]));
}

}  // namespace vlhip

using namespace vlhip;

extern "C" {

void vlhip_homkermap_evaluate(int dtype, void* out, void const* in, size_t n, double const* table, int order,
                              int numSubdivisions, int minExponent, int maxExponent) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  VLHIP_REQUIRE(order >= 0 && numSubdivisions >= 1 && minExponent < maxExponent, "invalid homogeneous kernel map");
  if (n == 0) return;
  const dim3 grid((unsigned)((n + 255) / 256));
  if (dtype == 1)
    VLHIP_LAUNCH("homkermap", homkermap_kernel<float>, grid, dim3(256), 0, (const float*)in, (long long)n, table,
                 order, numSubdivisions, minExponent, maxExponent, (float*)out);
  else if (dtype == 2)
    VLHIP_LAUNCH("homkermap", homkermap_kernel<double>, grid, dim3(256), 0, (const double*)in, (long long)n, table,
                 order, numSubdivisions, minExponent, maxExponent, (double*)out);
  else
    fatal("vlhip_homkermap_evaluate", __FILE__, __LINE__, "unknown data type");
}

}  // extern "C"
