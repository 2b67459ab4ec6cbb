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

// ORDER > 0: the order is known at compile time and the 2F table reads of a
// value are issued together; ORDER == 0: any order (runtime).
template <typename T, int ORDER>
__global__ __launch_bounds__(256) void homkermap_kernel(const T* __restrict__ in, long long n,
                                                        const double* __restrict__ table, int order,
                                                        int numSub, double sub, int minExp, int maxExp,
                                                        T* __restrict__ out) {
#pragma clang fp contract(off)
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= n) return;
  const int F = ORDER > 0 ? 2 * ORDER + 1 : 2 * order + 1;
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
  const double* lo = table + (size_t)(exponent - minExp) * (size_t)numSub * F;
  mantissa -= 1.0;
  while (mantissa >= sub) {
    mantissa -= sub;
    lo += F;
  }
  const double* hi = lo + F;
  const double scale = (double)numSub * mantissa;
  if constexpr (ORDER > 0) {
    double l[2 * ORDER + 1], h[2 * ORDER + 1];
#pragma unroll
    for (int j = 0; j < 2 * ORDER + 1; ++j) { l[j] = lo[j]; h[j] = hi[j]; }
#pragma unroll
    for (int j = 0; j < 2 * ORDER + 1; ++j) dst[j] = (T)((T)sign * ((h[j] - l[j]) * scale + l[j]));
  } else {
    for (int j = 0; j < F; ++j) dst[j] = (T)((T)sign * ((hi[j] - lo[j]) * scale + lo[j]));
  }
}

template <typename T>
static void homkermap_launch(const T* in, size_t n, const double* table, int order, int numSub, double sub, int minExp,
                             int maxExp, T* out) {
  const dim3 grid((unsigned)((n + 255) / 256));
  switch (order) {
#define HK_CASE(O)                                                                                              \
  case O:                                                                                                       \
    VLHIP_LAUNCH("homkermap", (homkermap_kernel<T, O>), grid, dim3(256), 0, in, (long long)n, table, order,      \
                 numSub, sub, minExp, maxExp, out);                                                             \
    break;
    HK_CASE(1) HK_CASE(2) HK_CASE(3) HK_CASE(4)
#undef HK_CASE
    default:
      VLHIP_LAUNCH("homkermap", (homkermap_kernel<T, 0>), grid, dim3(256), 0, in, (long long)n, table, order, numSub,
                   sub, minExp, maxExp, out);
  }
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
  VLHIP_REQUIRE((n + 255) / 256 < (size_t(1) << 31), "too many values for one homogeneous kernel map launch");
  const double sub = 1.0 / (double)numSubdivisions;  // the map's subdivision, as the reference computes it
  if (dtype == 1)
    homkermap_launch<float>((const float*)in, n, table, order, numSubdivisions, sub, minExponent, maxExponent,
                            (float*)out);
  else if (dtype == 2)
    homkermap_launch<double>((const double*)in, n, table, order, numSubdivisions, sub, minExponent, maxExponent,
                             (double*)out);
  else
    fatal("vlhip_homkermap_evaluate", __FILE__, __LINE__, "unknown data type");
}

}  // extern "C"
