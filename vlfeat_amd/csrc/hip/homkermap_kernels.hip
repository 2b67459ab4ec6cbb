// homkermap_kernels.hip -- homogeneous kernel map of whole arrays.
//
// Every value x of the input becomes the 2*order+1 values of its map
// (vl/homkermap.c:532-571): x = sign * mantissa * 2^exponent with the mantissa
// in [1, 2), zero when x is 0 or its exponent is outside (minExponent,
// maxExponent), otherwise linear interpolation between the two table rows that
// bracket the mantissa, the row found by repeated subtraction of the
// subdivision as the reference finds it.  Double arithmetic without
// contraction, so each value is the reference's bit for bit.  The table (at
// most a few tens of KB) stays in the caches.  homkermap_vec_kernel maps 16
// bytes of input per thread; homkermap_kernel (one value per thread) serves
// unaligned arrays and orders above 4.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip.h"

#include <algorithm>

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

// The map of one value (same operations as homkermap_kernel).  With POW2 the
// number of subdivisions is a power of two, so sub = 2^-s is exact and every
// step of the reference's repeated subtraction, mantissa - k sub with
// mantissa in [0, 1), is exact too (k sub is a multiple of 2^-s, itself a
// multiple of the mantissa's ulp, and the difference is smaller than the
// mantissa): k = floor(mantissa * numSub), computed exactly, gives the row and
// the remainder the loop would give.
template <typename T, int ORDER, bool POW2>
__device__ __forceinline__ void homkermap_one(T x, const double* __restrict__ table, int numSub, double sub,
                                              int minExp, int maxExp, T (&o)[2 * ORDER + 1]) {
#pragma clang fp contract(off)
  constexpr int F = 2 * ORDER + 1;
  int exponent;
  double mantissa = frexp((double)x, &exponent);
  const double sign = (mantissa >= 0.0) ? +1.0 : -1.0;
  mantissa *= 2 * sign;
  exponent--;
  if (mantissa == 0 || exponent <= minExp || exponent >= maxExp) {
#pragma unroll
    for (int j = 0; j < F; ++j) o[j] = T(0);
    return;
  }
  const double* lo = table + (size_t)(exponent - minExp) * (size_t)numSub * F;
  mantissa -= 1.0;
  if constexpr (POW2) {
    const int k = (int)(mantissa * (double)numSub);
    mantissa -= (double)k * sub;
    lo += (size_t)k * F;
  } else {
    while (mantissa >= sub) {
      mantissa -= sub;
      lo += F;
    }
  }
  const double* hi = lo + F;
  const double scale = (double)numSub * mantissa;
  double l[F], h[F];
#pragma unroll
  for (int j = 0; j < F; ++j) { l[j] = lo[j]; h[j] = hi[j]; }
#pragma unroll
  for (int j = 0; j < F; ++j) o[j] = (T)((T)sign * ((h[j] - l[j]) * scale + l[j]));
}

// 16 bytes of input (VPT values) at base (zero past n).
template <typename T, int VPT>
__device__ __forceinline__ void homkermap_load(const T* __restrict__ in, long long n, long long base, T (&v)[VPT]) {
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  if (base + VPT <= n) {
    const u32x4 w = *reinterpret_cast<const u32x4*>(in + base);
    __builtin_memcpy(v, &w, 16);
  } else {
#pragma unroll
    for (int j = 0; j < VPT; ++j) v[j] = base + j < n ? in[base + j] : T(0);
  }
}

// The VPT values v loaded from base, mapped; outputs written at out + base F.
template <typename T, int ORDER, bool POW2, int VPT, int W>
__device__ __forceinline__ void homkermap_map_store(const T (&v)[VPT], long long n, long long base,
                                                    const double* table, int numSub, double sub, int minExp,
                                                    int maxExp, T* __restrict__ out) {
  constexpr int F = 2 * ORDER + 1;
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  T o[VPT * F];
#pragma unroll
  for (int j = 0; j < VPT; ++j) {
    T m[F];
    homkermap_one<T, ORDER, POW2>(v[j], table, numSub, sub, minExp, maxExp, m);
#pragma unroll
    for (int f = 0; f < F; ++f) o[j * F + f] = m[f];
  }
  T* dst = out + base * F;
  if (base + VPT <= n) {
#pragma unroll
    for (int q = 0; q < W; ++q) {
      u32x4 w;
      __builtin_memcpy(&w, &o[q * (16 / (int)sizeof(T))], 16);
      *reinterpret_cast<u32x4*>(dst + q * (16 / (int)sizeof(T))) = w;
    }
  } else {
    for (int j = 0; j < VPT && base + j < n; ++j)
#pragma unroll
      for (int f = 0; f < F; ++f) dst[j * F + f] = o[j * F + f];
  }
}

// Persistent form: each workgroup copies the whole table (rows x F doubles,
// at most 84 KB for order 4) into LDS once, then maps 16 bytes of input per
// thread (4 floats or 2 doubles) in a grid-stride loop, reading the table rows
// from LDS.  The VPT x F outputs of a thread are contiguous (W = 2 ORDER + 1
// 16-byte words) and so are a wave's: they pass through LDS so that store q of
// lane l writes word l + 64 q of the wave's run, each store one contiguous
// KiB; a wave that reaches the ragged end writes value by value.  Each pass of
// the loop issues the loads of HU chunks (one grid stride apart) before
// mapping the first, so that more bytes are in flight per wave.  in and out
// 16-byte aligned.
constexpr int HU = 4;
template <typename T, int ORDER, bool POW2>
__global__ __launch_bounds__(256) void homkermap_vec_kernel(const T* __restrict__ in, long long n,
                                                            const double* __restrict__ table, long long tableLen,
                                                            int numSub, double sub, int minExp, int maxExp,
                                                            T* __restrict__ out) {
  extern __shared__ double tsh[];
  constexpr int F = 2 * ORDER + 1;
  constexpr int VPT = 16 / (int)sizeof(T);
  constexpr int W = VPT * F * (int)sizeof(T) / 16;  // 16-byte output words per thread
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  __shared__ u32x4 stage[4][64 * W];  // one run of output words per wave
  for (long long i = threadIdx.x; i < tableLen; i += 256) tsh[i] = table[i];
  __syncthreads();
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  const long long chunks = (n + VPT - 1) / VPT;
  const long long G = (long long)gridDim.x * 256;
  for (long long c0 = blockIdx.x * 256LL + threadIdx.x; c0 < chunks; c0 += G * HU) {
    T v[HU][VPT];
#pragma unroll
    for (int u = 0; u < HU; ++u)
      if (c0 + G * u < chunks) homkermap_load<T, VPT>(in, n, (c0 + G * u) * VPT, v[u]);
#pragma unroll
    for (int u = 0; u < HU; ++u) {
      const long long c = c0 + G * u;
      const long long cw = c - lane;  // the wave's first chunk
      if ((cw + 64) * VPT <= n) {
        // the whole wave maps full chunks: its outputs are one run of 64 W words
        T o[VPT * F];
#pragma unroll
        for (int j = 0; j < VPT; ++j) {
          T m[F];
          homkermap_one<T, ORDER, POW2>(v[u][j], tsh, numSub, sub, minExp, maxExp, m);
#pragma unroll
          for (int f = 0; f < F; ++f) o[j * F + f] = m[f];
        }
#pragma unroll
        for (int q = 0; q < W; ++q) {
          u32x4 w;
          __builtin_memcpy(&w, &o[q * (16 / (int)sizeof(T))], 16);
          stage[wv][lane * W + q] = w;
        }
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
        u32x4* dst = reinterpret_cast<u32x4*>(out + cw * VPT * F);
#pragma unroll
        for (int q = 0; q < W; ++q) dst[lane + 64 * q] = stage[wv][lane + 64 * q];
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
        __builtin_amdgcn_wave_barrier();
      } else if (c < chunks) {
        homkermap_map_store<T, ORDER, POW2, VPT, W>(v[u], n, c * VPT, tsh, numSub, sub, minExp, maxExp, out);
      }
    }
  }
}

template <typename T>
static void homkermap_launch(const T* in, size_t n, const double* table, int order, int numSub, double sub, int minExp,
                             int maxExp, T* out) {
  const bool aligned = ((reinterpret_cast<uintptr_t>(in) | reinterpret_cast<uintptr_t>(out)) & 15) == 0;
  const bool pow2 = (numSub & (numSub - 1)) == 0;
  const long long tableLen = (long long)numSub * (maxExp - minExp + 1) * (2 * order + 1);
  const size_t lds = (size_t)tableLen * sizeof(double);
  if (aligned && order >= 1 && order <= 4 && lds <= 96 * 1024) {
    constexpr int VPT = 16 / (int)sizeof(T);
    const long long blocksNeeded = ((long long)((n + VPT - 1) / VPT) + 255) / 256;
    auto go = [&](auto kern) {
      VLHIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
      int perCu = 0;
      VLHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, (const void*)kern, 256, lds));
      const long long G = std::min<long long>(blocksNeeded, (long long)std::max(perCu, 1) * ctx().num_cus);
      VLHIP_LAUNCH("homkermap", kern, dim3((unsigned)G), dim3(256), lds, in, (long long)n, table, tableLen, numSub,
                   sub, minExp, maxExp, out);
    };
    switch (order * 2 + (pow2 ? 1 : 0)) {
      case 2: go(homkermap_vec_kernel<T, 1, false>); return;
      case 3: go(homkermap_vec_kernel<T, 1, true>); return;
      case 4: go(homkermap_vec_kernel<T, 2, false>); return;
      case 5: go(homkermap_vec_kernel<T, 2, true>); return;
      case 6: go(homkermap_vec_kernel<T, 3, false>); return;
      case 7: go(homkermap_vec_kernel<T, 3, true>); return;
      case 8: go(homkermap_vec_kernel<T, 4, false>); return;
      default: go(homkermap_vec_kernel<T, 4, true>); return;
    }
  }
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
