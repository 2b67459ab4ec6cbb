// vlhip_device.h -- device helpers shared by the kernels.
#pragma once

#include <hip/hip_runtime.h>
#include <cstdint>
#include <cmath>

namespace vlhip {

constexpr double kLog2Pi = 1.8378770664093453;  // log(2*pi)

// Four consecutive LDS values (16-byte aligned for float, 32 for double).
__device__ __forceinline__ void lds4(const float* p, float (&r)[4]) {
  float4 v = *reinterpret_cast<const float4*>(p);
  r[0] = v.x; r[1] = v.y; r[2] = v.z; r[3] = v.w;
}
__device__ __forceinline__ void lds4(const double* p, double (&r)[4]) {
  double2 a = *reinterpret_cast<const double2*>(p);
  double2 b = *reinterpret_cast<const double2*>(p + 2);
  r[0] = a.x; r[1] = a.y; r[2] = b.x; r[3] = b.y;
}

template <typename T> __device__ __forceinline__ T dev_exp(T x);
template <> __device__ __forceinline__ float dev_exp<float>(float x) { return expf(x); }
template <> __device__ __forceinline__ double dev_exp<double>(double x) { return exp(x); }

template <typename T> __device__ __forceinline__ T dev_sqrt(T x);
template <> __device__ __forceinline__ float dev_sqrt<float>(float x) { return sqrtf(x); }
template <> __device__ __forceinline__ double dev_sqrt<double>(double x) { return sqrt(x); }

template <typename T> __device__ __forceinline__ T neg_inf();
template <> __device__ __forceinline__ float neg_inf<float>() { return -INFINITY; }
template <> __device__ __forceinline__ double neg_inf<double>() { return -(double)INFINITY; }

// Signed square root, fisher.c:463-472 / vlad.c:204-213.
template <typename T> __device__ __forceinline__ T signed_sqrt(T z) {
  return z >= T(0) ? dev_sqrt<T>(z) : -dev_sqrt<T>(-z);
}

// Block-wide sum of one double per thread (blockDim.x == 256), fixed order.
__device__ __forceinline__ double block_sum_256(double v, double* red /* >= 256 */) {
  const int t = threadIdx.x;
  red[t] = v;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] += red[t + s];
    __syncthreads();
  }
  double r = red[0];
  __syncthreads();
  return r;
}

// Reductions over the 16 lanes of a DPP row (lanes 16i .. 16i + 15), every
// lane receiving the result: swaps within pairs and quads, then the row's
// half mirror and mirror (four VALU operations, no LDS round trip).  IEEE
// min, max and addition are commutative, so all 16 lanes end with the same
// bits.
__device__ __forceinline__ float dpp_f(float v, int ctrl) {
  switch (ctrl) {  // (the control must be a constant)
    case 0: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0xB1, 0xF, 0xF, false));
    case 1: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x4E, 0xF, 0xF, false));
    case 2: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x141, 0xF, 0xF, false));
    default: return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), 0x140, 0xF, 0xF, false));
  }
}
__device__ __forceinline__ int dpp_i(int v, int ctrl) {
  switch (ctrl) {
    case 0: return __builtin_amdgcn_update_dpp(0, v, 0xB1, 0xF, 0xF, false);
    case 1: return __builtin_amdgcn_update_dpp(0, v, 0x4E, 0xF, 0xF, false);
    case 2: return __builtin_amdgcn_update_dpp(0, v, 0x141, 0xF, 0xF, false);
    default: return __builtin_amdgcn_update_dpp(0, v, 0x140, 0xF, 0xF, false);
  }
}
__device__ __forceinline__ float row16_max(float v) {
#pragma unroll
  for (int c = 0; c < 4; ++c) v = fmaxf(v, dpp_f(v, c));
  return v;
}
__device__ __forceinline__ float row16_min(float v) {
#pragma unroll
  for (int c = 0; c < 4; ++c) v = fminf(v, dpp_f(v, c));
  return v;
}
__device__ __forceinline__ float row16_sum(float v) {
#pragma unroll
  for (int c = 0; c < 4; ++c) v = v + dpp_f(v, c);
  return v;
}
// first maximum of (value, index) pairs: larger value, then smaller index
__device__ __forceinline__ void row16_argmax(float& v, int& i) {
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const float ov = dpp_f(v, c);
    const int oi = dpp_i(i, c);
    if (ov > v || (ov == v && oi < i)) { v = ov; i = oi; }
  }
}

}  // namespace vlhip
