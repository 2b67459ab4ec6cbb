// project_kernels.hip -- pre-encoding projection of local descriptors.
//
// The recognition application of the reference projects every descriptor
// before encoding it (apps/recognition/encodeImage.m:91-97, trained by
// trainEncoder.m:150-178):
//     y = P (x - c)                       P: D x Din (PCA rows, whitening folded in)
//     y = y * (1 / max(||y||, 1e-12))     when renormalisation is requested
// all in single precision.  Here it is one FP32 matrix-core contraction per
// workgroup of 64 descriptors (v_mfma_f32_16x16x4_f32: exact FP32 products,
// FP32 accumulation), with the centring done as the descriptors are staged
// and the renormalisation done on the accumulators before the single store,
// so a descriptor is read once and its projection written once.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

namespace vlhip {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int PJ_ROWS = 64;  // descriptors per workgroup, 16 per wave
constexpr int PJ_KC = 32;    // input dimensions per LDS chunk
constexpr int PJ_LS = PJ_KC + 2;  // LDS row stride: the fragment reads of a half-wave hit 32 distinct banks

// Workgroup of 4 waves; wave w owns descriptors row0 + 16w .. +15 and all
// NT x 16 output dimensions.  Fragment layout of v_mfma_f32_16x16x4_f32:
// lane l supplies A[l & 15][l >> 4] and B[l >> 4][l & 15] and holds
// C[4 (l >> 4) + r][l & 15], r = 0..3.
template <int NT>
__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ x, long long N, int Din, int D,
                                                      const float* __restrict__ P, const float* __restrict__ c,
                                                      int renorm, float* __restrict__ out) {
  __shared__ float xs[PJ_ROWS][PJ_LS];
  __shared__ float ps[NT * 16][PJ_LS];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const long long row0 = (long long)blockIdx.x * PJ_ROWS;
  f32x4 acc[NT];
#pragma unroll
  for (int t = 0; t < NT; ++t) acc[t] = f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < Din; k0 += PJ_KC) {
    __syncthreads();  // the previous chunk has been consumed
    for (int i = tid; i < PJ_ROWS * PJ_KC; i += 256) {
      const int r = i / PJ_KC, k = i - r * PJ_KC;
      const long long row = row0 + r;
      const int e = k0 + k;
      xs[r][k] = (row < N && e < Din) ? x[row * Din + e] - c[e] : 0.f;
    }
    for (int i = tid; i < NT * 16 * PJ_KC; i += 256) {
      const int d = i / PJ_KC, k = i - d * PJ_KC;
      const int e = k0 + k;
      ps[d][k] = (d < D && e < Din) ? P[(size_t)d * Din + e] : 0.f;
    }
    __syncthreads();
#pragma unroll
    for (int kk = 0; kk < PJ_KC; kk += 4) {
      const float a = xs[16 * w + lr][kk + lk];
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, ps[16 * t + lr][kk + lk], acc[t], 0, 0, 0);
    }
  }
  float scale[4] = {1.f, 1.f, 1.f, 1.f};
  if (renorm) {
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float s = 0.f;
#pragma unroll
      for (int t = 0; t < NT; ++t) s += acc[t][r] * acc[t][r];  // padding columns hold 0
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) s += __shfl_xor(s, o, 64);
      scale[r] = 1.f / fmaxf(1e-12f, sqrtf(s));
    }
  }
#pragma unroll
  for (int t = 0; t < NT; ++t) {
    const int col = 16 * t + lr;
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const long long row = row0 + 16 * w + 4 * lk + r;
      if (row < N && col < D) out[row * D + col] = acc[t][r] * scale[r];
    }
  }
}

void project_t(float* out, const float* x, size_t N, size_t Din, size_t D, const float* P, const float* c,
               int renorm) {
  VLHIP_REQUIRE(Din >= 1 && D >= 1 && D <= 128, "projection needs inDimension >= 1 and 1 <= outDimension <= 128");
  VLHIP_REQUIRE(Din < (1u << 30) && N < (uint64_t(1) << 40), "projection sizes out of range");
  if (N == 0) return;
  const dim3 grid((unsigned)((N + PJ_ROWS - 1) / PJ_ROWS));
  switch ((D + 15) / 16) {
#define PJ_CASE(NT)                                                                                      \
  case NT:                                                                                               \
    VLHIP_LAUNCH("project", project_kernel<NT>, grid, dim3(256), 0, x, (long long)N, (int)Din, (int)D, P, c, \
                 renorm, out);                                                                           \
    break;
    PJ_CASE(1) PJ_CASE(2) PJ_CASE(3) PJ_CASE(4) PJ_CASE(5) PJ_CASE(6) PJ_CASE(7) PJ_CASE(8)
#undef PJ_CASE
    default: fatal("vlhip_project", __FILE__, __LINE__, "outDimension must be <= 128");
  }
}

}  // namespace vlhip

using namespace vlhip;

extern "C" {

void vlhip_project(void* out, void const* data, size_t numData, size_t inDimension, size_t outDimension,
                   void const* projection, void const* center, int renormalize) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  project_t((float*)out, (const float*)data, numData, inDimension, outDimension, (const float*)projection,
            (const float*)center, renormalize);
}

}  // extern "C"
