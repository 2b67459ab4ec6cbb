// project_kernels.hip -- pre-encoding projection of local descriptors.
//
// The recognition application of the reference projects every descriptor
// before encoding it (apps/recognition/encodeImage.m:91-97, trained by
// trainEncoder.m:150-178):
//     y = P (x - c)                       P: D x Din (PCA rows, whitening folded in)
//     y = y * (1 / max(||y||, 1e-12))     when renormalisation is requested
// all in single precision.  Here it is one FP32 matrix-core contraction per
// workgroup of 256 descriptors (v_mfma_f32_16x16x4_f32: exact FP32 products,
// FP32 accumulation), with the centring done as the descriptors are staged
// and the renormalisation done on the accumulators before the single store,
// so a descriptor is read once and its projection written once.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

namespace vlhip {

typedef float f32x4 __attribute__((ext_vector_type(4)));

constexpr int PJ_RT = 4;                 // row tiles of 16 per wave
constexpr int PJ_ROWS = 4 * 16 * PJ_RT;  // descriptors per workgroup (4 waves)
constexpr int PJ_KC = 32;                // input dimensions per chunk
constexpr int PJ_LS = PJ_KC + 2;         // LDS row stride: the fragment reads of a half-wave hit 32 distinct banks
constexpr int PJ_XQ = PJ_ROWS * (PJ_KC / 4) / 256;  // 16-byte words of a descriptor chunk per thread

// Four consecutive input dimensions e0 .. e0 + 3 of `row` (zero outside the
// data), centred; vec: Din is a multiple of 4 and the arrays are 16-byte aligned.
__device__ __forceinline__ f32x4 pj_load4(const float* __restrict__ a, const float* __restrict__ c, long long row,
                                          long long rows, int e0, int Din, bool vec) {
  f32x4 v = f32x4{0.f, 0.f, 0.f, 0.f};
  if (row >= rows) return v;
  if (vec) {
    if (e0 < Din) {
      v = *reinterpret_cast<const f32x4*>(a + row * Din + e0);
      if (c) v -= *reinterpret_cast<const f32x4*>(c + e0);
    }
  } else {
#pragma unroll
    for (int u = 0; u < 4; ++u)
      if (e0 + u < Din) v[u] = a[row * Din + e0 + u] - (c ? c[e0 + u] : 0.f);
  }
  return v;
}

// Workgroup of 4 waves and PJ_ROWS descriptors; wave w owns descriptors
// row0 + 64w .. +63 (PJ_RT tiles of 16) and all NT x 16 output dimensions.
// The input dimensions are swept in chunks of PJ_KC staged in LDS (descriptor
// chunk centred on load, projection chunk); the next chunk is fetched into
// registers while the current one is multiplied.  Fragment layout of
// v_mfma_f32_16x16x4_f32: lane l supplies A[l & 15][l >> 4] and
// B[l >> 4][l & 15] and holds C[4 (l >> 4) + r][l & 15], r = 0..3.
template <int NT>
__global__ __launch_bounds__(256) void project_kernel(const float* __restrict__ x, long long N, int Din, int D,
                                                      const float* __restrict__ P, const float* __restrict__ c,
                                                      int renorm, int vec, float* __restrict__ out) {
  constexpr int PQ = (NT * 16 * (PJ_KC / 4) + 255) / 256;  // 16-byte words of a projection chunk per thread
  __shared__ float xs[PJ_ROWS][PJ_LS];
  __shared__ float ps[NT * 16][PJ_LS];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  const long long row0 = (long long)blockIdx.x * PJ_ROWS;
  f32x4 acc[PJ_RT][NT];
#pragma unroll
  for (int rt = 0; rt < PJ_RT; ++rt)
#pragma unroll
    for (int t = 0; t < NT; ++t) acc[rt][t] = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 xq[PJ_XQ], pq[PQ];
  auto fetch = [&](int k0) {
#pragma unroll
    for (int q = 0; q < PJ_XQ; ++q) {
      const int i = tid + 256 * q, r = i >> 3, c4 = (i & 7) * 4;
      xq[q] = pj_load4(x, c, row0 + r, N, k0 + c4, Din, vec);
    }
#pragma unroll
    for (int q = 0; q < PQ; ++q) {
      const int i = tid + 256 * q, d = i >> 3, c4 = (i & 7) * 4;
      pq[q] = (d < NT * 16) ? pj_load4(P, nullptr, d, D, k0 + c4, Din, vec) : f32x4{0.f, 0.f, 0.f, 0.f};
    }
  };
  fetch(0);
  for (int k0 = 0; k0 < Din; k0 += PJ_KC) {
    __syncthreads();  // the previous chunk has been consumed
#pragma unroll
    for (int q = 0; q < PJ_XQ; ++q) {
      const int i = tid + 256 * q, r = i >> 3, c4 = (i & 7) * 4;
#pragma unroll
      for (int u = 0; u < 4; ++u) xs[r][c4 + u] = xq[q][u];
    }
#pragma unroll
    for (int q = 0; q < PQ; ++q) {
      const int i = tid + 256 * q, d = i >> 3, c4 = (i & 7) * 4;
      if (d < NT * 16) {
#pragma unroll
        for (int u = 0; u < 4; ++u) ps[d][c4 + u] = pq[q][u];
      }
    }
    __syncthreads();
    if (k0 + PJ_KC < Din) fetch(k0 + PJ_KC);
#pragma unroll
    for (int kk = 0; kk < PJ_KC; kk += 4) {
      float b[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) b[t] = ps[16 * t + lr][kk + lk];
#pragma unroll
      for (int rt = 0; rt < PJ_RT; ++rt) {
        const float a = xs[64 * w + 16 * rt + lr][kk + lk];
#pragma unroll
        for (int t = 0; t < NT; ++t) acc[rt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b[t], acc[rt][t], 0, 0, 0);
      }
    }
  }
#pragma unroll
  for (int rt = 0; rt < PJ_RT; ++rt) {
    float scale[4] = {1.f, 1.f, 1.f, 1.f};
    if (renorm) {
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float s = 0.f;
#pragma unroll
        for (int t = 0; t < NT; ++t) s += acc[rt][t][r] * acc[rt][t][r];  // padding columns hold 0
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
        const long long row = row0 + 64 * w + 16 * rt + 4 * lk + r;
        if (row < N && col < D) out[row * D + col] = acc[rt][t][r] * scale[r];
      }
    }
  }
}

void project_t(float* out, const float* x, size_t N, size_t Din, size_t D, const float* P, const float* c,
               int renorm) {
  VLHIP_REQUIRE(Din >= 1 && D >= 1 && D <= 128, "projection needs inDimension >= 1 and 1 <= outDimension <= 128");
  VLHIP_REQUIRE(Din < (1u << 30) && N < (uint64_t(1) << 40), "projection sizes out of range");
  if (N == 0) return;
  const dim3 grid((unsigned)((N + PJ_ROWS - 1) / PJ_ROWS));
  const int vec = (Din % 4 == 0) && (reinterpret_cast<uintptr_t>(x) & 15) == 0 &&
                  (reinterpret_cast<uintptr_t>(P) & 15) == 0 && (reinterpret_cast<uintptr_t>(c) & 15) == 0;
  switch ((D + 15) / 16) {
#define PJ_CASE(NT)                                                                                      \
  case NT:                                                                                               \
    VLHIP_LAUNCH("project", project_kernel<NT>, grid, dim3(256), 0, x, (long long)N, (int)Din, (int)D, P, c, \
                 renorm, vec, out);                                                                      \
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
