// project_kernels.hip -- pre-encoding projection of local descriptors.
//
// The recognition application of the reference projects every descriptor
// before encoding it (apps/recognition/encodeImage.m:91-97, trained by
// trainEncoder.m:150-178):
//     y = P (x - c)                       P: D x Din (PCA rows, whitening folded in)
//     y = y * (1 / max(||y||, 1e-12))     when renormalisation is requested
// all in single precision.  Here it is an FP32 matrix-core contraction
// (v_mfma_f32_16x16x4_f32: exact FP32 products, FP32 accumulation), with the
// centring done on the loaded descriptors and the renormalisation on the
// accumulators before the single store, so a descriptor is read once and its
// projection written once: project_stream_kernel for 16-byte aligned data with
// 32 <= Din <= 128 a multiple of 16 (P in LDS, descriptors streamed from
// global memory), project_kernel (256 descriptors per workgroup, both operands
// through LDS chunks) otherwise.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

#include <algorithm>

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

// Streaming form, for 16-byte aligned data with Din = 16 DQ, 32 <= Din <= 128 (the bench's
// 128 -> 80 projection): P (padded rows) and c are staged in LDS once per
// workgroup of 8 waves; each wave streams its descriptors straight from
// global memory, 32 rows (2 row tiles) at a time, in a grid-stride loop over
// 32-row blocks.  The contraction over the input dimensions is taken in the
// order k = 16m + 4(l >> 4) + s for step s of chunk m, so lane l's A operands
// of four consecutive products are the four consecutive floats
// x[row][16m + 4(l >> 4) .. + 3] (one 16-byte global load) and its B operands
// the four floats P[col][same k] (one 16-byte LDS read per column tile and
// chunk).  Descriptor loads run two chunks ahead.  Same products as
// project_kernel (FP32, centred first); the k order, hence the rounding of
// the sums, differs.
template <int NT, int DQ>
__global__ __launch_bounds__(512, 4) void project_stream_kernel(const float* __restrict__ x, long long N, int D,
                                                                const float* __restrict__ P,
                                                                const float* __restrict__ c, int renorm,
                                                                float* __restrict__ out) {
  static_assert(DQ >= 2, "two chunks of prefetch need two chunks per row");
  constexpr int Din = 16 * DQ;
  constexpr int PS = Din + 4;  // padded LDS row: the 16 rows a half-wave reads start in different banks
  __shared__ __align__(16) float psh[NT * 16][PS];
  __shared__ __align__(16) float csh[Din];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lk = lane >> 4;
  for (int i = tid; i < Din; i += 512) csh[i] = c ? c[i] : 0.f;
  for (int i = tid; i < NT * 16 * (Din / 4); i += 512) {
    const int r = i / (Din / 4), q = i - r * (Din / 4);
    const f32x4 v = r < D ? *reinterpret_cast<const f32x4*>(P + (size_t)r * Din + 4 * q) : f32x4{0.f, 0.f, 0.f, 0.f};
    *reinterpret_cast<f32x4*>(&psh[r][4 * q]) = v;
  }
  __syncthreads();
  const long long nblk = (N + 31) / 32;
  const long long gw = (long long)blockIdx.x * 8 + w, GW = (long long)gridDim.x * 8;
  // block blk: a wave-uniform base and the lane's two row offsets (rows past
  // the data read the last row; their results are not stored)
  auto base_of = [&](long long blk) __attribute__((always_inline)) {
    const long long b = blk < nblk ? blk : nblk - 1;
    const int lo = __builtin_amdgcn_readfirstlane((int)(b & 0xffffffff));
    const int hi = __builtin_amdgcn_readfirstlane((int)(b >> 32));
    return ((long long)hi << 32) | (unsigned)lo;
  };
  auto offs_of = [&](long long b, int (&o)[2]) __attribute__((always_inline)) {
    const long long left = N - b * 32;  // rows of the block inside the data
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
      const int r = 16 * rt + lr;
      o[rt] = (r < left ? r : (int)(left - 1)) * Din + 4 * lk;
    }
  };
  auto load = [&](long long b, const int (&o)[2], int m, f32x4 (&v)[2]) __attribute__((always_inline)) {
    const float* bp = x + b * 32 * Din;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) v[rt] = *reinterpret_cast<const f32x4*>(bp + o[rt] + 16 * m);
  };
  long long bc = base_of(gw), bn;
  int oc[2], on[2];
  offs_of(bc, oc);
  f32x4 xa[2], xb[2], xc[2];
  load(bc, oc, 0, xa);
  load(bc, oc, 1, xb);
  for (long long blk = gw; blk < nblk; blk += GW) {
    bn = base_of(blk + GW);
    offs_of(bn, on);
    f32x4 acc[2][NT];
#pragma unroll
    for (int rt = 0; rt < 2; ++rt)
#pragma unroll
      for (int t = 0; t < NT; ++t) acc[rt][t] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int m = 0; m < DQ; ++m) {
      // chunk m + 2 of this block, or chunk m + 2 - DQ of the next
      if (m + 2 < DQ)
        load(bc, oc, m + 2, xc);
      else
        load(bn, on, m + 2 - DQ, xc);
      const f32x4 cm = *reinterpret_cast<const f32x4*>(&csh[16 * m + 4 * lk]);
      f32x4 pm[NT];
#pragma unroll
      for (int t = 0; t < NT; ++t) pm[t] = *reinterpret_cast<const f32x4*>(&psh[16 * t + lr][16 * m + 4 * lk]);
      f32x4 xv[2];
#pragma unroll
      for (int rt = 0; rt < 2; ++rt) xv[rt] = xa[rt] - cm;
      // P as the A operand: the accumulators hold out^T tiles, so that a lane
      // ends with four consecutive columns of one row (a 16-byte store)
#pragma unroll
      for (int s4 = 0; s4 < 4; ++s4)
#pragma unroll
        for (int rt = 0; rt < 2; ++rt)
#pragma unroll
          for (int t = 0; t < NT; ++t)
            acc[rt][t] = __builtin_amdgcn_mfma_f32_16x16x4f32(pm[t][s4], xv[rt][s4], acc[rt][t], 0, 0, 0);
#pragma unroll
      for (int rt = 0; rt < 2; ++rt) {
        xa[rt] = xb[rt];
        xb[rt] = xc[rt];
      }
      __builtin_amdgcn_sched_barrier(0);
    }
    bc = bn;
    oc[0] = on[0];
    oc[1] = on[1];
    const long long row0 = blk * 32;
    // lane (lr, lk) holds row 16 rt + lr, columns 16 t + 4 lk .. + 3
    const bool vecOut = (D & 3) == 0 && (reinterpret_cast<uintptr_t>(out) & 15) == 0;
#pragma unroll
    for (int rt = 0; rt < 2; ++rt) {
      const long long row = row0 + 16 * rt + lr;
      float scale = 1.f;
      if (renorm) {
        float sq = 0.f;
#pragma unroll
        for (int t = 0; t < NT; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) sq += acc[rt][t][r] * acc[rt][t][r];  // padding columns hold 0
        sq += __shfl_xor(sq, 16, 64);
        sq += __shfl_xor(sq, 32, 64);
        scale = 1.f / fmaxf(1e-12f, sqrtf(sq));
      }
      if (row < N) {
#pragma unroll
        for (int t = 0; t < NT; ++t) {
          const int col = 16 * t + 4 * lk;
          if (col >= D) continue;
          const f32x4 v = acc[rt][t] * scale;
          if (vecOut) {
            *reinterpret_cast<f32x4*>(out + row * D + col) = v;
          } else {
#pragma unroll
            for (int r = 0; r < 4; ++r)
              if (col + r < D) out[row * D + col + r] = v[r];
          }
        }
      }
    }
  }
}

void project_t(float* out, const float* x, size_t N, size_t Din, size_t D, const float* P, const float* c,
               int renorm) {
  VLHIP_REQUIRE(Din >= 1 && D >= 1 && D <= 128, "projection needs inDimension >= 1 and 1 <= outDimension <= 128");
  VLHIP_REQUIRE(Din < (1u << 30) && N < (uint64_t(1) << 40), "projection sizes out of range");
  if (N == 0) return;
  const int vec = (Din % 4 == 0) && (reinterpret_cast<uintptr_t>(x) & 15) == 0 &&
                  (reinterpret_cast<uintptr_t>(P) & 15) == 0 && (reinterpret_cast<uintptr_t>(c) & 15) == 0;
  const int NTn = (int)((D + 15) / 16);
  if (vec && Din % 16 == 0 && Din >= 32 && Din <= 128) {
    auto go = [&](auto kern) {
      int perCu = 0;
      VLHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, (const void*)kern, 512, 0));
      const long long blocksNeeded = ((long long)((N + 31) / 32) + 7) / 8;
      const long long G = std::min<long long>(blocksNeeded, (long long)std::max(perCu, 1) * ctx().num_cus);
      VLHIP_LAUNCH("project", kern, dim3((unsigned)G), dim3(512), 0, x, (long long)N, (int)D, P, c, renorm, out);
    };
    switch (NTn * 16 + (int)(Din / 16)) {
#define PR_GO(NT, DQ) \
  case NT * 16 + DQ:  \
    go(project_stream_kernel<NT, DQ>); \
    return;
#define PR_GO_NT(NT) PR_GO(NT, 2) PR_GO(NT, 3) PR_GO(NT, 4) PR_GO(NT, 5) PR_GO(NT, 6) PR_GO(NT, 7) PR_GO(NT, 8)
      PR_GO_NT(1) PR_GO_NT(2) PR_GO_NT(3) PR_GO_NT(4) PR_GO_NT(5) PR_GO_NT(6) PR_GO_NT(7) PR_GO_NT(8)
#undef PR_GO_NT
#undef PR_GO
      default: break;  // the general kernel below
    }
  }
  const dim3 grid((unsigned)((N + PJ_ROWS - 1) / PJ_ROWS));
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
