// gate_kernels.hip -- the learned gate of the Fisher encoder.
//
// A three-layer perceptron D -> H1 -> H2 -> K with GELU activations maps a
// descriptor to a distribution over the K mixture components; it is trained
// on the GPU, with the kernels below, to match the exact posteriors that the
// E-step kernel (gmm_kernels.hip, gmm.c:758-791) produces for the same
// descriptors (soft-target cross-entropy).  At encoding time its 16 most
// probable components per descriptor are the candidates for which exact
// posteriors are computed; rows whose predicted mass on those candidates is
// below a threshold fall back to the dense exact E-step.  There is no
// counterpart in the reference.
//
// Arithmetic: the matrix products run on the matrix cores
// (v_mfma_f32_16x16x32_bf16) with BF16 operands and FP32 accumulation;
// master weights, biases, gradients and the Adam state are FP32.
//
// Kernels: gate_input (standardise + round to BF16), gate_gemm (tiled
// 64 x 64 x 32 MFMA GEMM with optional transposed operands, bias and GELU
// epilogue), gate_xent (softmax, loss and its gradient per row), gate_gelu_bwd,
// gate_colsum (bias gradients, fixed-order two-stage reduction), gate_adam,
// gate_to_bf16, and for inference gate_transpose (weights to [out][in]),
// gate_input_pad, gate_gemm_nt (128 x 128 double-buffered MFMA GEMM with
// 16-byte loads; the gate's logits are rounded to BF16) and gate_select
// (softmax mass screen + exact top-M per row).
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

#include <cmath>
#include <cstring>
#include <vector>

namespace vlhip {

typedef float g_f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 g_bf16x8 __attribute__((ext_vector_type(8)));

constexpr int GT_M = 64, GT_N = 64, GT_K = 32, GT_PAD = 8;
constexpr int EPI_NONE = 0, EPI_GELU = 1;

__device__ __forceinline__ float gelu_f(float a) { return 0.5f * a * (1.f + erff(a * 0.70710678118654752f)); }
__device__ __forceinline__ float gelu_grad_f(float a) {
  const float cdf = 0.5f * (1.f + erff(a * 0.70710678118654752f));
  const float pdf = 0.39894228040143268f * expf(-0.5f * a * a);
  return cdf + a * pdf;
}

// x_in = bf16((x - mean) * istd), rows of D.
__global__ void gate_input_kernel(const float* __restrict__ x, long long n, int D, const float* __restrict__ mean,
                                  const float* __restrict__ istd, __bf16* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * D) return;
  const int d = (int)(i % D);
  out[i] = (__bf16)((x[i] - mean[d]) * istd[d]);
}

// C[M x N] = op(A)[M x K] op(B)[K x N] (+ bias[n]); op(A)(m,k) = TA ? A[k*lda+m] : A[m*lda+k],
// op(B)(k,n) = TB ? B[n*ldb+k] : B[k*ldb+n].  Cf receives the FP32 result
// (pre-activation), Cb the BF16 result after the epilogue (GELU or none).
// Workgroup: 4 waves, 64 x 64 tile, each wave 32 x 32 (2 x 2 MFMA tiles).
template <bool TA, bool TB>
__global__ __launch_bounds__(256) void gate_gemm_kernel(int M, int N, int K, const __bf16* __restrict__ A, int lda,
                                                        const __bf16* __restrict__ B, int ldb,
                                                        const float* __restrict__ bias, int epi,
                                                        float* __restrict__ Cf, int ldc, __bf16* __restrict__ Cb,
                                                        int ldcb) {
  __shared__ __align__(16) __bf16 As[GT_M][GT_K + GT_PAD];
  __shared__ __align__(16) __bf16 Bs[GT_N][GT_K + GT_PAD];
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int wm = w & 1, wn = w >> 1, lr = lane & 15, lg = lane >> 4;
  const int m0 = blockIdx.y * GT_M, n0 = blockIdx.x * GT_N;
  g_f32x4 acc[2][2];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u) acc[t][u] = g_f32x4{0.f, 0.f, 0.f, 0.f};
  for (int k0 = 0; k0 < K; k0 += GT_K) {
    // operand tiles, read along the contiguous dimension of each layout
#pragma unroll
    for (int q = 0; q < (GT_M * GT_K) / 256; ++q) {
      const int e = tid + 256 * q;
      int mm, kk;
      if (TA) { kk = e / GT_M; mm = e % GT_M; } else { mm = e / GT_K; kk = e % GT_K; }
      const int gm = m0 + mm, gk = k0 + kk;
      __bf16 v = (__bf16)0.f;
      if (gm < M && gk < K) v = TA ? A[(size_t)gk * lda + gm] : A[(size_t)gm * lda + gk];
      As[mm][kk] = v;
    }
#pragma unroll
    for (int q = 0; q < (GT_N * GT_K) / 256; ++q) {
      const int e = tid + 256 * q;
      int nn, kk;
      if (TB) { nn = e / GT_K; kk = e % GT_K; } else { kk = e / GT_N; nn = e % GT_N; }
      const int gn = n0 + nn, gk = k0 + kk;
      __bf16 v = (__bf16)0.f;
      if (gn < N && gk < K) v = TB ? B[(size_t)gn * ldb + gk] : B[(size_t)gk * ldb + gn];
      Bs[nn][kk] = v;
    }
    __syncthreads();
    g_bf16x8 a[2], b[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) a[t] = *reinterpret_cast<const g_bf16x8*>(&As[32 * wm + 16 * t + lr][8 * lg]);
#pragma unroll
    for (int u = 0; u < 2; ++u) b[u] = *reinterpret_cast<const g_bf16x8*>(&Bs[32 * wn + 16 * u + lr][8 * lg]);
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int u = 0; u < 2; ++u) acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], b[u], acc[t][u], 0, 0, 0);
    __syncthreads();
  }
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int u = 0; u < 2; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 32 * wm + 16 * t + 4 * lg + r, n = n0 + 32 * wn + 16 * u + lr;
        if (m >= M || n >= N) continue;
        float v = acc[t][u][r] + (bias ? bias[n] : 0.f);
        if (Cf) Cf[(size_t)m * ldc + n] = v;
        if (Cb) Cb[(size_t)m * ldcb + n] = (__bf16)(epi == EPI_GELU ? gelu_f(v) : v);
      }
}

// Per row of logits z (K): p = softmax(z), loss = -sum q log p,
// dz = (p - q) * scale, written in FP32 and BF16.  One workgroup per row.
__global__ __launch_bounds__(256) void gate_xent_kernel(const float* __restrict__ Z, const float* __restrict__ Q, int K,
                                                        float scale, float* __restrict__ dZf,
                                                        __bf16* __restrict__ dZb, float* __restrict__ loss) {
  __shared__ double red[256];
  const long long row = blockIdx.x;
  const float* z = Z + row * K;
  const float* q = Q + row * K;
  float mx = -INFINITY;
  for (int k = threadIdx.x; k < K; k += 256) mx = fmaxf(mx, z[k]);
  // block max (as double through the shared sum helper's buffer)
  red[threadIdx.x] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (threadIdx.x < s) red[threadIdx.x] = fmax(red[threadIdx.x], red[threadIdx.x + s]);
    __syncthreads();
  }
  const float m = (float)red[0];
  __syncthreads();
  double se = 0.0;
  for (int k = threadIdx.x; k < K; k += 256) se += (double)expf(z[k] - m);
  const double S = block_sum_256(se, red);
  const float logS = (float)log(S);
  double l = 0.0;
  for (int k = threadIdx.x; k < K; k += 256) {
    const float lp = z[k] - m - logS;
    const float p = expf(lp);
    const float g = (p - q[k]) * scale;
    dZf[row * K + k] = g;
    dZb[row * K + k] = (__bf16)g;
    if (q[k] > 0.f) l -= (double)q[k] * (double)lp;
  }
  const double L = block_sum_256(l, red);
  if (threadIdx.x == 0) loss[row] = (float)L;
}

// dA = dH * gelu'(A), FP32 and BF16.
__global__ void gate_gelu_bwd_kernel(const float* __restrict__ dH, const float* __restrict__ A, long long n,
                                     float* __restrict__ dAf, __bf16* __restrict__ dAb) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float g = dH[i] * gelu_grad_f(A[i]);
  dAf[i] = g;
  dAb[i] = (__bf16)g;
}

// Column sums of X [rows x N] in two fixed-order stages: partial sums over
// row chunks (grid.y chunks), then one thread per column adds the chunks.
__global__ void gate_colsum_partial_kernel(const float* __restrict__ X, int rows, int N, int chunk,
                                           float* __restrict__ part) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  const int r0 = blockIdx.y * chunk, r1 = min(rows, r0 + chunk);
  float s = 0.f;
  for (int r = r0; r < r1; ++r) s += X[(size_t)r * N + n];
  part[(size_t)blockIdx.y * N + n] = s;
}
__global__ void gate_colsum_final_kernel(const float* __restrict__ part, int chunks, int N, float* __restrict__ out) {
  const int n = blockIdx.x * blockDim.x + threadIdx.x;
  if (n >= N) return;
  float s = 0.f;
  for (int c = 0; c < chunks; ++c) s += part[(size_t)c * N + n];
  out[n] = s;
}

// Adam (Kingma & Ba 2015) on the FP32 master parameters.
__global__ void gate_adam_kernel(float* __restrict__ p, const float* __restrict__ g, float* __restrict__ m,
                                 float* __restrict__ v, long long n, float lr, float b1, float b2, float eps,
                                 float bc1, float bc2) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float gi = g[i];
  const float mi = b1 * m[i] + (1.f - b1) * gi;
  const float vi = b2 * v[i] + (1.f - b2) * gi * gi;
  m[i] = mi;
  v[i] = vi;
  p[i] -= lr * (mi / bc1) / (sqrtf(vi / bc2) + eps);
}

__global__ void gate_to_bf16_kernel(const float* __restrict__ in, long long n, __bf16* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) out[i] = (__bf16)in[i];
}

// ---- inference path --------------------------------------------------------
// Weights transposed to [out][in] (input index contiguous, zero-padded to a
// multiple of 8), so that both operands of every product are read along the
// reduction index with 16-byte loads.
__global__ void gate_transpose_kernel(const float* __restrict__ W, int in, int out, int inPad,
                                      __bf16* __restrict__ WT) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (long long)out * inPad) return;
  const int o = (int)(i / inPad), k = (int)(i % inPad);
  WT[i] = (__bf16)(k < in ? W[(size_t)k * out + o] : 0.f);
}

// x_in = bf16((x - mean) * istd) with rows padded to Dq (zeros).
__global__ void gate_input_pad_kernel(const float* __restrict__ x, long long n, int D, int Dq,
                                      const float* __restrict__ mean, const float* __restrict__ istd,
                                      __bf16* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n * Dq) return;
  const long long r = i / Dq;
  const int d = (int)(i % Dq);
  out[i] = (__bf16)(d < D ? (x[r * D + d] - mean[d]) * istd[d] : 0.f);
}

typedef unsigned g_u32x4 __attribute__((ext_vector_type(4)));
constexpr int NT_M = 128, NT_N = 128, NT_K = 32, NT_LD = NT_K + 8;

// One 128 x NT_K operand tile (rows r0.., reduction k0..) into registers as
// 16-byte words, NT_W per thread; rows or k past the end read as zero (K is a
// multiple of 8).
constexpr int NT_W = NT_M * NT_K / 8 / 256;
__device__ __forceinline__ void nt_load(g_u32x4 (&r)[NT_W], const __bf16* __restrict__ P, int ld, int rows, int r0,
                                        int K, int k0) {
#pragma unroll
  for (int q = 0; q < NT_W; ++q) {
    const int v = threadIdx.x + 256 * q;
    const int row = v / (NT_K / 8), kk = (v % (NT_K / 8)) * 8;
    const int gr = r0 + row, gk = k0 + kk;
    r[q] = (gr < rows && gk < K) ? *reinterpret_cast<const g_u32x4*>(P + (size_t)gr * ld + gk)
                                 : g_u32x4{0u, 0u, 0u, 0u};
  }
}
__device__ __forceinline__ void nt_store(const g_u32x4 (&r)[NT_W], __bf16 (*S)[NT_LD]) {
#pragma unroll
  for (int q = 0; q < NT_W; ++q) {
    const int v = threadIdx.x + 256 * q;
    *reinterpret_cast<g_u32x4*>(&S[v / (NT_K / 8)][(v % (NT_K / 8)) * 8]) = r[q];
  }
}

// C[M x N] = bf16(epilogue(A B^T + bias)), A [M x K] and B [N x K] row-major
// BF16 (K a multiple of 8, rows 16-byte aligned).  128 x 128 tile per
// workgroup, 4 waves of 64 x 64 (4 x 4 MFMA tiles), reduction steps of NT_K
// double-buffered in LDS with the next step fetched into registers.
__global__ __launch_bounds__(256) void gate_gemm_nt_kernel(int M, int N, int K, const __bf16* __restrict__ A, int lda,
                                                           const __bf16* __restrict__ B, int ldb,
                                                           const float* __restrict__ bias, int epi,
                                                           __bf16* __restrict__ C, int ldc) {
  __shared__ __align__(16) __bf16 As[2][NT_M][NT_LD];
  __shared__ __align__(16) __bf16 Bs[2][NT_N][NT_LD];
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  const int wm = w & 1, wn = w >> 1, lr = lane & 15, lg = lane >> 4;
  const int m0 = blockIdx.y * NT_M, n0 = blockIdx.x * NT_N;
  g_f32x4 acc[4][4];
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < 4; ++u) acc[t][u] = g_f32x4{0.f, 0.f, 0.f, 0.f};
  g_u32x4 ra[NT_W], rb[NT_W];
  nt_load(ra, A, lda, M, m0, K, 0);
  nt_load(rb, B, ldb, N, n0, K, 0);
  nt_store(ra, As[0]);
  nt_store(rb, Bs[0]);
  __syncthreads();
  int buf = 0;
  for (int k0 = 0; k0 < K; k0 += NT_K) {
    const bool more = k0 + NT_K < K;
    if (more) {
      nt_load(ra, A, lda, M, m0, K, k0 + NT_K);
      nt_load(rb, B, ldb, N, n0, K, k0 + NT_K);
    }
#pragma unroll
    for (int kk = 0; kk < NT_K; kk += 32) {
      g_bf16x8 a[4], b[4];
#pragma unroll
      for (int t = 0; t < 4; ++t)
        a[t] = *reinterpret_cast<const g_bf16x8*>(&As[buf][64 * wm + 16 * t + lr][kk + 8 * lg]);
#pragma unroll
      for (int u = 0; u < 4; ++u)
        b[u] = *reinterpret_cast<const g_bf16x8*>(&Bs[buf][64 * wn + 16 * u + lr][kk + 8 * lg]);
#pragma unroll
      for (int t = 0; t < 4; ++t)
#pragma unroll
        for (int u = 0; u < 4; ++u)
          acc[t][u] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[t], b[u], acc[t][u], 0, 0, 0);
    }
    if (more) {
      // the other buffer was last read before the previous barrier
      nt_store(ra, As[buf ^ 1]);
      nt_store(rb, Bs[buf ^ 1]);
    }
    __syncthreads();
    buf ^= 1;
  }
#pragma unroll
  for (int t = 0; t < 4; ++t)
#pragma unroll
    for (int u = 0; u < 4; ++u)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + 64 * wm + 16 * t + 4 * lg + r, n = n0 + 64 * wn + 16 * u + lr;
        if (m >= M || n >= N) continue;
        const float v = acc[t][u][r] + bias[n];
        C[(size_t)m * ldc + n] = (__bf16)(epi == EPI_GELU ? gelu_f(v) : v);
      }
}

// Candidate selection, one wave per row of BF16 logits z (K <= 64 * GS_VPL):
// p = softmax(z).  Screen: with c = #{p >= t} and S = sum_{p >= t} p, the mass
// of the M most probable components is at most S when c >= M and at most
// S + (M - c) t otherwise; rows whose bound is below `need` are marked
// dense without selecting.  The others select the M largest exactly (M
// rounds of a wave-wide argmax, ties to the lower index), and are confident
// when the selected mass reaches `need`.  idx receives M component indices
// (-1 for dense rows), mass the selected mass (or the bound).
constexpr int GS_MAXV = 32;  // components per lane at most (K <= 2048)
// Lane l holds components 8(l + 64q) + e, e < 8, q < GS_VPL / 8 (16-byte loads
// of the row's BF16 logits; the row stride ldz is a multiple of 8).
template <int GS_VPL>
__global__ __launch_bounds__(256) void gate_select_kernel(const __bf16* __restrict__ Z, int ldz, long long rows, int K,
                                                          int M, float t, float need, int* __restrict__ idx,
                                                          float* __restrict__ mass, unsigned char* __restrict__ conf) {
  const int lane = threadIdx.x & 63;
  const long long row = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (row >= rows) return;
  const __bf16* z = Z + row * ldz;
  float v[GS_VPL];
  float mx = -INFINITY;
#pragma unroll
  for (int q = 0; q < GS_VPL / 8; ++q) {
    const int k0 = 8 * (lane + 64 * q);
    g_bf16x8 zz = g_bf16x8{};
    if (k0 < K) zz = *reinterpret_cast<const g_bf16x8*>(z + k0);
#pragma unroll
    for (int e = 0; e < 8; ++e) {
      const int j = 8 * q + e;
      v[j] = (k0 + e < K) ? (float)zz[e] : -INFINITY;
      mx = fmaxf(mx, v[j]);
    }
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
  float s = 0.f;
#pragma unroll
  for (int j = 0; j < GS_VPL; ++j) {
    v[j] = __expf(v[j] - mx);  // unnormalised probabilities (0 for padding)
    s += v[j];
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  const float ts = t * s;
  float cs = 0.f, ms = 0.f;
#pragma unroll
  for (int j = 0; j < GS_VPL; ++j) {
    const bool above = v[j] >= ts;
    cs += above ? 1.f : 0.f;
    ms += above ? v[j] : 0.f;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    cs += __shfl_xor(cs, o, 64);
    ms += __shfl_xor(ms, o, 64);
  }
  const float bound = (cs >= (float)M) ? ms / s : ms / s + ((float)M - cs) * t;
  int* out = idx + row * M;
  if (bound < need) {
    if (lane < M) out[lane] = -1;
    if (lane == 0) { mass[row] = bound; conf[row] = 0; }
    return;
  }
  float sel = 0.f;
  for (int it = 0; it < M; ++it) {
    float bv = -1.f;
    int bj = 0;
#pragma unroll
    for (int j = 0; j < GS_VPL; ++j)
      if (v[j] > bv) { bv = v[j]; bj = j; }
    int bk = 8 * (lane + 64 * (bj >> 3)) + (bj & 7);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
      const float ov = __shfl_xor(bv, o, 64);
      const int ok = __shfl_xor(bk, o, 64);
      if (ov > bv || (ov == bv && ok < bk)) { bv = ov; bk = ok; }
    }
    sel += bv;
    if (lane == 0) out[it] = bk;
    if (((bk >> 3) & 63) == lane) {
#pragma unroll
      for (int j = 0; j < GS_VPL; ++j)
        if (8 * (lane + 64 * (j >> 3)) + (j & 7) == bk) v[j] = -1.f;
    }
  }
  if (lane == 0) {
    mass[row] = sel / s;
    conf[row] = (sel / s >= need) ? 1 : 0;
  }
}

// ---------------------------------------------------------------------------

static void* dev_alloc(size_t bytes) {
  void* p = nullptr;
  VLHIP_CHECK(hipMalloc(&p, bytes ? bytes : 16));
  return p;
}

struct Gate {
  int D, K, H1, H2;
  size_t nW1, nW2, nW3, nParams;
  float* P = nullptr;   // FP32 master parameters [W1 | b1 | W2 | b2 | W3 | b3]
  float* G = nullptr;   // gradients, same layout
  float* Am = nullptr;  // Adam first moments
  float* Av = nullptr;  // Adam second moments
  __bf16* Wb = nullptr; // BF16 copies of W1, W2, W3 (contiguous)
  int Dq = 0;           // D rounded up to 8
  int Kq = 0;           // K rounded up to 8 (row stride of the BF16 logits)
  __bf16* WT = nullptr; // inference copies [W1^T (H1 x Dq) | W2^T (H2 x H1) | W3^T (K x H2)]
  long long icap = 0;   // inference buffers for up to icap rows
  __bf16 *Xi = nullptr, *H1i = nullptr, *H2i = nullptr, *Zi = nullptr;
  float* mean = nullptr;
  float* istd = nullptr;
  long long step = 0;
  // work buffers for up to cap rows
  long long cap = 0;
  __bf16 *Xb = nullptr, *H1b = nullptr, *H2b = nullptr, *dZb = nullptr, *dA2b = nullptr, *dA1b = nullptr;
  float *A1 = nullptr, *A2 = nullptr, *Z = nullptr, *Q = nullptr, *dZf = nullptr, *dHf = nullptr, *dAf = nullptr;
  float *loss = nullptr, *part = nullptr;

  float* W1() { return P; }
  float* b1() { return P + nW1; }
  float* W2() { return P + nW1 + H1; }
  float* b2() { return P + nW1 + H1 + nW2; }
  float* W3() { return P + nW1 + H1 + nW2 + H2; }
  float* b3() { return P + nW1 + H1 + nW2 + H2 + nW3; }
  __bf16* W1T() { return WT; }
  __bf16* W2T() { return WT + (size_t)H1 * Dq; }
  __bf16* W3T() { return WT + (size_t)H1 * Dq + (size_t)H2 * H1; }
  __bf16* W1b() { return Wb; }
  __bf16* W2b() { return Wb + nW1; }
  __bf16* W3b() { return Wb + nW1 + nW2; }
};

static void free_infer(Gate& g);

static void free_work(Gate& g) {
  void* bufs[] = {g.Xb, g.H1b, g.H2b, g.dZb, g.dA2b, g.dA1b, g.A1, g.A2, g.Z, g.Q, g.dZf, g.dHf, g.dAf, g.loss, g.part};
  for (void* b : bufs)
    if (b) VLHIP_CHECK(hipFree(b));
  g.Xb = g.H1b = g.H2b = g.dZb = g.dA2b = g.dA1b = nullptr;
  g.A1 = g.A2 = g.Z = g.Q = g.dZf = g.dHf = g.dAf = g.loss = g.part = nullptr;
  g.cap = 0;
}

static void ensure_work(Gate& g, long long rows) {
  if (rows <= g.cap) return;
  sync();
  free_work(g);
  const long long r = rows;
  const int Hm = g.H1 > g.H2 ? g.H1 : g.H2;
  g.Xb = (__bf16*)dev_alloc(sizeof(__bf16) * r * g.D);
  g.H1b = (__bf16*)dev_alloc(sizeof(__bf16) * r * g.H1);
  g.H2b = (__bf16*)dev_alloc(sizeof(__bf16) * r * g.H2);
  g.dZb = (__bf16*)dev_alloc(sizeof(__bf16) * r * g.K);
  g.dA2b = (__bf16*)dev_alloc(sizeof(__bf16) * r * g.H2);
  g.dA1b = (__bf16*)dev_alloc(sizeof(__bf16) * r * g.H1);
  g.A1 = (float*)dev_alloc(sizeof(float) * r * g.H1);
  g.A2 = (float*)dev_alloc(sizeof(float) * r * g.H2);
  g.Z = (float*)dev_alloc(sizeof(float) * r * g.K);
  g.Q = (float*)dev_alloc(sizeof(float) * r * g.K);
  g.dZf = (float*)dev_alloc(sizeof(float) * r * g.K);
  g.dHf = (float*)dev_alloc(sizeof(float) * r * Hm);
  g.dAf = (float*)dev_alloc(sizeof(float) * r * Hm);
  g.loss = (float*)dev_alloc(sizeof(float) * r);
  g.part = (float*)dev_alloc(sizeof(float) * 64 * (size_t)(g.K > Hm ? g.K : Hm));
  g.cap = rows;
}

template <bool TA, bool TB>
static void gemm(const char* name, int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb,
                 const float* bias, int epi, float* Cf, int ldc, __bf16* Cb, int ldcb) {
  if (M <= 0 || N <= 0) return;
  dim3 grid((N + GT_N - 1) / GT_N, (M + GT_M - 1) / GT_M);
  VLHIP_LAUNCH(name, (gate_gemm_kernel<TA, TB>), grid, dim3(256), 0, M, N, K, A, lda, B, ldb, bias, epi, Cf, ldc,
               Cb, ldcb);
}

static void colsum(const float* X, int rows, int N, float* out, float* part) {
  const int chunks = rows < 64 ? (rows > 0 ? rows : 1) : 64;
  const int chunk = (rows + chunks - 1) / chunks;
  VLHIP_LAUNCH("gate_colsum", gate_colsum_partial_kernel, dim3((N + 255) / 256, chunks), dim3(256), 0, X, rows, N,
               chunk, part);
  VLHIP_LAUNCH("gate_colsum", gate_colsum_final_kernel, dim3((N + 255) / 256), dim3(256), 0, part, chunks, N, out);
}

static inline unsigned grid1(long long n) { return (unsigned)((n + 255) / 256); }

// Forward pass of `rows` descriptors (device FP32, rows x D) into g.Z.
static void forward(Gate& g, const float* x, long long rows) {
  const int r = (int)rows;
  VLHIP_LAUNCH("gate_input", gate_input_kernel, dim3(grid1(rows * g.D)), dim3(256), 0, x, rows, g.D, g.mean, g.istd,
               g.Xb);
  gemm<false, false>("gate_gemm_fwd", r, g.H1, g.D, g.Xb, g.D, g.W1b(), g.H1, g.b1(), EPI_GELU, g.A1, g.H1, g.H1b,
                     g.H1);
  gemm<false, false>("gate_gemm_fwd", r, g.H2, g.H1, g.H1b, g.H1, g.W2b(), g.H2, g.b2(), EPI_GELU, g.A2, g.H2, g.H2b,
                     g.H2);
  gemm<false, false>("gate_gemm_fwd", r, g.K, g.H2, g.H2b, g.H2, g.W3b(), g.K, g.b3(), EPI_NONE, g.Z, g.K, nullptr, 0);
}

// Backward pass for targets g.Q (rows x K) after forward(); gradients of the
// mean loss into g.G.
static void backward(Gate& g, long long rows) {
  const int r = (int)rows;
  VLHIP_LAUNCH("gate_xent", gate_xent_kernel, dim3((unsigned)rows), dim3(256), 0, g.Z, g.Q, g.K, 1.0f / (float)rows,
               g.dZf, g.dZb, g.loss);
  float* gW1 = g.G;
  float* gb1 = g.G + g.nW1;
  float* gW2 = gb1 + g.H1;
  float* gb2 = gW2 + g.nW2;
  float* gW3 = gb2 + g.H2;
  float* gb3 = gW3 + g.nW3;
  gemm<true, false>("gate_gemm_bwd", g.H2, g.K, r, g.H2b, g.H2, g.dZb, g.K, nullptr, EPI_NONE, gW3, g.K, nullptr, 0);
  colsum(g.dZf, r, g.K, gb3, g.part);
  gemm<false, true>("gate_gemm_bwd", r, g.H2, g.K, g.dZb, g.K, g.W3b(), g.K, nullptr, EPI_NONE, g.dHf, g.H2, nullptr,
                    0);
  VLHIP_LAUNCH("gate_gelu_bwd", gate_gelu_bwd_kernel, dim3(grid1(rows * g.H2)), dim3(256), 0, g.dHf, g.A2,
               rows * g.H2, g.dAf, g.dA2b);
  gemm<true, false>("gate_gemm_bwd", g.H1, g.H2, r, g.H1b, g.H1, g.dA2b, g.H2, nullptr, EPI_NONE, gW2, g.H2, nullptr,
                    0);
  colsum(g.dAf, r, g.H2, gb2, g.part);
  gemm<false, true>("gate_gemm_bwd", r, g.H1, g.H2, g.dA2b, g.H2, g.W2b(), g.H2, nullptr, EPI_NONE, g.dHf, g.H1,
                    nullptr, 0);
  VLHIP_LAUNCH("gate_gelu_bwd", gate_gelu_bwd_kernel, dim3(grid1(rows * g.H1)), dim3(256), 0, g.dHf, g.A1,
               rows * g.H1, g.dAf, g.dA1b);
  gemm<true, false>("gate_gemm_bwd", g.D, g.H1, r, g.Xb, g.D, g.dA1b, g.H1, nullptr, EPI_NONE, gW1, g.H1, nullptr, 0);
  colsum(g.dAf, r, g.H1, gb1, g.part);
}

static void refresh_bf16(Gate& g) {
  VLHIP_LAUNCH("gate_transpose", gate_transpose_kernel, dim3(grid1((long long)g.H1 * g.Dq)), dim3(256), 0, g.W1(), g.D,
               g.H1, g.Dq, g.W1T());
  VLHIP_LAUNCH("gate_transpose", gate_transpose_kernel, dim3(grid1((long long)g.H2 * g.H1)), dim3(256), 0, g.W2(), g.H1,
               g.H2, g.H1, g.W2T());
  VLHIP_LAUNCH("gate_transpose", gate_transpose_kernel, dim3(grid1((long long)g.K * g.H2)), dim3(256), 0, g.W3(), g.H2,
               g.K, g.H2, g.W3T());
  VLHIP_LAUNCH("gate_to_bf16", gate_to_bf16_kernel, dim3(grid1(g.nW1)), dim3(256), 0, g.W1(), (long long)g.nW1,
               g.W1b());
  VLHIP_LAUNCH("gate_to_bf16", gate_to_bf16_kernel, dim3(grid1(g.nW2)), dim3(256), 0, g.W2(), (long long)g.nW2,
               g.W2b());
  VLHIP_LAUNCH("gate_to_bf16", gate_to_bf16_kernel, dim3(grid1(g.nW3)), dim3(256), 0, g.W3(), (long long)g.nW3,
               g.W3b());
}

static double mean_loss(Gate& g, long long rows) {
  std::vector<float> h(rows);
  download(h.data(), g.loss, rows * sizeof(float));
  double s = 0.0;
  for (long long i = 0; i < rows; ++i) s += h[i];
  return rows ? s / (double)rows : 0.0;
}

// Mersenne-free deterministic initialisation: a 64-bit LCG (Knuth MMIX
// constants) drives U(-1/sqrt(fan_in), 1/sqrt(fan_in)) for weights and biases.
static void init_params(Gate& g, unsigned long long seed) {
  std::vector<float> h(g.nParams);
  unsigned long long s = seed * 2862933555777941757ULL + 3037000493ULL;
  auto uni = [&]() {
    s = s * 6364136223846793005ULL + 1442695040888963407ULL;
    return (double)(s >> 11) * (1.0 / 9007199254740992.0);
  };
  size_t o = 0;
  const int fan[3] = {g.D, g.H1, g.H2};
  const size_t nw[3] = {g.nW1, g.nW2, g.nW3};
  const int nb[3] = {g.H1, g.H2, g.K};
  for (int l = 0; l < 3; ++l) {
    const double a = 1.0 / std::sqrt((double)fan[l]);
    for (size_t i = 0; i < nw[l]; ++i) h[o++] = (float)((2.0 * uni() - 1.0) * a);
    for (int i = 0; i < nb[l]; ++i) h[o++] = (float)((2.0 * uni() - 1.0) * a);
  }
  upload(g.P, h.data(), g.nParams * sizeof(float));
  sync();
}

}  // namespace vlhip

using namespace vlhip;

extern "C" {

void* vlhip_gate_new(size_t dimension, size_t numClusters, size_t hidden1, size_t hidden2, unsigned long long seed) {
  VLHIP_REQUIRE(dimension >= 1 && numClusters >= 1 && hidden1 >= 1 && hidden2 >= 1, "gate sizes must be positive");
  VLHIP_REQUIRE(numClusters <= (size_t)64 * GS_MAXV, "the gate supports at most 2048 components");
  VLHIP_REQUIRE(hidden1 % 8 == 0 && hidden2 % 8 == 0, "hidden widths must be multiples of 8");
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  Gate* g = new Gate();
  g->D = (int)dimension;
  g->K = (int)numClusters;
  g->H1 = (int)hidden1;
  g->H2 = (int)hidden2;
  g->nW1 = (size_t)g->D * g->H1;
  g->nW2 = (size_t)g->H1 * g->H2;
  g->nW3 = (size_t)g->H2 * g->K;
  g->nParams = g->nW1 + g->H1 + g->nW2 + g->H2 + g->nW3 + g->K;
  g->P = (float*)dev_alloc(sizeof(float) * g->nParams);
  g->G = (float*)dev_alloc(sizeof(float) * g->nParams);
  g->Am = (float*)dev_alloc(sizeof(float) * g->nParams);
  g->Av = (float*)dev_alloc(sizeof(float) * g->nParams);
  g->Wb = (__bf16*)dev_alloc(sizeof(__bf16) * (g->nW1 + g->nW2 + g->nW3));
  g->Dq = (g->D + 7) / 8 * 8;
  g->Kq = (g->K + 7) / 8 * 8;
  g->WT = (__bf16*)dev_alloc(sizeof(__bf16) * ((size_t)g->H1 * g->Dq + (size_t)g->H2 * g->H1 + (size_t)g->K * g->H2));
  g->mean = (float*)dev_alloc(sizeof(float) * g->D);
  g->istd = (float*)dev_alloc(sizeof(float) * g->D);
  VLHIP_CHECK(hipMemsetAsync(g->Am, 0, sizeof(float) * g->nParams, ctx().stream));
  VLHIP_CHECK(hipMemsetAsync(g->Av, 0, sizeof(float) * g->nParams, ctx().stream));
  VLHIP_CHECK(hipMemsetAsync(g->G, 0, sizeof(float) * g->nParams, ctx().stream));
  std::vector<float> zero(g->D, 0.f), one(g->D, 1.f);
  upload(g->mean, zero.data(), sizeof(float) * g->D);
  upload(g->istd, one.data(), sizeof(float) * g->D);
  init_params(*g, seed);
  refresh_bf16(*g);
  sync();
  return g;
}

void vlhip_gate_delete(void* gate) {
  if (!gate) return;
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  Gate* g = (Gate*)gate;
  sync();
  free_work(*g);
  free_infer(*g);
  void* bufs[] = {g->P, g->G, g->Am, g->Av, g->Wb, g->WT, g->mean, g->istd};
  for (void* b : bufs)
    if (b) VLHIP_CHECK(hipFree(b));
  delete g;
}

size_t vlhip_gate_num_params(void* gate) { return ((Gate*)gate)->nParams; }

void vlhip_gate_get_params(void* gate, float* params_host, float* mean_host, float* istd_host) {
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  Gate* g = (Gate*)gate;
  if (params_host) download(params_host, g->P, g->nParams * sizeof(float));
  if (mean_host) download(mean_host, g->mean, g->D * sizeof(float));
  if (istd_host) download(istd_host, g->istd, g->D * sizeof(float));
}

void vlhip_gate_set_params(void* gate, float const* params_host, float const* mean_host, float const* istd_host) {
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  Gate* g = (Gate*)gate;
  if (params_host) upload(g->P, params_host, g->nParams * sizeof(float));
  if (mean_host) upload(g->mean, mean_host, g->D * sizeof(float));
  if (istd_host) upload(g->istd, istd_host, g->D * sizeof(float));
  refresh_bf16(*g);
  sync();
}

void vlhip_gate_set_standardization_from_data(void* gate, void const* data, size_t numData) {
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  Gate* g = (Gate*)gate;
  VLHIP_REQUIRE(numData >= 2, "standardisation needs at least two descriptors");
  std::vector<float> h((size_t)numData * g->D);
  download(h.data(), data, h.size() * sizeof(float));
  std::vector<double> mu(g->D, 0.0), var(g->D, 0.0);
  for (size_t i = 0; i < numData; ++i)
    for (int d = 0; d < g->D; ++d) mu[d] += h[i * g->D + d];
  for (int d = 0; d < g->D; ++d) mu[d] /= (double)numData;
  for (size_t i = 0; i < numData; ++i)
    for (int d = 0; d < g->D; ++d) {
      const double t = h[i * g->D + d] - mu[d];
      var[d] += t * t;
    }
  std::vector<float> m(g->D), s(g->D);
  for (int d = 0; d < g->D; ++d) {
    m[d] = (float)mu[d];
    const double sd = std::sqrt(var[d] / (double)(numData - 1));
    s[d] = (float)(sd > 1e-12 ? 1.0 / sd : 1.0);
  }
  upload(g->mean, m.data(), g->D * sizeof(float));
  upload(g->istd, s.data(), g->D * sizeof(float));
  sync();
}

// Loss and gradients of the mean soft-target cross-entropy for one batch
// (data rows x D, targets rows x K, both FP32 device arrays); grads_host
// receives nParams floats in the parameter layout.  Returns the mean loss.
double vlhip_gate_gradients(void* gate, void const* data, void const* targets, size_t rows, float* grads_host) {
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  Gate* g = (Gate*)gate;
  VLHIP_REQUIRE(rows >= 1, "empty batch");
  ensure_work(*g, (long long)rows);
  VLHIP_CHECK(hipMemcpyAsync(g->Q, targets, rows * g->K * sizeof(float), hipMemcpyDeviceToDevice, ctx().stream));
  forward(*g, (const float*)data, (long long)rows);
  backward(*g, (long long)rows);
  if (grads_host) download(grads_host, g->G, g->nParams * sizeof(float));
  return mean_loss(*g, (long long)rows);
}

// Training on numData descriptors (device FP32) against the exact posteriors
// of the GMM (device FP32 means, covariances, priors): `steps` Adam steps on
// consecutive batches of `batch` rows (wrapping around).  Returns the mean
// loss of the last step; loss_trace_host (optional, steps entries) receives
// every step's loss.
double vlhip_gate_train(void* gate, void const* data, size_t numData, void const* means, void const* covariances,
                        void const* priors, size_t steps, size_t batch, float lr, float* loss_trace_host) {
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  Gate* g = (Gate*)gate;
  VLHIP_REQUIRE(numData >= batch && batch >= 1, "training needs at least one full batch");
  ensure_work(*g, (long long)batch);
  GmmPrep<float> prep = gmm_prepare<float>((const float*)priors, (const float*)covariances, g->D, g->K);
  const float* x = (const float*)data;
  double last = 0.0;
  size_t pos = 0;
  for (size_t s = 0; s < steps; ++s) {
    if (pos + batch > numData) pos = 0;
    const float* xb = x + pos * g->D;
    run_estep<float>(xb, batch, g->D, g->K, (const float*)means, prep, g->Q, nullptr, nullptr, nullptr,
                     "gate_targets", false);
    forward(*g, xb, (long long)batch);
    backward(*g, (long long)batch);
    g->step += 1;
    const float b1 = 0.9f, b2 = 0.999f;
    const float bc1 = 1.f - std::pow(b1, (float)g->step), bc2 = 1.f - std::pow(b2, (float)g->step);
    VLHIP_LAUNCH("gate_adam", gate_adam_kernel, dim3(grid1((long long)g->nParams)), dim3(256), 0, g->P, g->G, g->Am,
                 g->Av, (long long)g->nParams, lr, b1, b2, 1e-8f, bc1, bc2);
    refresh_bf16(*g);
    if (loss_trace_host) loss_trace_host[s] = (float)mean_loss(*g, (long long)batch);
    pos += batch;
  }
  last = mean_loss(*g, (long long)batch);
  return last;
}

}  // extern "C"

namespace vlhip {

static void free_infer(Gate& g) {
  void* bufs[] = {g.Xi, g.H1i, g.H2i, g.Zi};
  for (void* b : bufs)
    if (b) VLHIP_CHECK(hipFree(b));
  g.Xi = g.H1i = g.H2i = g.Zi = nullptr;
  g.icap = 0;
}

static void ensure_infer(Gate& g, long long rows) {
  if (rows <= g.icap) return;
  sync();
  free_infer(g);
  g.Xi = (__bf16*)dev_alloc(sizeof(__bf16) * rows * g.Dq);
  g.H1i = (__bf16*)dev_alloc(sizeof(__bf16) * rows * g.H1);
  g.H2i = (__bf16*)dev_alloc(sizeof(__bf16) * rows * g.H2);
  g.Zi = (__bf16*)dev_alloc(sizeof(__bf16) * rows * g.Kq);
  g.icap = rows;
}

static void gemm_nt(int M, int N, int K, const __bf16* A, int lda, const __bf16* B, int ldb, const float* bias,
                    int epi, __bf16* C, int ldc) {
  dim3 grid((N + NT_N - 1) / NT_N, (M + NT_M - 1) / NT_M);
  VLHIP_LAUNCH("gate_gemm_nt", gate_gemm_nt_kernel, grid, dim3(256), 0, M, N, K, A, lda, B, ldb, bias, epi, C, ldc);
}

// Inference forward of rows descriptors into g.Zi (BF16 logits: the gate's output).
static void infer(Gate& g, const float* x, long long rows) {
  const int r = (int)rows;
  VLHIP_LAUNCH("gate_input", gate_input_pad_kernel, dim3(grid1(rows * g.Dq)), dim3(256), 0, x, rows, g.D, g.Dq,
               g.mean, g.istd, g.Xi);
  gemm_nt(r, g.H1, g.Dq, g.Xi, g.Dq, g.W1T(), g.Dq, g.b1(), EPI_GELU, g.H1i, g.H1);
  gemm_nt(r, g.H2, g.H1, g.H1i, g.H1, g.W2T(), g.H1, g.b2(), EPI_GELU, g.H2i, g.H2);
  gemm_nt(r, g.K, g.H2, g.H2i, g.H2, g.W3T(), g.H2, g.b3(), EPI_NONE, g.Zi, g.Kq);
}

__global__ void gate_bf16_to_f32_kernel(const __bf16* __restrict__ in, long long rows, int K, int ld,
                                        float* __restrict__ out) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < rows * K) out[i] = (float)in[(i / K) * ld + i % K];
}

}  // namespace vlhip

extern "C" {

// Candidates for rows descriptors (device FP32): idx (device, rows x M int32),
// mass (device, rows FP32) and conf (device, rows bytes) as gate_select_kernel
// defines them; processed in chunks of at most 65536 rows.
void vlhip_gate_select(void* gate, void const* data, size_t rows, int M, float need, int* idx, float* mass,
                       unsigned char* conf) {
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  Gate* g = (Gate*)gate;
  VLHIP_REQUIRE(M >= 1 && M <= 64 && M <= g->K, "candidate count must be in [1, min(64, K)]");
  VLHIP_REQUIRE(g->Kq / 8 <= 64 * (GS_MAXV / 8), "the gate supports at most 2048 components");
  const long long chunk = 262144;
  ensure_infer(*g, rows < (size_t)chunk ? (long long)rows : chunk);
  const float* x = (const float*)data;
  const float t = need / (float)(4 * M);  // screen threshold
  const int vpl = (g->Kq / 8 + 63) / 64 * 8;  // values per lane, a multiple of 8
  for (size_t r0 = 0; r0 < rows; r0 += (size_t)chunk) {
    const long long n = (long long)((rows - r0) < (size_t)chunk ? (rows - r0) : (size_t)chunk);
    infer(*g, x + r0 * g->D, n);
    const dim3 grid((unsigned)((n + 3) / 4)), block(256);
#define GATE_SELECT(V)                                                                                        \
    VLHIP_LAUNCH("gate_select", gate_select_kernel<V>, grid, block, 0, g->Zi, g->Kq, n, g->K, M, t, need,       \
                 idx + r0 * M,                                                                                \
                 mass + r0, conf + r0)
    if (vpl <= 8) GATE_SELECT(8);
    else if (vpl <= 16) GATE_SELECT(16);
    else GATE_SELECT(32);
#undef GATE_SELECT
  }
}

// Logits of the gate for rows descriptors into Z_host (rows x K, host).
void vlhip_gate_logits(void* gate, void const* data, size_t rows, float* logits_host) {
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  Gate* g = (Gate*)gate;
  ensure_infer(*g, (long long)rows);
  ensure_work(*g, (long long)rows);
  infer(*g, (const float*)data, (long long)rows);
  VLHIP_LAUNCH("gate_bf16_to_f32", gate_bf16_to_f32_kernel, dim3(grid1((long long)rows * g->K)), dim3(256), 0, g->Zi,
               (long long)rows, g->K, g->Kq, g->Z);
  download(logits_host, g->Z, rows * g->K * sizeof(float));
}

}  // extern "C"
