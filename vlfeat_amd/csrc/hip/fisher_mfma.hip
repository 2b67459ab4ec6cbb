// fisher_mfma.hip -- Fisher encoder on CDNA4 matrix cores (gfx950), with
// compensated split-16 arithmetic.
//
// The per-component Mahalanobis log-likelihood of fisher.c:363-488 (via the
// E-step of gmm.c:758-791) is rewritten as one dense contraction over the
// 2D features [x'^2, x'] (x' = x - g, g the prior-weighted mean of the means):
//     logit_ik = b_k + sum_d w2_kd x'_id^2 + sum_d w1_kd x'_id,
//     w2 = -1/2 sigma^-2,  w1 = sigma^-2 mu',  b = log pi - D/2 log 2pi
//          - 1/2 sum log sigma^2 - 1/2 sum sigma^-2 mu'^2   (FP64)
// and the Fisher moments as a second contraction of the posteriors with
// [x', x'^2]:  S0 = sum q,  S1 = sum q x',  S2 = sum q x'^2, giving
//     u = sigma^-1 (S1 - mu' S0) / (n sqrt pi)
//     v = (sigma^-2 (S2 - 2 mu' S1 + mu'^2 S0) - S0) / (n sqrt(2 pi)).
// Operands are 16-bit (bf16 or fp16, template H) split as v = hi + lo
// (hi = round(v), lo = round(v - hi)); the logits use the three products
// hi*hi + lo*hi + hi*lo, the moments hi and lo parts of x' and x'^2 against
// the fp16/bf16-rounded posterior (the same rounded value enters S0), all
// accumulated in FP32 by v_mfma_f32_16x16x32_{bf16,f16}.  fp16 operands are
// scaled by powers of two chosen from the mixture's range so that hi and lo
// parts stay normal.
//
// Kernels: mfma_center / mfma_tables (scales and split weights),
// mfma_lse (first pass: split operand records and row statistics per group of
// 32 descriptors), mfma_fisher (second pass: workgroup per image x component
// block: logits -> posteriors -> moments -> finalised slice of the FV).
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <type_traits>
#include <vector>

namespace vlhip {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
typedef _Float16 f16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

template <typename H> struct VecOf;
template <> struct VecOf<_Float16> { typedef f16x8 type; };
template <> struct VecOf<__bf16> { typedef bf16x8 type; };

__device__ __forceinline__ f32x4 mfma16(const f16x8& a, const f16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_f16(a, b, c, 0, 0, 0);
}
__device__ __forceinline__ f32x4 mfma16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

constexpr double kLog2eD = 1.4426950408889634;  // log2(e)

// Power-of-two operand scales and the feature centre (device resident).
struct MfmaScales {
  float s1, s2;     // x' and x'^2 multipliers
  float sq;         // posterior multiplier
  unsigned bound;   // bits of the (non-negative) range bound on |x'|, max-reduced atomically
  float shift;      // logit origin of the fused encoder (mfma_shift_kernel)
};

// Rows per group of the matrix-core passes and components per staged block
// of the first pass.
constexpr int GR = 32;
constexpr int MB_COMPS = 64;

// ---------------------------------------------------------------------------
// Tables.  Kp = K rounded up to 64, Dp = D rounded up to 16.
//   Whi/Wlo [Kp][2Dp]   features [0,Dp) = w2 * 2^-e2, [Dp,2Dp) = w1 * 2^-e1
//   Wst     [Kp/64][hi, lo][64][2Dp + 8]  the same split weights in the first
//                       pass's LDS layout (rows padded by 16 bytes of zeros),
//                       so that a block is one contiguous copy
//   bias    [Kp]        b_k (FP32), -inf for skipped / padding components
//   mup, isg, sinv [Kp][Dp]  mu', sigma^-1, sigma^-2 (FP32) for finalisation

// g_d = sum_k pi_k mu_kd / sum_k pi_k (FP64 sums), one workgroup per dimension
// (zero for the padding dimensions D <= d < Dp).
template <typename T>
__global__ __launch_bounds__(256) void mfma_center_kernel(const T* __restrict__ means, const T* __restrict__ priors,
                                                          int D, int K, float* __restrict__ g) {
  __shared__ double red[256];
  const int d = blockIdx.x, t = threadIdx.x;
  double a = 0.0, w = 0.0;
  if (d < D) {
    for (int k = t; k < K; k += 256) {
      a += (double)priors[k] * (double)means[(size_t)k * D + d];
      w += (double)priors[k];
    }
  }
  const double sa = block_sum_256(a, red);
  const double sw = block_sum_256(w, red);
  if (t == 0) g[d] = (d < D && sw > 0) ? (float)(sa / sw) : 0.f;
}

// Range bound on |x'|: max over components and dimensions of |mu' | + 6 sigma,
// max-reduced into sc->bound (zeroed beforehand; a non-negative float's bits
// order as unsigned integers).
template <typename T>
__global__ __launch_bounds__(256) void mfma_bound_kernel(const T* __restrict__ means, const T* __restrict__ covs,
                                                         int D, int K, const float* __restrict__ g,
                                                         MfmaScales* __restrict__ sc) {
  __shared__ float red[256];
  const int t = threadIdx.x;
  const size_t i = (size_t)blockIdx.x * 256 + t;
  float v = 0.f;
  if (i < (size_t)K * D) {
    const int d = (int)(i % D);
    const double x = fabs((double)means[i] - (double)g[d]) + 6.0 * sqrt((double)covs[i]);
    v = x > 0.0 ? (float)x : 0.f;  // NaN (a negative variance) does not raise the bound
  }
  red[t] = v;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (t < o) red[t] = fmaxf(red[t], red[t + o]);
    __syncthreads();
  }
  if (t == 0 && red[0] > 0.f) atomicMax(&sc->bound, __float_as_uint(red[0]));
}

// Scales from the bound: FP16 puts |x'| and |x'|^2 near 2^8 and the posterior
// (<= 1) at 2^14; BF16 needs none.
__device__ __forceinline__ MfmaScales mfma_scales(unsigned boundBits, bool isHalf) {
  MfmaScales s;
  const double b = (double)__uint_as_float(boundBits);
  const double xm = b > 1e-30 ? b : 1.0;
  if (isHalf) {
    s.s1 = (float)exp2(floor(log2(256.0 / xm)));
    s.s2 = (float)exp2(floor(log2(256.0 / (xm * xm))));
    s.sq = 16384.0f;
  } else {
    s.s1 = 1.0f;
    s.s2 = 1.0f;
    s.sq = 1.0f;
  }
  s.bound = boundBits;
  return s;
}

// One wave per component: lane l handles dimensions l and l + 64 (Dp <= 128);
// the log-determinant and the quadratic term of the bias are wave sums in FP64
// (fixed butterfly order).
template <typename T, typename H>
__global__ __launch_bounds__(64) void mfma_tables_kernel(const T* __restrict__ means, const T* __restrict__ covs,
                                                         const T* __restrict__ priors, int D, int Dp, int K, int Kp,
                                                         const float* __restrict__ g, MfmaScales* __restrict__ scp,
                                                         H* __restrict__ Whi, H* __restrict__ Wlo,
                                                         H* __restrict__ Wst, float* __restrict__ bias,
                                                         float* __restrict__ mup, float* __restrict__ isg,
                                                         float* __restrict__ sinv) {
  const int k = blockIdx.x, lane = threadIdx.x;
  // the bound does not change here (the scales are stored with the same bits)
  const MfmaScales sc = mfma_scales(scp->bound, std::is_same<H, _Float16>::value);
  if (k == 0 && lane == 0) *scp = sc;
  if (k >= Kp) return;
  const float e1 = 1.0f / sc.s1, e2 = 1.0f / sc.s2;
  H* wh = Whi + (size_t)k * 2 * Dp;
  H* wl = Wlo + (size_t)k * 2 * Dp;
  const int WS = 2 * Dp + 8;
  H* sh = Wst + ((size_t)(k / MB_COMPS) * 2 * MB_COMPS + (size_t)(k % MB_COMPS)) * WS;
  H* sl = sh + (size_t)MB_COMPS * WS;
  if (lane < 8) sh[2 * Dp + lane] = sl[2 * Dp + lane] = H(0.f);
  const bool real = k < K;
  double logdet = 0.0, quad = 0.0;
  for (int d = lane; d < Dp; d += 64) {
    float w2 = 0.f, w1 = 0.f, mu = 0.f, is = 0.f, si = 0.f;
    if (real && d < D) {
      const double c = (double)covs[(size_t)k * D + d];
      const double s = 1.0 / c;
      const double m = (double)means[(size_t)k * D + d] - (double)g[d];
      logdet += log(c);
      quad += s * m * m;
      w2 = (float)(-0.5 * s * kLog2eD) * e2;
      w1 = (float)(s * m * kLog2eD) * e1;
      mu = (float)m;
      is = (float)sqrt(s);
      si = (float)s;
    }
    const H h2 = H(w2), h1 = H(w1);
    const H l2 = H(w2 - (float)h2), l1 = H(w1 - (float)h1);
    wh[d] = sh[d] = h2;
    wl[d] = sl[d] = l2;
    wh[Dp + d] = sh[Dp + d] = h1;
    wl[Dp + d] = sl[Dp + d] = l1;
    mup[(size_t)k * Dp + d] = mu;
    isg[(size_t)k * Dp + d] = is;
    sinv[(size_t)k * Dp + d] = si;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) {
    logdet += __shfl_xor(logdet, o, 64);
    quad += __shfl_xor(quad, o, 64);
  }
  if (lane == 0) {
    if (!real) {
      bias[k] = -INFINITY;
    } else {
      const double p = (double)priors[k];
      const double logw = (p < 1e-6) ? -(double)INFINITY : log(p);
      bias[k] = (float)((logw - 0.5 * (double)D * kLog2Pi - 0.5 * logdet - 0.5 * quad) * kLog2eD);
    }
  }
}

// Split v into two 16-bit parts.
template <typename H>
__device__ __forceinline__ void split16(float v, H& hi, H& lo) {
  hi = H(v);
  lo = H(v - (float)hi);
}

// A fragment (16 rows x 32 features) of the logit contraction for one lane:
// row r, features f0..f0+7 of [x'^2 * s2, x' * s1] built from 8 consecutive
// fp32 centred coordinates.
template <typename H>
__device__ __forceinline__ void make_frag(const float (&xc)[8], bool square, float s1, float s2,
                                          typename VecOf<H>::type& hi, typename VecOf<H>::type& lo) {
#pragma unroll
  for (int v = 0; v < 8; ++v) {
    const float val = square ? xc[v] * xc[v] * s2 : xc[v] * s1;
    H h, l;
    split16<H>(val, h, l);
    hi[v] = h;
    lo[v] = l;
  }
}

// ---------------------------------------------------------------------------
// Group records.  Each image's descriptors are processed in groups of GR = 32
// consecutive rows starting at the image's first row (the last group of an
// image is padded), so an image's result does not depend on the other images
// of the batch.  Image i owns groups gstart[i] .. gstart[i+1]-1.  The first
// pass writes,
// per group, the split 16-bit operands in the register layout of the second
// pass's matrix instructions, and the row statistics:
//   A part  [t 2][j NS][hl 2][lane 64] x 8 halves: logit A fragments; lane l
//           holds row 16t + (l & 15), features 32j + 8(l >> 4) .. +7 of
//           [x'^2 s2 | x' s1], hi or lo part;
//   B part  [p 4][ft NS][lane 64] x 8 halves: moment B fragments; plane p of
//           (x' s1 hi, x' s1 lo, x'^2 s2 hi, x'^2 s2 lo), lane l holds feature
//           16ft + (l & 15) of rows rho(l >> 4, v), v = 0..7, where
//           rho(lg, v) = 16(v >> 2) + 4 lg + (v & 3);
//   stats   [3][32]: -m, sq / sum_k 2^(logit_k - m) (0 for the padding rows
//           after an image's last descriptor), argmax (-1 for padding rows;
//           written only when FAST encodings are requested).  m is the row's
//           largest logit with FAST encodings, otherwise a reference at most
//           64 above it (LseLazy); either gives the posteriors
//           2^(logit - m) * sq / sum.
// Logits are in log2 units (the weight tables carry the factor log2 e).
// rho is the order in which a lane holds rows of the logit accumulators
// (C layout: rows 4(l >> 4) + r of each 16-row tile), so the posteriors become
// A fragments of the moment contraction without leaving registers; the sum
// over rows does not depend on the order, only on A and B agreeing on it.
// groups (waves) per first-pass workgroup: 8, or 4 when D > 80 (registers)
template <int NS> struct LseGroups { static constexpr int value = NS <= 5 ? 8 : 4; };

template <int NS>
struct GroupRec {
  static constexpr size_t kA = (size_t)2 * NS * 2 * 64 * 16;
  static constexpr size_t kB = (size_t)4 * NS * 64 * 16;
  static constexpr size_t kStats = (size_t)3 * GR * 4;
  static constexpr size_t kBytes = kA + kB + kStats;
};

// Deterministic block sum for WAVES x 64 threads (fixed order).
template <int WAVES>
__device__ __forceinline__ double block_sum_waves(double v, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < WAVES; ++i) r += red[i];
  __syncthreads();
  return r;
}


// Group table of a batch: group gstart[i] - gstart[imgBegin] + k of image i
// starts at batch-relative row offsets[i] - offsets[imgBegin] + 32k, its image
// ends at offsets[i+1] - offsets[imgBegin]; entries {first row, image end,
// image index, 0}.  One thread per image.
__global__ void mfma_group_table_kernel(const uint64_t* __restrict__ offsets, const long long* __restrict__ gstart,
                                        int imgBegin, int numImages, int4* __restrict__ tab) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= numImages) return;
  const int img = imgBegin + i;
  const long long rb = (long long)offsets[imgBegin];
  const int r0 = (int)((long long)offsets[img] - rb), r1 = (int)((long long)offsets[img + 1] - rb);
  const long long g0 = gstart[img] - gstart[imgBegin];
  const int ng = (int)(gstart[img + 1] - gstart[img]);
  for (int k = 0; k < ng; ++k) tab[g0 + k] = make_int4(r0 + GR * k, r1, img, 0);
}

// Running log-sum-exp (log2 units) and first maximum of the lane's 8 rows.
struct LseState {
  float m[2][4], s[2][4], bv[2][4];
  int bi[2][4];
};

// Logits of 32 rows x 64 components from one block of staged weights (layout
// [hi, lo][comp][2Dp + 8]): 2 row tiles x 4 component tiles x NS feature
// steps x 3 products, eight independent accumulators between dependent matrix
// ops; the first product of each accumulator adds init[c] (the components'
// biases, or zero).  The B fragments of step j + 1 are read into the
// registers of step j as soon as its products are done with them (the hi part
// after 16 products, the lo part after 24), so only the first step of a block
// waits for LDS.  vwork(j) is vector work interleaved one to one with step j's
// products; between(j) runs after step j (it may branch).
template <typename H, int NS, typename VW, typename BW>
__device__ __forceinline__ void lse_products(f32x4 (&acc)[4][2], const typename VecOf<H>::type (&ahi)[2][NS],
                                             const typename VecOf<H>::type (&alo)[2][NS], const H* w,
                                             const f32x4 (&init)[4], int lr, int lg, VW&& vwork, BW&& between) {
  typedef typename VecOf<H>::type V8;
  constexpr int WS = 2 * NS * 16 + 8;
  const H* wh = w + lr * WS + 8 * lg;
  const H* wl = wh + MB_COMPS * WS;
  V8 bh[4], bl[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    bh[c] = *reinterpret_cast<const V8*>(wh + 16 * c * WS);
    bl[c] = *reinterpret_cast<const V8*>(wl + 16 * c * WS);
  }
#pragma unroll
  for (int j = 0; j < NS; ++j) {
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c][t] = mfma16(ahi[t][j], bh[c], j == 0 ? init[c] : acc[c][t]);
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c][t] = mfma16(alo[t][j], bh[c], acc[c][t]);
    if (j + 1 < NS) {
#pragma unroll
      for (int c = 0; c < 4; ++c) bh[c] = *reinterpret_cast<const V8*>(wh + 16 * c * WS + 32 * (j + 1));
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c][t] = mfma16(ahi[t][j], bl[c], acc[c][t]);
    if (j + 1 < NS) {
#pragma unroll
      for (int c = 0; c < 4; ++c) bl[c] = *reinterpret_cast<const V8*>(wl + 16 * c * WS + 32 * (j + 1));
    }
    vwork(j);
    if (j == 0) __builtin_amdgcn_sched_group_barrier(0x100, 12, 0);  // the block's first fragment and bias reads
#pragma unroll
    for (int i = 0; i < 16; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // one matrix op
      __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);  // one vector op
    }
    if (j + 1 < NS) __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);  // next step's hi fragments
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
      __builtin_amdgcn_sched_group_barrier(0x002, 1, 0);
    }
    if (j + 1 < NS) __builtin_amdgcn_sched_group_barrier(0x100, 4, 0);  // next step's lo fragments
    __builtin_amdgcn_sched_barrier(0);
    between(j);
  }
}

// Running log-sum-exp of a lane's 8 rows against a lazily moved reference
// (log2 units): s = sum of 2^(logit - ref) over the components seen so far;
// ref is the same in the 16 lanes that share a row and never below the row's
// largest logit, so every term is at most 1 and a block costs one
// subtraction, one exponential and one addition per logit, plus a check.
// When a block lifts a row's largest logit above ref, the row's ref moves to
// that logit + 64 (its term is then 2^-64, far from underflow; terms that
// underflow are more than 62 binades below it) and s is rescaled.  ref starts
// at -FLT_MAX, so the first finite logit of a row moves it.  The row
// statistics are then -ref and sq / sum, in place of -max and sq / sum over
// 2^(logit - max): the same posteriors 2^(logit - ref) * sq / sum.
struct LseLazy {
  float ref[2][4], s[2][4];
};
constexpr float kLseHeadroom = 64.f;

__device__ __forceinline__ void lse_lazy_init(LseLazy& st) {
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) { st.ref[t][r] = -3.402823466e38f; st.s[t][r] = 0.f; }
}
__device__ __forceinline__ float lse_block_max(const f32x4 (&a)[4][2], int t, int r) {
  return fmaxf(fmaxf(a[0][t][r], a[1][t][r]), fmaxf(a[2][t][r], a[3][t][r]));
}
// True when a block's logits a (bias included) lift a row of this lane above
// its reference (plain comparisons: a maximum would first quiet its operands).
__device__ __forceinline__ bool lse_lazy_check(const f32x4 (&a)[4][2], const LseLazy& st) {
  bool up = false;
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float ref = st.ref[t][r];
      up = up | (a[0][t][r] > ref) | (a[1][t][r] > ref) | (a[2][t][r] > ref) | (a[3][t][r] > ref);
    }
  return up;
}
// Moves the reference of every row whose largest logit in block a rose above
// it; all 64 lanes take part (row maxima across lanes).
__device__ __forceinline__ void lse_lazy_rescale(const f32x4 (&a)[4][2], LseLazy& st) {
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float rm = row16_max(lse_block_max(a, t, r));
      const float nr = rm > st.ref[t][r] ? rm + kLseHeadroom : st.ref[t][r];
      st.s[t][r] *= __builtin_amdgcn_exp2f(st.ref[t][r] - nr);
      st.ref[t][r] = nr;
    }
}
// The terms of row (t, r) of a block.
__device__ __forceinline__ void lse_lazy_row(const f32x4 (&a)[4][2], int t, int r, LseLazy& st) {
  const float ref = st.ref[t][r];
  st.s[t][r] += (__builtin_amdgcn_exp2f(a[0][t][r] - ref) + __builtin_amdgcn_exp2f(a[1][t][r] - ref)) +
                (__builtin_amdgcn_exp2f(a[2][t][r] - ref) + __builtin_amdgcn_exp2f(a[3][t][r] - ref));
}
// Whole update with a block's logits (the last block of a sweep).
__device__ __forceinline__ void lse_lazy_update(const f32x4 (&a)[4][2], LseLazy& st) {
  if (__ballot(lse_lazy_check(a, st)) != 0) lse_lazy_rescale(a, st);
#pragma unroll
  for (int q = 0; q < 8; ++q) lse_lazy_row(a, q >> 2, q & 3, st);
}

// Online log-sum-exp update with the logits of components cb .. cb + 63
// (lane holds components cb + 16c + lr; bblk holds their 64 biases, staged in
// LDS); with ARG also the first maximum.
template <bool ARG>
__device__ __forceinline__ void lse_update(const f32x4 (&acc)[4][2], const float* bblk, int cb, int lr,
                                           LseState& st) {
  float bc[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) bc[c] = bblk[16 * c + lr];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float lv[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        lv[c] = acc[c][t][r] + bc[c];
        if (ARG && lv[c] > st.bv[t][r]) { st.bv[t][r] = lv[c]; st.bi[t][r] = cb + 16 * c + lr; }
      }
      const float mx = fmaxf(fmaxf(lv[0], lv[1]), fmaxf(lv[2], lv[3]));
      const float mn = fmaxf(st.m[t][r], mx);
      const float ms = (mn == -INFINITY) ? 0.f : mn;
      st.s[t][r] = st.s[t][r] * __builtin_amdgcn_exp2f(st.m[t][r] - ms)
                 + ((__builtin_amdgcn_exp2f(lv[0] - ms) + __builtin_amdgcn_exp2f(lv[1] - ms))
                 + (__builtin_amdgcn_exp2f(lv[2] - ms) + __builtin_amdgcn_exp2f(lv[3] - ms)));
      st.m[t][r] = mn;
    }
}

// Log-sum-exp update of one row (t, r) with the logits acc of components
// cb + 16c + lr (bias bc already loaded).
template <bool ARG>
__device__ __forceinline__ void lse_row(const f32x4 (&acc)[4][2], const float (&bc)[4], int t, int r, int cb, int lr,
                                        LseState& st) {
  float lv[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    lv[c] = acc[c][t][r] + bc[c];
    if (ARG && lv[c] > st.bv[t][r]) { st.bv[t][r] = lv[c]; st.bi[t][r] = cb + 16 * c + lr; }
  }
  const float mx = fmaxf(fmaxf(lv[0], lv[1]), fmaxf(lv[2], lv[3]));
  const float mn = fmaxf(st.m[t][r], mx);
  const float ms = (mn == -INFINITY) ? 0.f : mn;
  st.s[t][r] = st.s[t][r] * __builtin_amdgcn_exp2f(st.m[t][r] - ms)
             + ((__builtin_amdgcn_exp2f(lv[0] - ms) + __builtin_amdgcn_exp2f(lv[1] - ms))
             + (__builtin_amdgcn_exp2f(lv[2] - ms) + __builtin_amdgcn_exp2f(lv[3] - ms)));
  st.m[t][r] = mn;
}

// ---------------------------------------------------------------------------
// First pass.  Workgroup: 8 waves (4 when D > 80), wave = one group of 32 rows.  The wave
// builds its A fragments from the fp32 descriptors, stores them and the B
// fragments in the group record, then sweeps all components: the split
// weights of 64 components at a time are double-buffered in LDS (the next
// block is copied straight into LDS while the current one is multiplied), and
// every wave runs 2 row tiles x 4 component tiles x NS steps x 3 products per
// block.  Without FAST encodings the accumulators start at the components'
// biases and each row keeps a log-sum-exp against a lazily moved reference
// (LseLazy); with them, a lane-local online log-sum-exp and the first index
// of the maximum.  Both are merged across the 16 lanes that share rows at the
// end.  The update with a block is issued with the products of the next.
template <typename H, int NS, bool ARG>
__global__ __launch_bounds__(LseGroups<NS>::value * 64, 1) void mfma_lse_kernel(
    const float* __restrict__ x, int D, int Kp, const H* __restrict__ Wst, const float* __restrict__ bias,
    const MfmaScales* __restrict__ scp, const float* __restrict__ g,
    unsigned char* __restrict__ rec, long long numGroups, const int4* __restrict__ gtab,
    const unsigned char* __restrict__ gflag) {
  typedef typename VecOf<H>::type V8;
  typedef GroupRec<NS> R;
  constexpr int Dp = NS * 16;
  constexpr int WS = 2 * Dp + 8;                 // LDS row stride (halves) of the staged weights
  constexpr int LG = LseGroups<NS>::value;
  constexpr int NT = LG * 64;
  // The two weight buffers ([hi, lo][comp][feature]) are separate LDS
  // objects, used by every step in a parity fixed at compile time, so that
  // the compiler sees that a step's fragment reads do not alias the copy of
  // the next block into the other buffer and does not wait for that copy
  // before the step's products.
  __shared__ __align__(16) H wsh0[2][MB_COMPS][WS];
  __shared__ __align__(16) H wsh1[2][MB_COMPS][WS];
  // Biases of the staged blocks (block B in slot B & 3), four copies each,
  // loaded with the block's copy and stored at the end of the step before it,
  // so that no global load is waited for inside the sweep.
  __shared__ __align__(16) f32x4 bsh[4][MB_COMPS];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const long long grp = (long long)blockIdx.x * LG + w;  // relative to the batch
  const bool live = grp < numGroups;  // uniform per wave; dead waves still take part in barriers
  // batch-relative first row of the group and end of its image
  long long row0 = 0, rowEnd = 0;
  if (live) {
    const int4 gr = gtab[grp];
    row0 = gr.x;
    rowEnd = gr.y;
  }
  const float s1 = scp->s1, s2 = scp->s2;
  V8* Ap = reinterpret_cast<V8*>(rec + (size_t)grp * R::kBytes);
  V8* Bp = reinterpret_cast<V8*>(rec + (size_t)grp * R::kBytes + R::kA);
  float* Sf = reinterpret_cast<float*>(rec + (size_t)grp * R::kBytes + R::kA + R::kB);

  // Weight staging: a block of Wst (its LDS layout [hi, lo][comp][2Dp + 8])
  // is copied in 1 KiB pieces, wave w taking pieces w, w + LG, ... and lane l
  // bytes 16l .. 16l + 15 of each.  Blocks after the first are copied straight
  // into LDS (global_load_lds, 16 bytes per lane) a whole block ahead, through
  // no registers.
  constexpr size_t BLOCK_BYTES = sizeof(H) * 2 * MB_COMPS * WS;
  constexpr int PIECES = (int)(BLOCK_BYTES / 1024);
  constexpr int PW = (PIECES + LG - 1) / LG;       // pieces per wave, at most
  static_assert(PIECES * 1024 == BLOCK_BYTES, "pieces must tile a block");
  const unsigned voff = 1024u * (unsigned)w + 16u * (unsigned)lane;
  // block 0 goes through registers, loaded now and stored once the staging
  // area below is free
  u32x4 w0[PW];
#pragma unroll
  for (int k = 0; k < PW; ++k)
    if (w + LG * k < PIECES)
      w0[k] = *reinterpret_cast<const u32x4*>(reinterpret_cast<const char*>(Wst) + voff + 1024u * LG * k);
  float bnext = tid < MB_COMPS ? bias[tid] : 0.f;

  // Stage the wave's 32 centred descriptors x' = x - g (zero past D and past
  // the image) in LDS with 16-byte loads; the weight buffers are free until
  // the sweep starts, and each holds the rows of half of the waves (32 rows x
  // (Dp + 4) floats per wave).  Both operand layouts are then read from LDS.
  constexpr int XS = Dp + 4;
  constexpr int HALF = (LG + 1) / 2;  // waves staging in each weight buffer
  static_assert(HALF * GR * XS * sizeof(float) <= sizeof(wsh0), "staging must fit the weight buffers");
  const bool vecx = (D & 3) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  float* xsh = reinterpret_cast<float*>(w < HALF ? &wsh0[0][0][0] : &wsh1[0][0][0]) + (w % HALF) * GR * XS;
  {
    // every load of the lane issued before the first is used
    constexpr int XW = GR * (Dp / 4);  // 16-byte words of the wave's rows
    constexpr int XN = (XW + 63) / 64;
    // (the choice of load width is uniform and taken outside the loop: one
    // load per word, so that no load waits for an earlier one)
    f32x4 xv[XN], gv[XN];
#pragma unroll
    for (int u = 0; u < XN; ++u) {
      const int i = lane + 64 * u;
      const int c4 = (i - (i / (Dp / 4)) * (Dp / 4)) * 4;
      gv[u] = i < XW ? *reinterpret_cast<const f32x4*>(g + c4) : f32x4{0.f, 0.f, 0.f, 0.f};  // g: Dp entries, zero past D
    }
    if (vecx) {
#pragma unroll
      for (int u = 0; u < XN; ++u) {
        const int i = lane + 64 * u;
        const int r = i / (Dp / 4), c4 = (i - r * (Dp / 4)) * 4;
        const long long row = row0 + r;
        f32x4 v = f32x4{0.f, 0.f, 0.f, 0.f};
        if (i < XW && row < rowEnd && c4 < D) v = *reinterpret_cast<const f32x4*>(x + row * D + c4);
        xv[u] = v;
      }
    } else {
#pragma unroll
      for (int u = 0; u < XN; ++u) {
        const int i = lane + 64 * u;
        const int r = i / (Dp / 4), c4 = (i - r * (Dp / 4)) * 4;
        const long long row = row0 + r;
        f32x4 v = f32x4{0.f, 0.f, 0.f, 0.f};
        if (i < XW && row < rowEnd) {
#pragma unroll
          for (int e = 0; e < 4; ++e) v[e] = c4 + e < D ? x[row * D + c4 + e] : 0.f;
        }
        xv[u] = v;
      }
    }
#pragma unroll
    for (int u = 0; u < XN; ++u) {
      const int i = lane + 64 * u;
      if (i >= XW) continue;
      const int r = i / (Dp / 4), c4 = (i - r * (Dp / 4)) * 4;
      f32x4 v = xv[u];
      if (row0 + r < rowEnd) v = v - gv[u];
      *reinterpret_cast<f32x4*>(xsh + r * XS + c4) = v;
    }
  }
  __syncthreads();
  // A fragments of this wave's 32 rows, resident for the whole sweep
  V8 ahi[2][NS], alo[2][NS];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    const int f0 = 32 * j + 8 * lg;
    const bool sq = f0 < Dp;
    const int d0 = sq ? f0 : f0 - Dp;
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const f32x4 a = *reinterpret_cast<const f32x4*>(xsh + (16 * t + lr) * XS + d0);
      const f32x4 b = *reinterpret_cast<const f32x4*>(xsh + (16 * t + lr) * XS + d0 + 4);
      float xc[8];
#pragma unroll
      for (int v = 0; v < 4; ++v) { xc[v] = a[v]; xc[4 + v] = b[v]; }
      make_frag<H>(xc, sq, s1, s2, ahi[t][j], alo[t][j]);
      if (live) {
        Ap[((t * NS + j) * 2 + 0) * 64 + lane] = ahi[t][j];
        Ap[((t * NS + j) * 2 + 1) * 64 + lane] = alo[t][j];
      }
    }
  }
  // B fragments: feature 16ft + lr of rows rho(lg, v)
  if (live) {
#pragma unroll
    for (int ft = 0; ft < NS; ++ft) {
      const int d = 16 * ft + lr;
      V8 b1h, b1l, b2h, b2l;
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const float xc = xsh[(16 * (v >> 2) + 4 * lg + (v & 3)) * XS + d];
        H h, l;
        split16<H>(xc * s1, h, l);
        b1h[v] = h;
        b1l[v] = l;
        split16<H>(xc * xc * s2, h, l);
        b2h[v] = h;
        b2l[v] = l;
      }
      Bp[(0 * NS + ft) * 64 + lane] = b1h;
      Bp[(1 * NS + ft) * 64 + lane] = b1l;
      Bp[(2 * NS + ft) * 64 + lane] = b2h;
      Bp[(3 * NS + ft) * 64 + lane] = b2l;
    }
  }
  __syncthreads();  // the staging area becomes the weight buffers

  // Software pipeline over the component blocks: the matrix products of block
  // b are issued together with the log-sum-exp update of block b - 1, so the
  // vector work overlaps the matrix cores; one barrier per block.
  // groups the gate serves sparsely need no row statistics; a workgroup whose
  // groups are all sparse skips the sweep
  const bool need = live && !(gflag != nullptr && gflag[grp] != 0);
  if (!__syncthreads_or(need)) return;
  const int nb = Kp / MB_COMPS;
  H* const buf0 = &wsh0[0][0][0];  // even blocks
  H* const buf1 = &wsh1[0][0][0];  // odd blocks
  // copy of block b into its buffer d (every wave takes part, needed or not)
  auto copy_block = [&](int b, H* d) __attribute__((always_inline)) {
    const char* src = reinterpret_cast<const char*>(Wst) + (size_t)b * BLOCK_BYTES;
    unsigned char* dst = reinterpret_cast<unsigned char*>(d);
#pragma unroll
    for (int k = 0; k < PW; ++k)
      if (w + LG * k < PIECES)
        __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + voff + 1024u * LG * k),
                                         (__attribute__((address_space(3))) void*)(dst + 1024 * (w + LG * k)), 16, 0,
                                         0);
  };
  // biases of block b, four copies each (an accumulator's initial value)
  auto store_bias = [&](int b) __attribute__((always_inline)) {
    if (tid < MB_COMPS) bsh[b & 3][tid] = f32x4{bnext, bnext, bnext, bnext};
  };
#pragma unroll
  for (int k = 0; k < PW; ++k)
    if (w + LG * k < PIECES) reinterpret_cast<u32x4*>(buf0)[64 * (w + LG * k) + lane] = w0[k];
  store_bias(0);
  if (nb > 1) {
    copy_block(1, buf1);
    if (tid < MB_COMPS) bnext = bias[MB_COMPS + tid];
  }
  __syncthreads();
  f32x4 accA[4][2], accB[4][2];
  const f32x4 zero4[4] = {f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f}, f32x4{0.f, 0.f, 0.f, 0.f},
                          f32x4{0.f, 0.f, 0.f, 0.f}};
  auto no_work = [](int) __attribute__((always_inline)) {};
  // block b's initial accumulators: its biases (or zero with ARG, whose
  // update adds them, so that its first maximum compares the same values as
  // before)
  auto init_of = [&](int b, f32x4 (&bq)[4]) __attribute__((always_inline)) {
#pragma unroll
    for (int c = 0; c < 4; ++c) bq[c] = ARG ? zero4[c] : bsh[b & 3][16 * c + lr];
  };
  // end of step b: block b + 1 and its biases are in place for every wave
  auto finish_step = [&](int b) __attribute__((always_inline)) {
    if (b + 1 < nb) store_bias(b + 1);
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies have landed
    __syncthreads();
  };

  const float sq = scp->sq;
  if constexpr (!ARG) {
    LseLazy st;
    lse_lazy_init(st);
    // the update of block b - 1 with the products of block b: row maxima and
    // the reference check with step 0, the branch (rare) after it, then the
    // terms of the 8 rows spread over steps 1 .. NS - 1
    auto step = [&](f32x4 (&accNew)[4][2], f32x4 (&accOld)[4][2], int b, auto parity)
                    __attribute__((always_inline)) {
      constexpr bool odd = decltype(parity)::value != 0;  // b's parity
      if (b + 1 < nb) {
        copy_block(b + 1, odd ? buf0 : buf1);
        if (tid < MB_COMPS) bnext = bias[(b + 1) * MB_COMPS + tid];
      }
      if (need) {
        f32x4 bq[4];
        init_of(b, bq);
        bool up = false;
        lse_products<H, NS>(
            accNew, ahi, alo, odd ? buf1 : buf0, bq, lr, lg,
            [&](int j) __attribute__((always_inline)) {
              if (j == 0) {
                up = lse_lazy_check(accOld, st);
              } else {
#pragma unroll
                for (int q = (8 * (j - 1)) / (NS - 1 > 0 ? NS - 1 : 1); q < (8 * j) / (NS - 1 > 0 ? NS - 1 : 1); ++q)
                  lse_lazy_row(accOld, q >> 2, q & 3, st);
              }
            },
            [&](int j) __attribute__((always_inline)) {
              if (j == 0 && __ballot(up) != 0) lse_lazy_rescale(accOld, st);
            });
        if (NS == 1) {
#pragma unroll
          for (int q = 0; q < 8; ++q) lse_lazy_row(accOld, q >> 2, q & 3, st);
        }
      }
      finish_step(b);
    };
    if (need) {
      f32x4 bq[4];
      init_of(0, bq);
      lse_products<H, NS>(accA, ahi, alo, buf0, bq, lr, lg, no_work, no_work);
    }
    finish_step(0);
    int b = 1;
    for (; b + 1 < nb; b += 2) {
      step(accB, accA, b, std::integral_constant<int, 1>{});
      step(accA, accB, b + 1, std::integral_constant<int, 0>{});
    }
    if (b < nb) {
      step(accB, accA, b, std::integral_constant<int, 1>{});
      if (need) lse_lazy_update(accB, st);
    } else {
      if (need) lse_lazy_update(accA, st);
    }
    // sums over the 16 lanes that hold the same rows (one reference per row)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const float sum = row16_sum(st.s[t][r]);
        const int row = 16 * t + 4 * lg + r;
        if (need && lr == 0) {
          const bool pad = row0 + row >= rowEnd;
          Sf[row] = pad ? 0.f : -st.ref[t][r];
          Sf[GR + row] = pad ? 0.f : sq / sum;
        }
      }
  } else {
    LseState st;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) { st.m[t][r] = -INFINITY; st.s[t][r] = 0.f; st.bv[t][r] = -INFINITY; st.bi[t][r] = 0; }
    auto bias_of = [&](int b, float (&bc)[4]) __attribute__((always_inline)) {
#pragma unroll
      for (int c = 0; c < 4; ++c) bc[c] = bsh[b & 3][16 * c + lr][0];
    };
    // the update of block b - 1 (the online log-sum-exp and first maximum),
    // its 8 rows spread over the NS steps of block b's products
    auto step = [&](f32x4 (&accNew)[4][2], f32x4 (&accOld)[4][2], int b, auto parity)
                    __attribute__((always_inline)) {
      constexpr bool odd = decltype(parity)::value != 0;  // b's parity
      if (b + 1 < nb) {
        copy_block(b + 1, odd ? buf0 : buf1);
        if (tid < MB_COMPS) bnext = bias[(b + 1) * MB_COMPS + tid];
      }
      if (need) {
        float bc[4];
        bias_of(b - 1, bc);
        lse_products<H, NS>(
            accNew, ahi, alo, odd ? buf1 : buf0, zero4, lr, lg,
            [&](int j) __attribute__((always_inline)) {
#pragma unroll
              for (int q = (8 * j) / NS; q < (8 * (j + 1)) / NS; ++q)
                lse_row<true>(accOld, bc, q >> 2, q & 3, (b - 1) * MB_COMPS, lr, st);
            },
            no_work);
      }
      finish_step(b);
    };
    if (need) lse_products<H, NS>(accA, ahi, alo, buf0, zero4, lr, lg, no_work, no_work);
    finish_step(0);
    int b = 1;
    for (; b + 1 < nb; b += 2) {
      step(accB, accA, b, std::integral_constant<int, 1>{});
      step(accA, accB, b + 1, std::integral_constant<int, 0>{});
    }
    float bc[4];
    bias_of(nb - 1, bc);
    if (b < nb) {
      step(accB, accA, b, std::integral_constant<int, 1>{});
      if (need) {
#pragma unroll
        for (int q = 0; q < 8; ++q) lse_row<true>(accB, bc, q >> 2, q & 3, (nb - 1) * MB_COMPS, lr, st);
      }
    } else {
      if (need) {
#pragma unroll
        for (int q = 0; q < 8; ++q) lse_row<true>(accA, bc, q >> 2, q & 3, (nb - 1) * MB_COMPS, lr, st);
      }
    }
    // merge the 16 lanes that hold the same rows (lanes lg*16 .. lg*16+15)
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        float m = st.m[t][r];
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
        const float ms = (m == -INFINITY) ? 0.f : m;
        float s = st.s[t][r] * __builtin_amdgcn_exp2f(st.m[t][r] - ms);
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) s += __shfl_xor(s, o, 64);
        int bi = st.bi[t][r];
        float bv = st.bv[t][r];
#pragma unroll
        for (int o = 1; o < 16; o <<= 1) {
          const float ov = __shfl_xor(bv, o, 64);
          const int oi = __shfl_xor(bi, o, 64);
          if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
        }
        const int row = 16 * t + 4 * lg + r;
        if (need && lr == 0) {
          const bool pad = row0 + row >= rowEnd;
          Sf[row] = pad ? 0.f : -m;
          Sf[GR + row] = pad ? 0.f : sq / s;
          Sf[2 * GR + row] = __int_as_float(pad ? -1 : bi);
        }
      }
  }
}

// Gated groups.  One wave per group, after the first pass has written every
// group's exact row statistics (-m, sq / sum over all K components).
// For each row the exact logits of its M <= 16 candidate components (the same
// FP32 tables as the dense path: bias + log2 e * sum_d sigma^-2 (mu' x' - x'^2 / 2))
// give exact posteriors q = 2^(logit - m) / sum -- normalised over all
// components by the first pass, not over the candidates -- and their mass.
// The group is served sparsely only if every row passed the network's
// pre-screen (conf) and the candidates' exact mass of every row is >= need;
// with FAST encodings the largest candidate must also exceed 1 - mass, so that
// it is the row's true first maximum.  A served group lists {component, q * sq}
// in list[group][32][M] (q < 1e-6 dropped, fisher.c:436-441; padding rows:
// {-1, 0}) and counts its terms into numTerms of its image; the posterior mass
// outside the candidates, at most 1 - need per row, is dropped.
template <int NS>
__global__ __launch_bounds__(256) void mfma_sparse_kernel(
    const float* __restrict__ x, int D, long long numGroups, const int4* __restrict__ gtab,
    const unsigned char* __restrict__ conf, const int* __restrict__ cand, int M, const float* __restrict__ bias,
    const float* __restrict__ mup, const float* __restrict__ sinv, const float* __restrict__ g,
    const MfmaScales* __restrict__ scp, const unsigned char* __restrict__ rec, float need, int flags,
    unsigned char* __restrict__ gflag, int2* __restrict__ list, unsigned long long* __restrict__ numTerms) {
  typedef GroupRec<NS> R;
  constexpr int Dp = NS * 16;
  const int lane = threadIdx.x & 63;
  const long long grp = (long long)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (grp >= numGroups) return;
  const int4 gr = gtab[grp];
  const long long row0 = gr.x;
  const int nrow = min(GR, gr.y - gr.x);
  // the network's pre-screen: a row it does not expect to be served is not tried
  const bool rowOk = lane >= nrow || conf[row0 + lane] != 0;
  if (__all(rowOk) == 0) {
    if (lane == 0) gflag[grp] = 0;
    return;
  }
  const float* Sf = reinterpret_cast<const float*>(rec + (size_t)grp * R::kBytes + R::kA + R::kB);
  const bool fast = (flags & 4) != 0;
  const float sq = scp->sq;
  const float invsq = 1.0f / sq;
  const float qminS = __uint_as_float(0x358637beu) * sq;  // next float above float(1e-6), scaled
  const float kL2e = 1.4426950408889634f;
  const int j = lane & 15, sub = lane >> 4;
  float qv[GR / 4];  // scaled exact posterior q * sq of (row 4 it + sub, candidate j)
  int cv[GR / 4];
  bool ok = true;
#pragma unroll
  for (int it = 0; it < GR / 4; ++it) {
    const int r = 4 * it + sub;
    const bool valid = r < nrow && j < M;
    const int comp = valid ? cand[(row0 + r) * M + j] : -1;
    float v = 0.f;
    if (valid && comp >= 0) {
      const float* xr = x + (row0 + r) * D;
      const float* mu = mup + (size_t)comp * Dp;
      const float* si = sinv + (size_t)comp * Dp;
      float acc = 0.f;
      for (int d = 0; d < D; ++d) {
        const float xc = xr[d] - g[d];
        acc += si[d] * (mu[d] * xc - 0.5f * xc * xc);
      }
      // 2^(logit - max) * sq / sum: the row statistics of the first pass
      v = __builtin_amdgcn_exp2f(bias[comp] + kL2e * acc + Sf[r]) * Sf[GR + r];
    }
    qv[it] = v;
    cv[it] = valid ? comp : -1;
    // certified mass and largest posterior of the row (lanes 16 sub .. 16 sub + 15)
    float ms = v, mx = v;
#pragma unroll
    for (int o = 1; o < 16; o <<= 1) {
      ms += __shfl_xor(ms, o, 64);
      mx = fmaxf(mx, __shfl_xor(mx, o, 64));
    }
    const float mass = ms * invsq;
    const bool rowServed = r >= nrow || (mass >= need && (!fast || mx * invsq > 1.0f - mass));
    ok = ok && rowServed;
  }
  const bool served = __all(ok) != 0;
  if (lane == 0) gflag[grp] = served ? 1 : 0;
  if (!served) return;
  unsigned terms = 0;
#pragma unroll
  for (int it = 0; it < GR / 4; ++it) {
    const int r = 4 * it + sub;
    const int comp = cv[it];
    float v = qv[it];
    if (fast) {
      // one-hot on the first largest candidate (value, then lowest index)
      float mx = v;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) mx = fmaxf(mx, __shfl_xor(mx, o, 64));
      int best = (comp >= 0 && v == mx && v > 0.f) ? comp : 0x7fffffff;
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) best = min(best, __shfl_xor(best, o, 64));
      v = (comp >= 0 && comp == best) ? sq : 0.f;
    } else if (!(v >= qminS)) {
      v = 0.f;
    }
    if (comp >= 0 && v > 0.f) terms += 1;
    if (j < M) list[((size_t)grp * GR + r) * M + j] = make_int2(comp, __float_as_int(v));
  }
  if (numTerms) {
    unsigned t = terms;
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) t += __shfl_xor(t, o, 64);
    if (lane == 0 && t) atomicAdd(&numTerms[gr.z], (unsigned long long)t);
  }
}

// Block mapping of the second pass (component-block schedule).  Workgroups
// are dealt to the 8 XCDs round robin.
//   0  XCD-aware, image-major: all component blocks of one image land on the
//      same XCD, one after the other, and share its L2 copy of the image's
//      group records (the default);
//   1  linear, image-major (blocks of one image consecutive in launch order,
//      spread over the XCDs);
//   2  XCD-aware, component-major: an XCD walks its images for component
//      block 0, then block 1, ... (weights reused, records re-read).
// With fewer than 8 images every schedule is the linear one.  The grid is
// padImages x numKB (padImages: numImages rounded up to 8 from 8 images on);
// blocks past the last image return.
__device__ __forceinline__ void fisher_mfma_coords(int numKB, int numImages, int sched, int& img, int& kb) {
  const int b = blockIdx.x;
  if (numImages < 8 || sched == 1) {
    img = b / numKB;
    kb = b % numKB;
  } else if (sched == 2) {
    const int xcd = b & 7, slot = b >> 3, perX = (numImages + 7) / 8;
    kb = slot / perX;
    img = (slot - kb * perX) * 8 + xcd;
  } else {
    const int xcd = b & 7, slot = b >> 3;
    img = (slot / numKB) * 8 + xcd;
    kb = slot % numKB;
  }
}

// Copy of one group record between HBM, registers and LDS (NT threads, 16 bytes each).
template <int PER, int NT, int NV>
__device__ __forceinline__ void rec_load(u32x4 (&pre)[PER], const unsigned char* src) {
  const u32x4* s = reinterpret_cast<const u32x4*>(src);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int i = threadIdx.x + NT * q;
    pre[q] = s[i < NV ? i : NV - 1];  // unconditional, so that pre stays in registers
  }
}
template <int PER, int NT, int NV>
__device__ __forceinline__ void rec_store(const u32x4 (&pre)[PER], unsigned char* dst) {
  u32x4* d = reinterpret_cast<u32x4*>(dst);
#pragma unroll
  for (int q = 0; q < PER; ++q) {
    const int i = threadIdx.x + NT * q;
    if (i < NV) d[i] = pre[q];
  }
}

// Fisher-vector slice of one (image, component block) from the moment
// accumulators in registers (fisher.c:419-485 in FP64): lane l holds features
// 16(ft0 + ft) + (l & 15), ft < NF, of components kw + 16c + 4(l >> 4) + r.
// Writes u and v, the slice's squared norm into normpart[img * numKB + kb] and
// the workgroup's kept terms (per-wave lane-mask counts) into numTerms[img].
template <typename T, int NS, int CT, int WAVES, int NF = NS>
__device__ __forceinline__ void fisher_slice_finalize(
    const f32x4 (&macc0)[CT], const f32x4 (&macc)[CT][2][NF], const MfmaScales& sc, int kw, int lr, int lg, int tid,
    int lane, int img, int kb, int numKB, long long nImg, int D, int K, const float* __restrict__ mup,
    const float* __restrict__ isg, const float* __restrict__ sinv, const T* __restrict__ priors, int flags,
    T* __restrict__ enc, double* __restrict__ normpart, unsigned long long* __restrict__ numTerms, unsigned terms,
    double* red, int ft0 = 0) {
  constexpr int Dp = NS * 16;
  const double invq = 1.0 / (double)sc.sq;
  const double inv1 = invq / (double)sc.s1, inv2 = invq / (double)sc.s2;
  const bool sqrtFlag = (flags & 1) != 0;
  T* encU = enc + (size_t)img * 2 * K * D;
  T* encV = encU + (size_t)K * D;
  double nrm = 0.0;
#pragma unroll
  for (int c = 0; c < CT; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const double S0 = (double)macc0[c][r] * invq;
      const int k = kw + 16 * c + 4 * lg + r;
      const double prior = k < K ? (double)priors[k] : 0.0;
      const bool use = k < K && prior >= 1e-6;
      double up = 1.0, vp = 1.0;
      if (nImg > 0 && use) {
        up = 1.0 / ((double)nImg * sqrt(prior));
        vp = 1.0 / ((double)nImg * sqrt(2.0 * prior));
      }
#pragma unroll
      for (int ft = 0; ft < NF; ++ft) {
        const int d = 16 * (ft0 + ft) + lr;  // >= D past the last tile
        if (k >= K || d >= D) continue;
        T u = T(0), v = T(0);
        if (use) {
          const double S1 = (double)macc[c][0][ft][r] * inv1;
          const double S2 = (double)macc[c][1][ft][r] * inv2;
          const double mu = (double)mup[(size_t)k * Dp + d];
          const double uu = (double)isg[(size_t)k * Dp + d] * (S1 - mu * S0);
          const double vv = (double)sinv[(size_t)k * Dp + d] * (S2 - 2.0 * mu * S1 + mu * mu * S0) - S0;
          u = (T)(uu * up);
          v = (T)(vv * vp);
        }
        if (sqrtFlag) { u = signed_sqrt(u); v = signed_sqrt(v); }
        encU[(size_t)k * D + d] = u;
        encV[(size_t)k * D + d] = v;
        nrm += (double)u * (double)u + (double)v * (double)v;
      }
      // one (component tile, row) at a time: its loads are not hoisted above
      // the previous one's, which keeps the FP64 temporaries few (the fused
      // kernel calls this with its weights and accumulators live)
      __builtin_amdgcn_sched_barrier(0);
    }
  const double bn = block_sum_waves<WAVES>(nrm, red);
  if (tid == 0) normpart[(size_t)img * numKB + kb] = bn;
  if (numTerms) {
    // terms is the wave's count (lane masks), entered once per wave
    const double bt = block_sum_waves<WAVES>(lane == 0 ? (double)terms : 0.0, red);
    if (tid == 0) atomicAdd(&numTerms[img], (unsigned long long)bt);
  }
}

// ---------------------------------------------------------------------------
// Second pass.  Workgroup = (image, WAVES x 16CT components); wave w owns the
// CT component tiles 16CT w .. with their split weights in registers and
// walks the image's groups (records double-buffered in LDS, the next one
// prefetched into registers):
//   logits   2 row tiles x CT component tiles x NS steps x 3 products;
//   q        2^(logit - m) / sum, FAST one-hot, < 1e-6 dropped
//            (fisher.c:436-441), rows outside the image masked, rounded to
//            16 bits (scaled) -- the same rounded value enters S0;
//   moments  q (A, straight from the accumulators) x the 4 planes of B:
//            S1 += q x'_hi + q x'_lo,  S2 += q x'^2_hi + q x'^2_lo.
// The finalisation works on the accumulators in registers: lane l holds
// features 16ft + (l & 15) of components 16c + 4(l >> 4) + r.
//
// With EM the "images" are consecutive chunks of an EM data set and the
// kernel produces the M-step moments of gmm.c:1093-1149 instead of a Fisher
// vector: posteriors below emThr (the M-step threshold 0.01/K, scaled by sq)
// are dropped, the unthresholded mass is summed on the side (FP32 per group,
// FP64 across groups), and the chunk's raw moments about g,
//   [S0 K | mass K | sum q x' (K x D) | sum q x'^2 (K x D)]   (FP64),
// are written to emOut[chunk].
template <typename T, typename H, int NS, int CT, int WAVES, bool EM = false>
__global__ __launch_bounds__(WAVES * 64, 1) void mfma_fisher_kernel(
    const unsigned char* __restrict__ rec, const uint64_t* __restrict__ offsets, const long long* __restrict__ gstart,
    int imgBegin, int numImages, int D, int K, const H* __restrict__ Whi, const H* __restrict__ Wlo,
    const float* __restrict__ bias, const MfmaScales* __restrict__ scp, const float* __restrict__ mup,
    const float* __restrict__ isg, const float* __restrict__ sinv, const T* __restrict__ priors, int flags,
    T* __restrict__ enc, double* __restrict__ normpart, unsigned long long* __restrict__ numTerms, int numKB,
    const unsigned char* __restrict__ gflag, const int2* __restrict__ list, int M, float emThr,
    double* __restrict__ emOut, int sched) {
  typedef typename VecOf<H>::type V8;
  typedef GroupRec<NS> R;
  constexpr int Dp = NS * 16;
  constexpr int NT = WAVES * 64;
  constexpr int BC = WAVES * 16 * CT;
  constexpr int NV = (int)(R::kBytes / 16);  // 16-byte words per record
  constexpr int PER = (NV + NT - 1) / NT;
  extern __shared__ __align__(16) unsigned char smem_raw[];  // 2 records
  __shared__ double red[WAVES];

  int li, kb;
  fisher_mfma_coords(numKB, numImages, sched, li, kb);
  if (li >= numImages) return;
  const int img = imgBegin + li;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int kw = kb * BC + 16 * CT * w;  // first component of this wave
  const long long nImg = (long long)offsets[img + 1] - (long long)offsets[img];
  const long long gBase = gstart[imgBegin];
  const long long gB = gstart[img] - gBase, gE = gstart[img + 1] - gBase;  // this image's records
  const bool fast = !EM && (flags & 4) != 0;
  const MfmaScales sc = *scp;

  V8 whi[CT][NS], wlo[CT][NS];
  float bc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int comp = kw + 16 * c + lr;
    bc[c] = bias[comp];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      whi[c][j] = *reinterpret_cast<const V8*>(Whi + (size_t)comp * 2 * Dp + 32 * j + 8 * lg);
      wlo[c][j] = *reinterpret_cast<const V8*>(Wlo + (size_t)comp * 2 * Dp + 32 * j + 8 * lg);
    }
  }
  f32x4 macc[CT][2][NS];  // [comp tile][S1, S2][feature tile]
#pragma unroll
  for (int c = 0; c < CT; ++c)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int ft = 0; ft < NS; ++ft) macc[c][s][ft] = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 macc0[CT];  // posterior mass S0 (every column)
#pragma unroll
  for (int c = 0; c < CT; ++c) macc0[c] = f32x4{0.f, 0.f, 0.f, 0.f};
  V8 ones;
#pragma unroll
  for (int v = 0; v < 8; ++v) ones[v] = H(1.f);
  unsigned terms = 0;  // kept (row, component) pairs of this wave, counted from lane masks
  double massd[CT];  // EM: unthresholded posterior mass of component kw + 16c + lr
#pragma unroll
  for (int c = 0; c < CT; ++c) massd[c] = 0.0;
  // (double)q < 1e-6 for a float q is q < the next float above float(1e-6);
  // the posteriors are compared after the exact power-of-two scaling by sq
  const float qminS = EM ? emThr : __uint_as_float(0x358637beu) * sc.sq;  // 1.0000001e-6 * sq

  u32x4 pre[PER];
  if (gB < gE) {
    rec_load<PER, NT, NV>(pre, rec + (size_t)gB * R::kBytes);
    rec_store<PER, NT, NV>(pre, smem_raw);
    if (gB + 1 < gE) rec_load<PER, NT, NV>(pre, rec + (size_t)(gB + 1) * R::kBytes);
  }
  __syncthreads();

  for (long long gi = gB; gi < gE; ++gi) {
    const int buf = (int)((gi - gB) & 1);
    // read before the prefetch of group gi + 2 is issued, so that waiting for
    // the flag does not wait for that prefetch
    const bool gateServed = gflag != nullptr && gflag[gi] != 0;
    if (gi + 1 < gE) {
      // every wave finished reading this buffer (group gi - 1) before the last barrier
      rec_store<PER, NT, NV>(pre, smem_raw + (size_t)(buf ^ 1) * R::kBytes);
      if (gi + 2 < gE) rec_load<PER, NT, NV>(pre, rec + (size_t)(gi + 2) * R::kBytes);
    }
    const unsigned char* cur = smem_raw + (size_t)buf * R::kBytes;
    const V8* Ap = reinterpret_cast<const V8*>(cur);
    const V8* Bp = reinterpret_cast<const V8*>(cur + R::kA);
    V8 aq[CT];
    if (gateServed) {
      // gate-served group: scatter its listed posteriors into a rows x BC tile
      float* tile = reinterpret_cast<float*>(smem_raw + 2 * R::kBytes);
#pragma unroll
      for (int q = 0; q < (GR * BC / 4) / NT; ++q)
        reinterpret_cast<f32x4*>(tile)[tid + NT * q] = f32x4{0.f, 0.f, 0.f, 0.f};
      __syncthreads();
      const int2* L = list + (size_t)gi * GR * M;
      for (int e = tid; e < GR * M; e += NT) {
        const int2 en = L[e];
        const int c = en.x - kb * BC;
        if (c >= 0 && c < BC) tile[(e / M) * BC + c] = __int_as_float(en.y);
      }
      __syncthreads();
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < CT; ++c)
            aq[c][4 * t + r] = H(tile[(16 * t + 4 * lg + r) * BC + 16 * CT * w + 16 * c + lr]);
    } else {
      const float* Sf = reinterpret_cast<const float*>(cur + R::kA + R::kB);
      // rows 16t + 4lg + r of this lane: -m, sq / sum
      f32x4 nm4[2], rs4[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        nm4[t] = *reinterpret_cast<const f32x4*>(&Sf[16 * t + 4 * lg]);
        rs4[t] = *reinterpret_cast<const f32x4*>(&Sf[GR + 16 * t + 4 * lg]);
      }

      // ---- logits - m (the accumulators start at bias - m)
      f32x4 acc[2][CT];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c)
          acc[t][c] = f32x4{bc[c] + nm4[t][0], bc[c] + nm4[t][1], bc[c] + nm4[t][2], bc[c] + nm4[t][3]};
#pragma unroll
      for (int j = 0; j < NS; ++j) {
        V8 ah[2], al[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          ah[t] = Ap[((t * NS + j) * 2 + 0) * 64 + lane];
          al[t] = Ap[((t * NS + j) * 2 + 1) * 64 + lane];
        }
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int c = 0; c < CT; ++c) acc[t][c] = mfma16(ah[t], whi[c][j], acc[t][c]);
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int c = 0; c < CT; ++c) acc[t][c] = mfma16(al[t], whi[c][j], acc[t][c]);
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int c = 0; c < CT; ++c) acc[t][c] = mfma16(ah[t], wlo[c][j], acc[t][c]);
      }

      // ---- scaled posteriors q * sq, straight into A fragments of the moment
      // contraction (padding rows have sq / sum = 0 and FAST argmax -1)
      if (fast) {
        f32x4 ar4[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) ar4[t] = *reinterpret_cast<const f32x4*>(&Sf[2 * GR + 16 * t + 4 * lg]);
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < CT; ++c) {
              const bool hit = kw + 16 * c + lr == __float_as_int(ar4[t][r]) && bc[c] != -INFINITY;
              terms += (unsigned)__popcll(__ballot(hit));
              aq[c][4 * t + r] = hit ? H(sc.sq) : H(0.f);
            }
      } else {
        float msum[CT];
#pragma unroll
        for (int c = 0; c < CT; ++c) msum[c] = 0.f;
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r)
#pragma unroll
            for (int c = 0; c < CT; ++c) {
              const float v = __builtin_amdgcn_exp2f(acc[t][c][r]) * rs4[t][r];
              const bool keep = v >= qminS;
              terms += (unsigned)__popcll(__ballot(keep));
              // padding rows (rs4 = 0) may have 2^acc = inf: keep them out of the mass
              if (EM) msum[c] += rs4[t][r] > 0.f ? v : 0.f;
              aq[c][4 * t + r] = H(keep ? v : 0.f);
            }
        if (EM) {
#pragma unroll
          for (int c = 0; c < CT; ++c) massd[c] += (double)msum[c];
        }
      }
    }
    // posterior mass: every column of q x ones is the row sum
#pragma unroll
    for (int c = 0; c < CT; ++c) macc0[c] = mfma16(aq[c], ones, macc0[c]);

    // ---- moments
#pragma unroll
    for (int p = 0; p < 4; ++p)
#pragma unroll
      for (int ft = 0; ft < NS; ++ft) {
        const V8 bx = Bp[(p * NS + ft) * 64 + lane];
#pragma unroll
        for (int c = 0; c < CT; ++c) macc[c][p >> 1][ft] = mfma16(aq[c], bx, macc[c][p >> 1][ft]);
      }
    __syncthreads();
  }

  // ---- finalisation in registers
  const double invq = 1.0 / (double)sc.sq;
  const double inv1 = invq / (double)sc.s1, inv2 = invq / (double)sc.s2;
  if constexpr (EM) {
    const size_t KD = (size_t)K * D;
    double* P = emOut + (size_t)img * (2 * (size_t)K + 2 * KD);
#pragma unroll
    for (int c = 0; c < CT; ++c) {
      double m = massd[c];
      m += __shfl_xor(m, 16, 64);
      m += __shfl_xor(m, 32, 64);
      const int kl = kw + 16 * c + lr;
      if (lg == 0 && kl < K) P[K + kl] = m * invq;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = kw + 16 * c + 4 * lg + r;
        if (k >= K) continue;
        if (lr == 0) P[k] = (double)macc0[c][r] * invq;
#pragma unroll
        for (int ft = 0; ft < NS; ++ft) {
          const int d = 16 * ft + lr;
          if (d >= D) continue;
          P[2 * (size_t)K + (size_t)k * D + d] = (double)macc[c][0][ft][r] * inv1;
          P[2 * (size_t)K + KD + (size_t)k * D + d] = (double)macc[c][1][ft][r] * inv2;
        }
      }
    }
    return;
  }
  fisher_slice_finalize<T, NS, CT, WAVES>(macc0, macc, sc, kw, lr, lg, tid, lane, img, kb, numKB, nImg, D, K, mup,
                                          isg, sinv, priors, flags, enc, normpart, numTerms, terms, red);
}

// ---------------------------------------------------------------------------
// Pipelined second pass (Fisher vectors without the gate).  The arithmetic of
// mfma_fisher_kernel, group by group in the same order, so the encodings are
// the same bit for bit; but iteration i of a wave issues the logit products of
// group i and then the moment products of group i - 1, with the posteriors of
// group i (exponentials, threshold, rounding) computed under those moment
// products, so that the vector work of one group overlaps the matrix work of
// its neighbour instead of alternating with it.  Group records are copied
// straight into LDS (global_load_lds, 16 bytes per lane, 1 KiB per wave-wide
// copy) by parts: iteration i reads the A part and statistics of group i
// (logits, posteriors) and the B part of group i - 1 (moments) while the A part
// of group i + 1 and the B part of group i are copied, so two slots of each
// part suffice (the statistics' copy repeats the record's last word into its
// slot's padding).  One barrier per group.
template <int NS>
constexpr size_t pipe_a_bytes() {  // A part + statistics, whole KiB
  return GroupRec<NS>::kA + 1024;
}
template <int NS>
constexpr size_t pipe_lds_bytes() {  // two A slots, two B slots
  return 2 * (pipe_a_bytes<NS>() + GroupRec<NS>::kB);
}
// D > 112 keeps the double-buffered kernel: with 32 components per wave the
// pipelined form needs more than 512 registers there
template <int NS>
constexpr bool pipe_fits() {
  return NS <= 7 && pipe_lds_bytes<NS>() <= 160 * 1024;
}

template <typename T, typename H, int NS, int CT, int WAVES, bool FAST>
__global__ __launch_bounds__(WAVES * 64, 1) void mfma_fisher_pipe_kernel(
    const unsigned char* __restrict__ rec, const uint64_t* __restrict__ offsets, const long long* __restrict__ gstart,
    int imgBegin, int numImages, int D, int K, const H* __restrict__ Whi, const H* __restrict__ Wlo,
    const float* __restrict__ bias, const MfmaScales* __restrict__ scp, const float* __restrict__ mup,
    const float* __restrict__ isg, const float* __restrict__ sinv, const T* __restrict__ priors, int flags,
    T* __restrict__ enc, double* __restrict__ normpart, unsigned long long* __restrict__ numTerms, int numKB,
    int sched) {
  typedef typename VecOf<H>::type V8;
  typedef GroupRec<NS> R;
  constexpr int BC = WAVES * 16 * CT;
  constexpr int NV = (int)(R::kBytes / 16);   // 16-byte words per record
  static_assert(R::kA == (size_t)4 * NS * 1024 && R::kB == (size_t)4 * NS * 1024 && R::kStats <= 1024,
                "record parts in whole KiB, statistics in one");
  constexpr int PA = 4 * NS + 1;              // 1 KiB copies of an A slot (the last: statistics)
  constexpr int PB = 4 * NS;                  // 1 KiB copies of a B slot
  constexpr int PWA = (PA + WAVES - 1) / WAVES, PWB = (PB + WAVES - 1) / WAVES;
  constexpr size_t AS = pipe_a_bytes<NS>();
  extern __shared__ __align__(16) unsigned char smem_raw[];  // A slots 0, 1, then B slots 0, 1
  __shared__ double red[WAVES];

  int li, kb;
  fisher_mfma_coords(numKB, numImages, sched, li, kb);
  if (li >= numImages) return;
  const int img = imgBegin + li;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int kw = kb * BC + 16 * CT * w;  // first component of this wave
  const long long nImg = (long long)offsets[img + 1] - (long long)offsets[img];
  const long long gBase = gstart[imgBegin];
  const long long gB = gstart[img] - gBase;  // this image's records
  const int n = (int)(gstart[img + 1] - gstart[img]);
  const MfmaScales sc = *scp;

  auto abuf = [&](int i) __attribute__((always_inline)) { return smem_raw + (size_t)(i & 1) * AS; };
  auto bbuf = [&](int i) __attribute__((always_inline)) {
    return smem_raw + 2 * AS + (size_t)(i & 1) * R::kB;
  };
  // 1 KiB copy q of group gB + i's record into LDS at dst (every wave takes part)
  auto copy_piece = [&](int i, int q, unsigned char* dst) __attribute__((always_inline)) {
    const unsigned char* src = rec + (size_t)(gB + i) * R::kBytes;
    const int word = 64 * q + lane < NV ? 64 * q + lane : NV - 1;
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + 16 * (size_t)word),
                                     (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
  };
  // A part and statistics of group i into A slot i % 2
  auto copy_a = [&](int i) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < PWA; ++k) {
      const int u = w + WAVES * k;
      if (u < PA) copy_piece(i, u < 4 * NS ? u : 8 * NS, abuf(i) + 1024 * u);
    }
  };
  // B part of group i into B slot i % 2
  auto copy_b = [&](int i) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < PWB; ++k) {
      const int u = w + WAVES * k;
      if (u < PB) copy_piece(i, 4 * NS + u, bbuf(i) + 1024 * u);
    }
  };
  if (n > 0) copy_a(0);

  V8 whi[CT][NS], wlo[CT][NS];
  float bc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int comp = kw + 16 * c + lr;
    bc[c] = bias[comp];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      whi[c][j] = *reinterpret_cast<const V8*>(Whi + (size_t)comp * 2 * NS * 16 + 32 * j + 8 * lg);
      wlo[c][j] = *reinterpret_cast<const V8*>(Wlo + (size_t)comp * 2 * NS * 16 + 32 * j + 8 * lg);
    }
  }
  f32x4 macc[CT][2][NS];  // [comp tile][S1, S2][feature tile]
#pragma unroll
  for (int c = 0; c < CT; ++c)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int ft = 0; ft < NS; ++ft) macc[c][s][ft] = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 macc0[CT];  // posterior mass S0 (every column)
#pragma unroll
  for (int c = 0; c < CT; ++c) macc0[c] = f32x4{0.f, 0.f, 0.f, 0.f};
  V8 ones;
#pragma unroll
  for (int v = 0; v < 8; ++v) ones[v] = H(1.f);
  unsigned terms = 0;  // kept (row, component) pairs of this wave, counted from lane masks
  // (double)q < 1e-6 for a float q is q < the next float above float(1e-6),
  // compared after the exact power-of-two scaling by sq
  const float qminS = __uint_as_float(0x358637beu) * sc.sq;

  // logits - m of the group in A slot b (the accumulators start at bias - m)
  auto logits = [&](const unsigned char* b, f32x4 (&acc)[2][CT]) __attribute__((always_inline)) {
    const V8* Ap = reinterpret_cast<const V8*>(b);
    const float* Sf = reinterpret_cast<const float*>(b + R::kA);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const f32x4 nm = *reinterpret_cast<const f32x4*>(&Sf[16 * t + 4 * lg]);
#pragma unroll
      for (int c = 0; c < CT; ++c) acc[t][c] = f32x4{bc[c] + nm[0], bc[c] + nm[1], bc[c] + nm[2], bc[c] + nm[3]};
    }
    // the logit products run at raised wave priority: the other wave of the
    // SIMD then fills the matrix pipe's gaps with its vector and LDS work
    __builtin_amdgcn_s_setprio(1);
    // the A fragments of step j + 1 are read while step j's products run
    V8 ah[2][2], al[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      ah[0][t] = Ap[((t * NS + 0) * 2 + 0) * 64 + lane];
      al[0][t] = Ap[((t * NS + 0) * 2 + 1) * 64 + lane];
    }
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int u = j & 1;
      if (j + 1 < NS) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          ah[u ^ 1][t] = Ap[((t * NS + j + 1) * 2 + 0) * 64 + lane];
          al[u ^ 1][t] = Ap[((t * NS + j + 1) * 2 + 1) * 64 + lane];
        }
      }
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[t][c] = mfma16(ah[u][t], whi[c][j], acc[t][c]);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[t][c] = mfma16(al[u][t], whi[c][j], acc[t][c]);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < CT; ++c) acc[t][c] = mfma16(ah[u][t], wlo[c][j], acc[t][c]);
      __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_s_setprio(0);
  };
  // scaled posteriors q * sq of the group in A slot b as A fragments of the
  // moment contraction (padding rows have sq / sum = 0 and FAST argmax -1)
  auto posteriors = [&](const unsigned char* b, const f32x4 (&acc)[2][CT], V8 (&aq)[CT])
                        __attribute__((always_inline)) {
    const float* Sf = reinterpret_cast<const float*>(b + R::kA);
    if constexpr (FAST) {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const f32x4 ar = *reinterpret_cast<const f32x4*>(&Sf[2 * GR + 16 * t + 4 * lg]);
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            const bool hit = kw + 16 * c + lr == __float_as_int(ar[r]) && bc[c] != -INFINITY;
            terms += (unsigned)__popcll(__ballot(hit));
            aq[c][4 * t + r] = hit ? H(sc.sq) : H(0.f);
          }
      }
    } else {
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const f32x4 rs = *reinterpret_cast<const f32x4*>(&Sf[GR + 16 * t + 4 * lg]);
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            const float v = __builtin_amdgcn_exp2f(acc[t][c][r]) * rs[r];
            const bool keep = v >= qminS;
            terms += (unsigned)__popcll(__ballot(keep));
            aq[c][4 * t + r] = H(keep ? v : 0.f);
          }
      }
    }
  };
  // moments of the group in B slot b with its posteriors aq: 4 planes x NS
  // feature tiles, each fragment read three steps ahead of its products; with
  // post, the posteriors of the next group (acc, A slot pa) are formed one
  // element per step under these products (steps 2 ..), the rest after them
  constexpr int MS = 4 * NS;       // moment steps
  constexpr int PREF = 3;          // fragment reads in flight
  auto moments = [&](const unsigned char* b, const V8 (&aq)[CT], bool post, const unsigned char* pa,
                     const f32x4 (&acc)[2][CT], V8 (&aqn)[CT]) __attribute__((always_inline)) {
    const V8* Bp = reinterpret_cast<const V8*>(b);
    const float* Sn = reinterpret_cast<const float*>(pa + R::kA);
    V8 bx[PREF];
#pragma unroll
    for (int s = 0; s < PREF && s < MS; ++s) bx[s] = Bp[s * 64 + lane];
    f32x4 rs[2];
    if (!FAST && post) {
#pragma unroll
      for (int t = 0; t < 2; ++t) rs[t] = *reinterpret_cast<const f32x4*>(&Sn[GR + 16 * t + 4 * lg]);
    }
    f32x4 ar[2];
    if (FAST && post) {
#pragma unroll
      for (int t = 0; t < 2; ++t) ar[t] = *reinterpret_cast<const f32x4*>(&Sn[2 * GR + 16 * t + 4 * lg]);
    }
    // one posterior element (t, r, c) of the next group
    auto element = [&](int e) __attribute__((always_inline)) {
      const int c = e % CT, r = (e / CT) & 3, t = e / (4 * CT);
      if constexpr (FAST) {
        const bool hit = kw + 16 * c + lr == __float_as_int(ar[t][r]) && bc[c] != -INFINITY;
        terms += (unsigned)__popcll(__ballot(hit));
        aqn[c][4 * t + r] = hit ? H(sc.sq) : H(0.f);
      } else {
        const float v = __builtin_amdgcn_exp2f(acc[t][c][r]) * rs[t][r];
        const bool keep = v >= qminS;
        terms += (unsigned)__popcll(__ballot(keep));
        aqn[c][4 * t + r] = H(keep ? v : 0.f);
      }
    };
    constexpr int NE = 8 * CT;  // posterior elements per lane
    // posterior mass: every column of q x ones is the row sum
#pragma unroll
    for (int c = 0; c < CT; ++c) macc0[c] = mfma16(aq[c], ones, macc0[c]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      const int p = s / NS, ft = s - (s / NS) * NS;
      const V8 cur = bx[s % PREF];
      if (s + PREF < MS) bx[s % PREF] = Bp[(s + PREF) * 64 + lane];
#pragma unroll
      for (int c = 0; c < CT; ++c) macc[c][p >> 1][ft] = mfma16(aq[c], cur, macc[c][p >> 1][ft]);
      if (post && s >= 2 && s - 2 < NE) element(s - 2);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (post) {
#pragma unroll
      for (int e = MS - 2 > 0 ? MS - 2 : 0; e < NE; ++e) element(e);
    }
  };
  // end of iteration i (logits and posteriors of group i, moments of group
  // i - 1): every wave is done with A slot i % 2 and B slot (i - 1) % 2, and the
  // copies of A(i + 1) and B(i) have landed; the slots take A(i + 2) and B(i + 1)
  auto next = [&](int i) __attribute__((always_inline)) {
    __builtin_amdgcn_s_waitcnt(0);  // this wave's copies have landed
    __syncthreads();
    if (i + 2 < n) copy_a(i + 2);
    if (i + 1 < n) copy_b(i + 1);
  };

  if (n > 1) copy_a(1);
  if (n > 0) copy_b(0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  V8 aqPrev[CT];
  if (n > 0) {
    f32x4 acc[2][CT];
    if constexpr (!FAST) logits(abuf(0), acc);
    posteriors(abuf(0), acc, aqPrev);
    next(0);
  }
  for (int i = 1; i < n; ++i) {
    f32x4 acc[2][CT];
    V8 aq[CT];
    if constexpr (!FAST) logits(abuf(i), acc);
    moments(bbuf(i - 1), aqPrev, true, abuf(i), acc, aq);
#pragma unroll
    for (int c = 0; c < CT; ++c) aqPrev[c] = aq[c];
    next(i);
  }
  if (n > 0) {
    f32x4 acc[2][CT];
    V8 aq[CT];
    moments(bbuf(n - 1), aqPrev, false, abuf(n - 1), acc, aq);
  }

  fisher_slice_finalize<T, NS, CT, WAVES>(macc0, macc, sc, kw, lr, lg, tid, lane, img, kb, numKB, nImg, D, K, mup,
                                          isg, sinv, priors, flags, enc, normpart, numTerms, terms, red);
}

// ---------------------------------------------------------------------------
// Paired second pass: 128-component blocks on eight waves.  With 32
// components per wave, D > 80 needs more than 256 registers per lane for the
// split weights and the moment accumulators, so the four-wave tilings run one
// wave per SIMD and nothing hides a wave's LDS reads and vector work.  Here
// waves w and w + 4 form a pair that owns 32 components: wave h of the pair
// forms the logits and posteriors of components 16h .. 16h + 15 (every
// feature, its split weights in registers), the pair exchanges posteriors
// through LDS (1 KiB per wave and group, two slots), and each wave
// accumulates the moments of all 32 components over its half of the feature
// tiles.  Per lane: weights 8 NS and moment accumulators 16 ceil(NS / 2)
// registers -- two waves per SIMD at D = 128.  Records are copied in parts as
// in the pipelined kernel, and iteration i issues the logits of group i, then
// the moments of group i - 1 with this wave's posteriors of group i formed
// one element per product.  Logits, posteriors and moment sums are those of
// mfma_fisher_kernel operation for operation, so the unnormalised encodings
// and the EM moments are the same bit for bit (the slice norm is summed over
// the waves in another order).  With EM the kernel writes the EM moments of
// a chunk as mfma_fisher_kernel<..., EM = true> does.
template <int NS>
constexpr size_t pair_lds_bytes() {  // A and B slots, two posterior slots of 8 KiB
  return pipe_lds_bytes<NS>() + 2 * 8 * 1024;
}

template <typename T, typename H, int NS, bool FAST, bool EM>
__global__ __launch_bounds__(512, 1) void mfma_fisher_pair_kernel(
    const unsigned char* __restrict__ rec, const uint64_t* __restrict__ offsets, const long long* __restrict__ gstart,
    int imgBegin, int numImages, int D, int K, const H* __restrict__ Whi, const H* __restrict__ Wlo,
    const float* __restrict__ bias, const MfmaScales* __restrict__ scp, const float* __restrict__ mup,
    const float* __restrict__ isg, const float* __restrict__ sinv, const T* __restrict__ priors, int flags,
    T* __restrict__ enc, double* __restrict__ normpart, unsigned long long* __restrict__ numTerms, int numKB,
    float emThr, double* __restrict__ emOut, int sched) {
  static_assert(!(FAST && EM), "EM takes soft posteriors");
  typedef typename VecOf<H>::type V8;
  typedef GroupRec<NS> R;
  constexpr int WAVES = 8, CT = 2;
  constexpr int BC = 128;              // four pairs x 32 components
  constexpr int NF = (NS + 1) / 2;     // feature tiles of a wave's moments (the second half may lack one)
  constexpr int NV = (int)(R::kBytes / 16);
  static_assert(R::kA == (size_t)4 * NS * 1024 && R::kB == (size_t)4 * NS * 1024 && R::kStats <= 1024,
                "record parts in whole KiB, statistics in one");
  constexpr int PA = 4 * NS + 1, PB = 4 * NS;
  constexpr int PWA = (PA + WAVES - 1) / WAVES, PWB = (PB + WAVES - 1) / WAVES;
  constexpr size_t AS = pipe_a_bytes<NS>();
  extern __shared__ __align__(16) unsigned char smem_raw[];  // A slots 0, 1, B slots 0, 1, posterior slots 0, 1
  __shared__ double red[WAVES];

  int li, kb;
  fisher_mfma_coords(numKB, numImages, sched, li, kb);
  if (li >= numImages) return;
  const int img = imgBegin + li;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int h = w >> 2;                      // half of the pair
  const int kw = kb * BC + 32 * (w & 3);     // first component of the pair
  const int kown = kw + 16 * h;              // this wave's logit components kown + lr
  const int ft0 = h * NF;                    // this wave's first moment feature tile
  const long long nImg = (long long)offsets[img + 1] - (long long)offsets[img];
  const long long gB = gstart[img] - gstart[imgBegin];
  const int n = (int)(gstart[img + 1] - gstart[img]);
  const MfmaScales sc = *scp;

  auto abuf = [&](int i) __attribute__((always_inline)) { return smem_raw + (size_t)(i & 1) * AS; };
  auto bbuf = [&](int i) __attribute__((always_inline)) {
    return smem_raw + 2 * AS + (size_t)(i & 1) * R::kB;
  };
  auto xbuf = [&](int i) __attribute__((always_inline)) {
    return reinterpret_cast<V8*>(smem_raw + 2 * AS + 2 * R::kB + (size_t)(i & 1) * 8192);
  };
  auto copy_piece = [&](int i, int q, unsigned char* dst) __attribute__((always_inline)) {
    const unsigned char* src = rec + (size_t)(gB + i) * R::kBytes;
    const int word = 64 * q + lane < NV ? 64 * q + lane : NV - 1;
    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void*)(src + 16 * (size_t)word),
                                     (__attribute__((address_space(3))) void*)dst, 16, 0, 0);
  };
  auto copy_a = [&](int i) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < PWA; ++k) {
      const int u = w + WAVES * k;
      if (u < PA) copy_piece(i, u < 4 * NS ? u : 8 * NS, abuf(i) + 1024 * u);
    }
  };
  auto copy_b = [&](int i) __attribute__((always_inline)) {
#pragma unroll
    for (int k = 0; k < PWB; ++k) {
      const int u = w + WAVES * k;
      if (u < PB) copy_piece(i, 4 * NS + u, bbuf(i) + 1024 * u);
    }
  };
  if (n > 0) copy_a(0);

  V8 whi[NS], wlo[NS];
  const float bc = bias[kown + lr];
#pragma unroll
  for (int j = 0; j < NS; ++j) {
    whi[j] = *reinterpret_cast<const V8*>(Whi + (size_t)(kown + lr) * 2 * NS * 16 + 32 * j + 8 * lg);
    wlo[j] = *reinterpret_cast<const V8*>(Wlo + (size_t)(kown + lr) * 2 * NS * 16 + 32 * j + 8 * lg);
  }
  f32x4 macc[CT][2][NF];  // [comp tile][S1, S2][feature tile ft0 + f]
#pragma unroll
  for (int c = 0; c < CT; ++c)
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int f = 0; f < NF; ++f) macc[c][s][f] = f32x4{0.f, 0.f, 0.f, 0.f};
  f32x4 macc0[CT];  // posterior mass S0 of both component tiles (both waves)
#pragma unroll
  for (int c = 0; c < CT; ++c) macc0[c] = f32x4{0.f, 0.f, 0.f, 0.f};
  V8 ones;
#pragma unroll
  for (int v = 0; v < 8; ++v) ones[v] = H(1.f);
  unsigned terms = 0;  // kept (row, component) pairs of this wave's components
  double massd = 0.0;  // EM: unthresholded mass of component kown + lr over this lane's rows
  float msum = 0.f;
  const float qminS = EM ? emThr : __uint_as_float(0x358637beu) * sc.sq;  // 1.0000001e-6 * sq

  // logits - m of this wave's components for the group in A slot b
  auto logits = [&](const unsigned char* b, f32x4 (&acc)[2]) __attribute__((always_inline)) {
    const V8* Ap = reinterpret_cast<const V8*>(b);
    const float* Sf = reinterpret_cast<const float*>(b + R::kA);
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const f32x4 nm = *reinterpret_cast<const f32x4*>(&Sf[16 * t + 4 * lg]);
      acc[t] = f32x4{bc + nm[0], bc + nm[1], bc + nm[2], bc + nm[3]};
    }
    __builtin_amdgcn_s_setprio(1);  // logit products at raised wave priority, as above
    V8 ah[2][2], al[2][2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      ah[0][t] = Ap[((t * NS + 0) * 2 + 0) * 64 + lane];
      al[0][t] = Ap[((t * NS + 0) * 2 + 1) * 64 + lane];
    }
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int u = j & 1;
      if (j + 1 < NS) {
#pragma unroll
        for (int t = 0; t < 2; ++t) {
          ah[u ^ 1][t] = Ap[((t * NS + j + 1) * 2 + 0) * 64 + lane];
          al[u ^ 1][t] = Ap[((t * NS + j + 1) * 2 + 1) * 64 + lane];
        }
      }
#pragma unroll
      for (int t = 0; t < 2; ++t) acc[t] = mfma16(ah[u][t], whi[j], acc[t]);
#pragma unroll
      for (int t = 0; t < 2; ++t) acc[t] = mfma16(al[u][t], whi[j], acc[t]);
#pragma unroll
      for (int t = 0; t < 2; ++t) acc[t] = mfma16(ah[u][t], wlo[j], acc[t]);
      __builtin_amdgcn_sched_barrier(0);
    }
    __builtin_amdgcn_s_setprio(0);
  };
  // statistics of this lane's rows 16t + 4lg + r of the group in A slot b
  auto row_stats = [&](const unsigned char* b, f32x4 (&rs)[2]) __attribute__((always_inline)) {
    const float* Sf = reinterpret_cast<const float*>(b + R::kA);
#pragma unroll
    for (int t = 0; t < 2; ++t)
      rs[t] = *reinterpret_cast<const f32x4*>(&Sf[(FAST ? 2 * GR : GR) + 16 * t + 4 * lg]);
  };
  // posterior element e = 4t + r of this wave's component for this lane
  auto element = [&](int e, const f32x4 (&acc)[2], const f32x4 (&rs)[2], V8& aq) __attribute__((always_inline)) {
    const int t = e >> 2, r = e & 3;
    if constexpr (FAST) {
      const bool hit = kown + lr == __float_as_int(rs[t][r]) && bc != -INFINITY;
      terms += (unsigned)__popcll(__ballot(hit));
      aq[e] = hit ? H(sc.sq) : H(0.f);
    } else {
      const float v = __builtin_amdgcn_exp2f(acc[t][r]) * rs[t][r];
      const bool keep = v >= qminS;
      terms += (unsigned)__popcll(__ballot(keep));
      // padding rows (sq / sum = 0) may have 2^acc = inf: keep them out of the mass
      if (EM) msum += rs[t][r] > 0.f ? v : 0.f;
      aq[e] = H(keep ? v : 0.f);
    }
  };
  auto mass_done = [&]() __attribute__((always_inline)) {
    if (EM) {
      massd += (double)msum;
      msum = 0.f;
    }
  };
  // moments of the group in B slot b with the posteriors q0, q1 of the pair's
  // two component tiles over this wave's feature tiles; with post, this wave's
  // posteriors of the next group are formed one element per product (steps
  // 2 ..), the rest after them
  constexpr int MS = 4 * NF;
  constexpr int PREF = 3;
  auto moments = [&](const unsigned char* b, const V8& q0, const V8& q1, bool post, const f32x4 (&acc)[2],
                     const f32x4 (&rs)[2], V8& aqn) __attribute__((always_inline)) {
    const V8* Bp = reinterpret_cast<const V8*>(b);
    // fragment of step s (plane s / NF, tile ft0 + s % NF; past the last tile, the last one, unused)
    auto frag = [&](int s) __attribute__((always_inline)) {
      const int ft = ft0 + s % NF < NS ? ft0 + s % NF : NS - 1;
      return Bp[((s / NF) * NS + ft) * 64 + lane];
    };
    V8 bx[PREF];
#pragma unroll
    for (int s = 0; s < PREF && s < MS; ++s) bx[s] = frag(s);
    macc0[0] = mfma16(q0, ones, macc0[0]);
    macc0[1] = mfma16(q1, ones, macc0[1]);
    __builtin_amdgcn_sched_barrier(0);
#pragma unroll
    for (int s = 0; s < MS; ++s) {
      const int p = s / NF, f = s % NF;
      const V8 cur = bx[s % PREF];
      if (s + PREF < MS) bx[s % PREF] = frag(s + PREF);
      if ((NS & 1) == 0 || ft0 + f < NS) {
        macc[0][p >> 1][f] = mfma16(q0, cur, macc[0][p >> 1][f]);
        macc[1][p >> 1][f] = mfma16(q1, cur, macc[1][p >> 1][f]);
      }
      if (post && s >= 2 && s - 2 < 8) element(s - 2, acc, rs, aqn);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (post) {
#pragma unroll
      for (int e = MS - 2 > 0 ? MS - 2 : 0; e < 8; ++e) element(e, acc, rs, aqn);
      mass_done();
    }
  };
  auto next = [&](int i) __attribute__((always_inline)) {
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    if (i + 2 < n) copy_a(i + 2);
    if (i + 1 < n) copy_b(i + 1);
  };

  if (n > 1) copy_a(1);
  if (n > 0) copy_b(0);
  __builtin_amdgcn_s_waitcnt(0);
  __syncthreads();
  V8 own;  // this wave's posteriors of the previous group
  if (n > 0) {
    f32x4 acc[2], rs[2];
    if constexpr (!FAST) logits(abuf(0), acc);
    row_stats(abuf(0), rs);
#pragma unroll
    for (int e = 0; e < 8; ++e) element(e, acc, rs, own);
    mass_done();
    xbuf(0)[w * 64 + lane] = own;
    next(0);
  }
  for (int i = 1; i < n; ++i) {
    const V8 other = xbuf(i - 1)[(w ^ 4) * 64 + lane];
    f32x4 acc[2], rs[2];
    V8 aq;
    if constexpr (!FAST) logits(abuf(i), acc);
    row_stats(abuf(i), rs);
    moments(bbuf(i - 1), h ? other : own, h ? own : other, true, acc, rs, aq);
    xbuf(i)[w * 64 + lane] = aq;
    own = aq;
    next(i);
  }
  if (n > 0) {
    const V8 other = xbuf(n - 1)[(w ^ 4) * 64 + lane];
    f32x4 acc[2], rs[2];
    V8 aq;
    moments(bbuf(n - 1), h ? other : own, h ? own : other, false, acc, rs, aq);
  }

  if constexpr (EM) {
    const double invq = 1.0 / (double)sc.sq;
    const double inv1 = invq / (double)sc.s1, inv2 = invq / (double)sc.s2;
    const size_t KD = (size_t)K * D;
    double* P = emOut + (size_t)img * (2 * (size_t)K + 2 * KD);
    double m = massd;
    m += __shfl_xor(m, 16, 64);
    m += __shfl_xor(m, 32, 64);
    if (lg == 0 && kown + lr < K) P[K + kown + lr] = m * invq;
#pragma unroll
    for (int c = 0; c < CT; ++c)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int k = kw + 16 * c + 4 * lg + r;
        if (k >= K) continue;
        if (lr == 0 && c == h) P[k] = (double)macc0[c][r] * invq;
#pragma unroll
        for (int f = 0; f < NF; ++f) {
          const int d = 16 * (ft0 + f) + lr;
          if (d >= D) continue;
          P[2 * (size_t)K + (size_t)k * D + d] = (double)macc[c][0][f][r] * inv1;
          P[2 * (size_t)K + KD + (size_t)k * D + d] = (double)macc[c][1][f][r] * inv2;
        }
      }
    return;
  } else {
    fisher_slice_finalize<T, NS, CT, WAVES, NF>(macc0, macc, sc, kw, lr, lg, tid, lane, img, kb, numKB, nImg, D, K,
                                                mup, isg, sinv, priors, flags, enc, normpart, numTerms, terms, red,
                                                ft0);
  }
}

// ---------------------------------------------------------------------------
// Fused single-pass encoder.  The two-pass form computes every logit twice:
// once for the row log-sum-exp (first pass) and once more for the posteriors
// (second pass), because a workgroup owns only a block of components and the
// normaliser of a row needs all of them.  Here the numKB workgroups that own
// the component blocks of an image ("partners") walk the image's groups in
// step and exchange, per group, each row's partial statistics over their own
// block (max logit, sum of 2^(logit - max) and, for FAST, the first argmax):
//
//   iteration n:  build the split fragments of group n in LDS, logits(n),
//                 publish the block's partial statistics of group n;
//                 moments of group n - 1 (its posteriors held in registers),
//                 which hides the partners' latency;
//                 wait for the partners' partials of group n, combine them
//                 in block order (so the result does not depend on arrival
//                 order), posteriors of group n -> registers.
//
// Partials are tagged 64-bit words {tag = n + 1 | value} written and read with
// device-scope relaxed atomics, so a reader that sees the tag sees the value:
// no fence is needed, and a ring of FX_RING slots per (stream, block) is
// enough because a workgroup can publish group n + 2 only after every partner
// has consumed its group n.  All workgroups are co-resident (cooperative
// launch), so every wait ends; each wait is still bounded (FX_SPIN_MAX polls,
// or a partner's abort), after which the workgroup raises the abort flag and
// returns, and the host reports the failure.
//
// Blocks are dealt to the XCDs round robin, and the partners of a stream sit
// on one XCD (they read the same descriptors).  Each stream encodes its list
// of images (balanced by size on the host) one after the other; the
// arithmetic of a row, a group and an image is that of the second pass, with
// the normaliser combined from the block partials, so an image's encoding
// does not depend on the batch or on the stream it lands on.
// Logit origin of the fused encoder: max_k b_k - (D / 2) log2 e, about where
// a row's largest logit falls (a component whose squared Mahalanobis distance
// is D).  The fused encoder accumulates logit - shift, so that the logits that
// matter are formed near zero and rounded there, as the second pass forms
// logit - row maximum; the shift cancels in the posteriors.
__global__ __launch_bounds__(256) void mfma_shift_kernel(const float* __restrict__ bias, int Kp, int D,
                                                         MfmaScales* __restrict__ scp) {
  __shared__ float red[256];
  float m = -INFINITY;
  for (int k = threadIdx.x; k < Kp; k += 256) m = fmaxf(m, bias[k]);
  red[threadIdx.x] = m;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if (threadIdx.x < o) red[threadIdx.x] = fmaxf(red[threadIdx.x], red[threadIdx.x + o]);
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const float mb = red[0];
    scp->shift = (mb == -INFINITY) ? 0.f : (float)((double)mb - 0.5 * (double)D * kLog2eD);
  }
}

constexpr int FX_RING = 4;
constexpr int FX_SPIN_MAX = 1 << 20;

__device__ __forceinline__ unsigned long long fx_word(unsigned tag, unsigned bits) {
  return ((unsigned long long)tag << 32) | (unsigned long long)bits;
}
__device__ __forceinline__ void fx_store(unsigned long long* p, unsigned long long v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long fx_load(const unsigned long long* p) {
  return __hip_atomic_load(const_cast<unsigned long long*>(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The fused kernel's output for one (image, component block): the raw FP32
// moment accumulators, still scaled by the operand scales, S1 and S2 in the
// places of the image's u and v entries and S0 in s0raw[image][K]; lane l
// holds features 16ft + (l & 15) of components kw + 16c + 4(l >> 4) + r.
// mfma_fused_finalize_kernel turns them into the Fisher vector.
template <int NS, int CT>
__device__ __forceinline__ void fused_store_raw(const f32x4 (&macc0)[CT], const f32x4 (&macc)[CT][2][NS], int kw,
                                                int lr, int lg, int img, int D, int K, float* __restrict__ enc,
                                                float* __restrict__ s0raw) {
  // the lane's coordinates pass through an empty asm, so that the store
  // offsets are formed here and not hoisted out of the caller's loops (where
  // eighty of them would stay live across the whole sweep)
  int k0 = kw + 4 * lg, l0 = lr;
  asm volatile("" : "+v"(k0), "+v"(l0));
  const size_t KD = (size_t)K * D;
  float* e = enc + (size_t)img * 2 * KD;
#pragma unroll
  for (int c = 0; c < CT; ++c)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const int k = k0 + 16 * c + r;
      if (k >= K) continue;
      if (l0 == 0) s0raw[(size_t)img * K + k] = macc0[c][r];
      float* eu = e + (size_t)k * D + l0;
#pragma unroll
      for (int ft = 0; ft < NS; ++ft) {
        if (16 * ft + l0 >= D) continue;
        eu[16 * ft] = macc[c][0][ft][r];
        eu[KD + 16 * ft] = macc[c][1][ft][r];
      }
    }
}

// Staging of a group's rows: the 32 rows are contiguous in memory, so with D
// a multiple of 4 and 16-byte aligned data they are copied straight into LDS
// (global_load_lds_dwordx4: lane l of a wave writes bytes 16l .. 16l + 15 of
// the wave's 1 KiB piece), issued a whole iteration ahead and holding no
// registers; otherwise they are loaded through registers when needed.  Rows
// past the image are masked when the fragments are built; addresses past the
// end of the data are clamped to its last 16 bytes.
constexpr int FX_PIECE = 1024;  // bytes per wave-wide copy

template <typename H, int NS, int CT, int WAVES, bool FAST, bool DMA>
__global__ __launch_bounds__(WAVES * 64, 1) void mfma_fused_kernel(
    const float* __restrict__ x, unsigned long long xBytes, const uint64_t* __restrict__ offsets,
    const long long* __restrict__ streamOff, const int* __restrict__ streamImg, int numKB, int D, int K,
    const H* __restrict__ Whi, const H* __restrict__ Wlo, const float* __restrict__ bias,
    const MfmaScales* __restrict__ scp, const float* __restrict__ g, float* __restrict__ enc,
    float* __restrict__ s0raw, unsigned long long* __restrict__ numTerms, unsigned long long* __restrict__ xch,
    unsigned* __restrict__ abortFlag) {
  typedef typename VecOf<H>::type V8;
  typedef GroupRec<NS> R;
  constexpr int Dp = NS * 16;
  constexpr int NT = WAVES * 64;
  constexpr int BC = WAVES * 16 * CT;
  constexpr size_t FRAG = R::kA + R::kB;         // A and B fragments of one group
  constexpr int XR = (GR * Dp * 4 + FX_PIECE - 1) / FX_PIECE * (FX_PIECE / 4) + 8;  // staged floats (+ slack)
  constexpr int NA = 2 * NS * 64, NB = NS * 64;  // fragment items (A: hi + lo, B: 4 planes)
  constexpr int NI = (NA + NB + NT - 1) / NT;
  extern __shared__ __align__(16) unsigned char smem_raw[];  // 2 x FRAG
  __shared__ __align__(16) float xr[XR];         // the group's rows, [row][D], uncentred
  __shared__ __align__(16) float gsh[Dp];
  __shared__ float pm[WAVES][GR], ps[WAVES][GR];
  __shared__ int pa[WAVES][GR];
  __shared__ float fm[GR], frs[GR];
  __shared__ int fa[GR];
  __shared__ int fabort;

  // block -> (stream, component block); the partners of a stream share an XCD
  const int b = blockIdx.x, xcd = b & 7, slot = b >> 3;
  const int kb = slot % numKB, stream = (slot / numKB) * 8 + xcd;
  const long long i0 = streamOff[stream], i1 = streamOff[stream + 1];
  if (i0 >= i1) return;  // the partners have the same (empty) list

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int kw = kb * BC + 16 * CT * w;  // first component of this wave
  const MfmaScales sc = *scp;
  const float qminS = __uint_as_float(0x358637beu) * sc.sq;  // next float above float(1e-6), scaled

  V8 whi[CT][NS], wlo[CT][NS];
  float bc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int comp = kw + 16 * c + lr;
    bc[c] = bias[comp];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      whi[c][j] = *reinterpret_cast<const V8*>(Whi + (size_t)comp * 2 * Dp + 32 * j + 8 * lg);
      wlo[c][j] = *reinterpret_cast<const V8*>(Wlo + (size_t)comp * 2 * Dp + 32 * j + 8 * lg);
    }
  }
  float bcs[CT];  // biases relative to the logit origin
#pragma unroll
  for (int c = 0; c < CT; ++c) bcs[c] = bc[c] - sc.shift;
  for (int d = tid; d < Dp; d += NT) gsh[d] = g[d];
  V8 ones;
#pragma unroll
  for (int v = 0; v < 8; ++v) ones[v] = H(1.f);
  f32x4 macc[CT][2][NS], macc0[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    macc0[c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
      for (int ft = 0; ft < NS; ++ft) macc[c][s2][ft] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  unsigned terms = 0;
  V8 aqp[CT];  // posteriors (scaled, 16-bit) of the pending group, A fragments of the moments
  unsigned long long* own = xch + (size_t)(stream * numKB + kb) * FX_RING * 3 * GR;

  // rows [rowA, rowA + 32) into xr (see above)
  auto stage = [&](long long rowA) __attribute__((always_inline)) {
    if constexpr (DMA) {
      constexpr int Q = (GR * Dp * 4 + FX_PIECE - 1) / FX_PIECE;
      const unsigned long long base = (unsigned long long)rowA * (unsigned long long)D * 4ull;
      const int qn = (GR * D * 4 + FX_PIECE - 1) / FX_PIECE;
#pragma unroll
      for (int q0 = 0; q0 < Q; q0 += WAVES) {
        const int q = q0 + w;
        if (q < qn) {
          unsigned long long off = base + (unsigned long long)q * FX_PIECE + 16ull * lane;
          if (off + 16 > xBytes) off = xBytes - 16;
          __builtin_amdgcn_global_load_lds(
              (const __attribute__((address_space(1))) void*)(reinterpret_cast<const char*>(x) + off),
              (__attribute__((address_space(3))) void*)(xr + q * (FX_PIECE / 4)), 16, 0, 0);
        }
      }
    } else {
      const long long rowLim = (long long)(xBytes / 4ull / (unsigned long long)D);
      for (int i = tid; i < GR * D; i += NT) {
        const long long row = rowA + i / D;
        xr[i] = row < rowLim ? x[rowA * D + i] : 0.f;
      }
    }
  };

  int pendImg = -1;            // image of the pending group (moments not yet added)
  bool pendLast = false;       // the pending group is its image's last
  int pendBuf = 0;
  long long n = 0;             // groups of this stream so far (the exchange sequence)

  // moments of the pending group; when it was its image's last, the image's raw slice
  auto flush = [&]() __attribute__((always_inline)) {
    const V8* Bp = reinterpret_cast<const V8*>(smem_raw + (size_t)pendBuf * FRAG + R::kA);
#pragma unroll
    for (int c = 0; c < CT; ++c) macc0[c] = mfma16(aqp[c], ones, macc0[c]);
#pragma unroll
    for (int p = 0; p < 4; ++p) {
#pragma unroll
      for (int ft = 0; ft < NS; ++ft) {
        const V8 bx = Bp[(p * NS + ft) * 64 + lane];
#pragma unroll
        for (int c = 0; c < CT; ++c) macc[c][p >> 1][ft] = mfma16(aqp[c], bx, macc[c][p >> 1][ft]);
      }
    }
    if (pendLast) {
      fused_store_raw<NS, CT>(macc0, macc, kw, lr, lg, pendImg, D, K, enc, s0raw);
      // terms: the wave's count (lane masks, the same in every lane)
      if (numTerms && lane == 0 && terms) atomicAdd(&numTerms[pendImg], (unsigned long long)terms);
      terms = 0;
#pragma unroll
      for (int c = 0; c < CT; ++c) {
        macc0[c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
        for (int s2 = 0; s2 < 2; ++s2)
#pragma unroll
          for (int ft = 0; ft < NS; ++ft) macc[c][s2][ft] = f32x4{0.f, 0.f, 0.f, 0.f};
      }
    }
    pendImg = -1;
  };

  // the stream's groups as one sequence (images without rows are skipped:
  // the finalisation writes their zero encodings); position (li, gi)
  auto skip_empty = [&](long long l) __attribute__((always_inline)) {
    while (l < i1) {
      const int im = streamImg[l];
      if (offsets[im + 1] > offsets[im]) break;
      ++l;
    }
    return l;
  };
  long long li = skip_empty(i0), gi = 0;
  if (li < i1) stage((long long)offsets[streamImg[li]]);
  if constexpr (DMA) __builtin_amdgcn_s_waitcnt(0);  // this wave's copies have landed
  __syncthreads();                                  // xr holds the first group's rows
  while (true) {
    const bool have = li < i1;
    int img = 0;
    int nrow = 0, buf = 0;
    bool lastOfImg = false;
    long long nli = li, ngi = gi;
    f32x4 acc[2][CT];
    const unsigned tag = (unsigned)(n + 1);  // exchange sequence of this group
    const int rs = (int)(n % FX_RING);
    const int xrow = lane & (GR - 1);
    float bm = -INFINITY, bs = 0.f;          // the block's partial of row xrow (wave 0)
    int ba = 0;
    if (have) {
      img = streamImg[li];
      const long long rB = (long long)offsets[img], rE = (long long)offsets[img + 1];
      const long long row0 = rB + GR * gi;
      nrow = (int)min((long long)GR, rE - row0);
      buf = (int)(n & 1);
      unsigned char* fr = smem_raw + (size_t)buf * FRAG;
      // (xr holds the group's rows: copied before the previous iteration's last barrier)
      // ---- split fragments of the group (the first pass's record layout),
      // centred x' = x - g, zero past D and past the image
      {
        V8* Ap = reinterpret_cast<V8*>(fr);
        V8* Bp = reinterpret_cast<V8*>(fr + R::kA);
        // the thread index passes through an empty asm: the item addresses
        // (runtime row stride D) are formed here on every group instead of
        // being hoisted out of the loops and held in registers throughout
        int tv = tid;
        asm volatile("" : "+v"(tv));
#pragma unroll
        for (int q = 0; q < NI; ++q) {
          const int it = tv + NT * q;
          if (it < NA) {
            const int t = it / (NS * 64), rem = it - t * (NS * 64), j = rem >> 6, l = rem & 63;
            const int f0 = 32 * j + 8 * (l >> 4);
            const bool sqf = f0 < Dp;
            const int d0 = sqf ? f0 : f0 - Dp;
            const int row = 16 * t + (l & 15);
            float xc[8];
            if (DMA && row < nrow && d0 + 8 <= D) {
              // (D is a multiple of 4 here: 16-byte aligned reads)
              const f32x4 a0 = *reinterpret_cast<const f32x4*>(&xr[row * D + d0]) -
                               *reinterpret_cast<const f32x4*>(&gsh[d0]);
              const f32x4 a1 = *reinterpret_cast<const f32x4*>(&xr[row * D + d0 + 4]) -
                               *reinterpret_cast<const f32x4*>(&gsh[d0 + 4]);
#pragma unroll
              for (int v = 0; v < 4; ++v) { xc[v] = a0[v]; xc[4 + v] = a1[v]; }
            } else {
#pragma unroll
              for (int v = 0; v < 8; ++v) {
                const int d = d0 + v;
                xc[v] = (row < nrow && d < D) ? xr[row * D + d] - gsh[d] : 0.f;
              }
            }
            V8 hi, lo;
            make_frag<H>(xc, sqf, sc.s1, sc.s2, hi, lo);
            Ap[((t * NS + j) * 2 + 0) * 64 + l] = hi;
            Ap[((t * NS + j) * 2 + 1) * 64 + l] = lo;
          } else if (it < NA + NB) {
            const int ib = it - NA, ft = ib >> 6, l = ib & 63;
            const int d = 16 * ft + (l & 15), lg2 = l >> 4;
            const float gd = d < D ? gsh[d] : 0.f;
            V8 b1h, b1l, b2h, b2l;
#pragma unroll
            for (int v = 0; v < 8; ++v) {
              const int row = 16 * (v >> 2) + 4 * lg2 + (v & 3);
              const float xc = (row < nrow && d < D) ? xr[row * D + d] - gd : 0.f;
              H h, lo;
              split16<H>(xc * sc.s1, h, lo);
              b1h[v] = h;
              b1l[v] = lo;
              split16<H>(xc * xc * sc.s2, h, lo);
              b2h[v] = h;
              b2l[v] = lo;
            }
            Bp[(0 * NS + ft) * 64 + l] = b1h;
            Bp[(1 * NS + ft) * 64 + l] = b1l;
            Bp[(2 * NS + ft) * 64 + l] = b2h;
            Bp[(3 * NS + ft) * 64 + l] = b2l;
          }
        }
      }
      __syncthreads();  // fragments complete; xr is free
      // the next group's rows (possibly of the next image), copied while this
      // one is computed
      lastOfImg = row0 + GR >= rE;
      if (lastOfImg) {
        nli = skip_empty(li + 1);
        ngi = 0;
      } else {
        ngi = gi + 1;
      }
      if (nli < i1) stage((long long)offsets[streamImg[nli]] + GR * ngi);
      // ---- logits of the group minus the logit origin (log2 units)
      {
        const V8* Ap = reinterpret_cast<const V8*>(fr);
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int c = 0; c < CT; ++c) acc[t][c] = f32x4{bcs[c], bcs[c], bcs[c], bcs[c]};
#pragma unroll
        for (int j = 0; j < NS; ++j) {
          V8 ah[2], al[2];
#pragma unroll
          for (int t = 0; t < 2; ++t) {
            ah[t] = Ap[((t * NS + j) * 2 + 0) * 64 + lane];
            al[t] = Ap[((t * NS + j) * 2 + 1) * 64 + lane];
          }
#pragma unroll
          for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[t][c] = mfma16(ah[t], whi[c][j], acc[t][c]);
#pragma unroll
          for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[t][c] = mfma16(al[t], whi[c][j], acc[t][c]);
#pragma unroll
          for (int t = 0; t < 2; ++t)
#pragma unroll
            for (int c = 0; c < CT; ++c) acc[t][c] = mfma16(ah[t], wlo[c][j], acc[t][c]);
        }
      }
      // ---- the wave's partial statistics of each row (16 lanes share a row)
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float m = acc[t][0][r];
          float bv = acc[t][0][r];
          int bi = kw + lr;
#pragma unroll
          for (int c = 1; c < CT; ++c) {
            m = fmaxf(m, acc[t][c][r]);
            if (FAST && acc[t][c][r] > bv) { bv = acc[t][c][r]; bi = kw + 16 * c + lr; }
          }
          m = row16_max(m);
          const float ms = (m == -INFINITY) ? 0.f : m;
          float s = 0.f;
#pragma unroll
          for (int c = 0; c < CT; ++c) s += __builtin_amdgcn_exp2f(acc[t][c][r] - ms);
          s = row16_sum(s);
          if constexpr (FAST) row16_argmax(bv, bi);
          if (lr == 0) {
            const int row = 16 * t + 4 * lg + r;
            pm[w][row] = m;
            ps[w][row] = s;
            pa[w][row] = bi;
          }
        }
      __syncthreads();
      // ---- the block's partial of each row, published for the partners
      // (wave 0; lanes l and l + 32 both hold row l & 31, so that the wave's
      // control flow stays uniform)
      if (w == 0) {
#pragma unroll
        for (int v = 0; v < WAVES; ++v) bm = fmaxf(bm, pm[v][xrow]);
        const float bms = (bm == -INFINITY) ? 0.f : bm;
        bool found = false;
#pragma unroll
        for (int v = 0; v < WAVES; ++v) {
          const float mv = pm[v][xrow];
          const float sv = ps[v][xrow];
          if (sv != 0.f) bs += sv * __builtin_amdgcn_exp2f(((mv == -INFINITY) ? 0.f : mv) - bms);
          if (!found && mv == bm) { ba = pa[v][xrow]; found = true; }
        }
        if (lane < GR) {
          unsigned long long* sl = own + (size_t)rs * 3 * GR;
          fx_store(sl + xrow, fx_word(tag, __float_as_uint(bm)));
          fx_store(sl + GR + xrow, fx_word(tag, __float_as_uint(bs)));
          fx_store(sl + 2 * GR + xrow, fx_word(tag, (unsigned)ba));
        }
      }
    }
    // ---- moments of the pending group, while the partners publish (the
    // only place the accumulators are updated)
    if (pendImg >= 0) flush();
    if (!have) break;
    {
      // ---- the partners' partials, combined in block order: a running
      // maximum M and sum S = sum_p s_p 2^(m_p - M), and the first block whose
      // maximum is M gives the first argmax (blocks ascend in component index)
      if (w == 0) {
        bool ok = true;
        float M = -INFINITY, S = 0.f;
        int A = 0;
        for (int p = 0; p < numKB && ok; ++p) {
          float mp = bm, sp = bs;
          int ap = ba;
          if (p != kb) {
            const unsigned long long* sl = xch + ((size_t)(stream * numKB + p) * FX_RING + rs) * 3 * GR;
            int spins = 0;
            while (true) {  // uniform over the wave: it leaves when every lane's row has arrived
              const unsigned long long a0 = fx_load(sl + xrow), a1 = fx_load(sl + GR + xrow),
                                       a2 = fx_load(sl + 2 * GR + xrow);
              const bool ready =
                  (unsigned)(a0 >> 32) == tag && (unsigned)(a1 >> 32) == tag && (unsigned)(a2 >> 32) == tag;
              if (__all(ready)) {
                mp = __uint_as_float((unsigned)a0);
                sp = __uint_as_float((unsigned)a1);
                ap = (int)(unsigned)a2;
                break;
              }
              if (++spins > FX_SPIN_MAX ||
                  __hip_atomic_load(abortFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
                ok = false;
                break;
              }
              __builtin_amdgcn_s_sleep(1);
            }
            if (!ok) break;
          }
          const float mps = (mp == -INFINITY) ? 0.f : mp;
          if (mp > M) {
            const float Mso = (M == -INFINITY) ? 0.f : M;
            S = (S != 0.f ? S * __builtin_amdgcn_exp2f(Mso - mps) : 0.f) + sp;
            M = mp;
            A = ap;
          } else if (sp != 0.f) {
            S += sp * __builtin_amdgcn_exp2f(mps - ((M == -INFINITY) ? 0.f : M));
          }
        }
        if (ok && lane < GR) {
          const bool real = lane < nrow;
          fm[lane] = (M == -INFINITY) ? 0.f : M;
          frs[lane] = real ? sc.sq / S : 0.f;
          fa[lane] = real ? A : -1;
        }
        if (lane == 0) {
          fabort = ok ? 0 : 1;
          if (!ok) __hip_atomic_store(abortFlag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        }
      }
      // the next group's rows have landed (they are read after this barrier)
      if constexpr (DMA) __builtin_amdgcn_s_waitcnt(0);
      __syncthreads();
      if (fabort) return;  // uniform: every wave read the same flag after the barrier
      // ---- posteriors of this group -> the pending A fragments
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int row = 16 * t + 4 * lg + r;
          const float Ms = fm[row], rsv = frs[row];
          const int A = fa[row];
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            if constexpr (FAST) {
              const bool hit = kw + 16 * c + lr == A && bc[c] != -INFINITY;
              terms += (unsigned)__popcll(__ballot(hit));
              aqp[c][4 * t + r] = hit ? H(sc.sq) : H(0.f);
            } else {
              const float v = __builtin_amdgcn_exp2f(acc[t][c][r] - Ms) * rsv;
              const bool keep = v >= qminS;
              terms += (unsigned)__popcll(__ballot(keep));
              aqp[c][4 * t + r] = H(keep ? v : 0.f);
            }
          }
        }
      pendImg = img;
      pendLast = lastOfImg;
      pendBuf = buf;
    }
    li = nli;
    gi = ngi;
    ++n;
  }
}

// Fisher vectors from the fused kernel's raw accumulators (fisher.c:419-485 in
// FP64, the same arithmetic as fisher_slice_finalize): workgroup per image;
// the per-component factors S0, 1/(n sqrt(pi)), 1/(n sqrt(2 pi)) in LDS; with
// the L2 normalisation the entries are computed twice, once for the squared
// norm (FP64, fixed order) and once to be divided by it, so that each raw
// value is read twice and each result written once.
constexpr int FF_THREADS = 512;
__global__ __launch_bounds__(FF_THREADS) void mfma_fused_finalize_kernel(
    float* __restrict__ enc, const float* __restrict__ s0raw, const MfmaScales* __restrict__ scp,
    const float* __restrict__ mup, const float* __restrict__ isg, const float* __restrict__ sinv,
    const float* __restrict__ priors, const uint64_t* __restrict__ offsets, int D, int Dp, int K, int flags) {
  extern __shared__ double kf[];  // [S0 | u factor | v factor] x K; u factor < 0: component skipped
  __shared__ double red[FF_THREADS / 64];
  __shared__ float nsh;
  const int img = blockIdx.x, tid = threadIdx.x;
  const long long nImg = (long long)offsets[img + 1] - (long long)offsets[img];
  if (nImg == 0) {
    // the fused kernel skips images without descriptors: all-zero encodings
    float* e = enc + (size_t)img * 2 * (size_t)K * D;
    for (size_t i = tid; i < 2 * (size_t)K * D; i += FF_THREADS) e[i] = 0.f;
    return;
  }
  const MfmaScales sc = *scp;
  const double invq = 1.0 / (double)sc.sq;
  const double inv1 = invq / (double)sc.s1, inv2 = invq / (double)sc.s2;
  for (int k = tid; k < K; k += FF_THREADS) {
    const double prior = (double)priors[k];
    const bool use = prior >= 1e-6;
    double up = 1.0, vp = 1.0;
    if (nImg > 0 && use) {
      up = 1.0 / ((double)nImg * sqrt(prior));
      vp = 1.0 / ((double)nImg * sqrt(2.0 * prior));
    }
    kf[k] = (double)s0raw[(size_t)img * K + k] * invq;
    kf[K + k] = use ? up : -1.0;
    kf[2 * K + k] = vp;
  }
  __syncthreads();
  const size_t KD = (size_t)K * D;
  float* e = enc + (size_t)img * 2 * KD;
  const bool sqrtFlag = (flags & 1) != 0;
  auto entry = [&](size_t i, float& u, float& v) __attribute__((always_inline)) {
    const int k = (int)(i / (size_t)D), d = (int)(i - (size_t)k * D);
    const double up = kf[K + k];
    u = 0.f;
    v = 0.f;
    if (up >= 0.0) {
      const double S0 = kf[k];
      const double S1 = (double)e[i] * inv1;
      const double S2 = (double)e[KD + i] * inv2;
      const double mu = (double)mup[(size_t)k * Dp + d];
      const double uu = (double)isg[(size_t)k * Dp + d] * (S1 - mu * S0);
      const double vv = (double)sinv[(size_t)k * Dp + d] * (S2 - 2.0 * mu * S1 + mu * mu * S0) - S0;
      u = (float)(uu * up);
      v = (float)(vv * kf[2 * K + k]);
    }
    if (sqrtFlag) { u = signed_sqrt(u); v = signed_sqrt(v); }
  };
  float n = 1.f;
  const bool l2 = (flags & 2) != 0;
  if (l2) {
    double acc = 0.0;
    for (size_t i = tid; i < KD; i += FF_THREADS) {
      float u, v;
      entry(i, u, v);
      acc += (double)u * (double)u + (double)v * (double)v;
    }
    const double tot = block_sum_waves<FF_THREADS / 64>(acc, red);
    if (tid == 0) {
      const float nn = (float)sqrt(tot);
      nsh = (float)(((double)nn > 1e-12) ? (double)nn : 1e-12);
    }
    __syncthreads();
    n = nsh;
  }
  for (size_t i = tid; i < KD; i += FF_THREADS) {
    float u, v;
    entry(i, u, v);
    e[i] = l2 ? u / n : u;
    e[KD + i] = l2 ? v / n : v;
  }
}

// ---------------------------------------------------------------------------

// Bytes of group records materialised at once (batches of whole images).
constexpr size_t kRecordBudget = size_t(8) << 30;
// Rows of a batch on which the learned gate is tried before it runs on the rest
// (2048 groups of 32 descriptors decide whether it serves at least 1% of them).
constexpr size_t kGateProbeRows = 16384;

// Device tables of one encode, carved from one buffer:
// [scales][g Dp][bias Kp][mup, isg, sinv Kp*Dp][Whi, Wlo Kp*2Dp][Wst Kp*2(2Dp+8)].
template <typename H>
struct MfmaTables {
  MfmaScales* scp;
  float *g, *bias, *mup, *isg, *sinv;
  H *Whi, *Wlo, *Wst;
};

static inline size_t al256(size_t b) { return (b + 255) & ~size_t(255); }

template <typename H>
static size_t mfma_tables_bytes(int Dp, int Kp) {
  return al256(sizeof(MfmaScales)) + al256(sizeof(float) * Dp) + al256(sizeof(float) * Kp) +
         3 * al256(sizeof(float) * (size_t)Kp * Dp) + 2 * al256(sizeof(H) * (size_t)Kp * 2 * Dp) +
         al256(sizeof(H) * (size_t)Kp * 2 * (2 * Dp + 8));
}

template <typename H>
static MfmaTables<H> mfma_tables_carve(unsigned char* base, int Dp, int Kp) {
  const size_t bS = al256(sizeof(MfmaScales)), bG = al256(sizeof(float) * Dp), bB = al256(sizeof(float) * Kp),
               bT = al256(sizeof(float) * (size_t)Kp * Dp), bW = al256(sizeof(H) * (size_t)Kp * 2 * Dp);
  MfmaTables<H> t;
  t.scp = (MfmaScales*)base;
  t.g = (float*)(base + bS);
  t.bias = (float*)(base + bS + bG);
  t.mup = (float*)(base + bS + bG + bB);
  t.isg = t.mup + bT / sizeof(float);
  t.sinv = t.isg + bT / sizeof(float);
  t.Whi = (H*)(base + bS + bG + bB + 3 * bT);
  t.Wlo = (H*)(base + bS + bG + bB + 3 * bT + bW);
  t.Wst = (H*)(base + bS + bG + bB + 3 * bT + 2 * bW);
  return t;
}

// Zero fill by a kernel (capturable in a graph, unlike a memset node).
__global__ void mfma_zero_kernel(unsigned* __restrict__ p, size_t n) {
  const size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0u;
}
static void zero_words(void* p, size_t words) {
  if (words == 0) return;
  VLHIP_LAUNCH("mfma_zero", mfma_zero_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0, (unsigned*)p, words);
}

template <typename H>
static void mfma_tables_launch(const MfmaTables<H>& t, const float* means, const float* covs, const float* priors,
                               int D, int Dp, int K, int Kp) {
  zero_words(&t.scp->bound, 1);
  VLHIP_LAUNCH("mfma_center", (mfma_center_kernel<float>), dim3(Dp), dim3(256), 0, means, priors, D, K, t.g);
  VLHIP_LAUNCH("mfma_bound", (mfma_bound_kernel<float>), dim3((unsigned)(((size_t)K * D + 255) / 256)), dim3(256), 0,
               means, covs, D, K, t.g, t.scp);
  VLHIP_LAUNCH("mfma_tables", (mfma_tables_kernel<float, H>), dim3(Kp), dim3(64), 0, means, covs, priors, D, Dp, K,
               Kp, t.g, t.scp, t.Whi, t.Wlo, t.Wst, t.bias, t.mup, t.isg, t.sinv);
}

// The two passes over one batch of images [b0, b0 + n) whose descriptors start
// at xb (nG groups, group table gtab with batch-relative rows).
template <typename H, int NS>
static void mfma_first_pass(const MfmaTables<H>& t, const float* xb, int D, int Kp, int flags, unsigned char* rec,
                            size_t nG, const int4* gtab) {
  if (nG == 0) return;
  constexpr int LG = LseGroups<NS>::value;
  const unsigned lb = (unsigned)((nG + LG - 1) / LG);
  if (flags & 4) {
    VLHIP_LAUNCH("mfma_lse", (mfma_lse_kernel<H, NS, true>), dim3(lb), dim3(LG * 64), 0, xb, D, Kp, t.Wst,
                 t.bias, t.scp, t.g, rec, (long long)nG, gtab, (const unsigned char*)nullptr);
  } else {
    VLHIP_LAUNCH("mfma_lse", (mfma_lse_kernel<H, NS, false>), dim3(lb), dim3(LG * 64), 0, xb, D, Kp, t.Wst,
                 t.bias, t.scp, t.g, rec, (long long)nG, gtab, (const unsigned char*)nullptr);
  }
}

template <typename H, int NS, int CT, int WAVES>
static void mfma_second_pass(const MfmaTables<H>& t, int D, int K, int flags, unsigned char* rec,
                             const uint64_t* offsets, const long long* gstart, int b0, int n, const float* priors,
                             float* enc, double* normpart, unsigned long long* numTerms, int numKB,
                             const unsigned char* gflag, const int2* list, int M, size_t shmem, int sched = 0) {
  const int padImages = n >= 8 ? ((n + 7) / 8) * 8 : n;
  const dim3 grid((unsigned)((size_t)padImages * numKB));
  if constexpr (CT == 2 && WAVES == 4) {
    if (gflag == nullptr) {
      // 128-component blocks without the gate: the paired kernel on eight waves
      const size_t sh = pair_lds_bytes<NS>();
      if (flags & 4)
        VLHIP_LAUNCH("mfma_fisher", (mfma_fisher_pair_kernel<float, H, NS, true, false>), grid, dim3(512), sh, rec,
                     offsets, gstart, b0, n, D, K, t.Whi, t.Wlo, t.bias, t.scp, t.mup, t.isg, t.sinv, priors, flags,
                     enc, normpart, numTerms, numKB, 0.f, (double*)nullptr, sched);
      else
        VLHIP_LAUNCH("mfma_fisher", (mfma_fisher_pair_kernel<float, H, NS, false, false>), grid, dim3(512), sh, rec,
                     offsets, gstart, b0, n, D, K, t.Whi, t.Wlo, t.bias, t.scp, t.mup, t.isg, t.sinv, priors, flags,
                     enc, normpart, numTerms, numKB, 0.f, (double*)nullptr, sched);
      return;
    }
  }
  if constexpr (pipe_fits<NS>()) {
    if (gflag == nullptr) {
      // no gate-served group: the pipelined kernel (same encodings)
      const size_t sh = pipe_lds_bytes<NS>();
      if (flags & 4)
        VLHIP_LAUNCH("mfma_fisher", (mfma_fisher_pipe_kernel<float, H, NS, CT, WAVES, true>), grid, dim3(WAVES * 64),
                     sh, rec, offsets, gstart, b0, n, D, K, t.Whi, t.Wlo, t.bias, t.scp, t.mup, t.isg, t.sinv, priors,
                     flags, enc, normpart, numTerms, numKB, sched);
      else
        VLHIP_LAUNCH("mfma_fisher", (mfma_fisher_pipe_kernel<float, H, NS, CT, WAVES, false>), grid, dim3(WAVES * 64),
                     sh, rec, offsets, gstart, b0, n, D, K, t.Whi, t.Wlo, t.bias, t.scp, t.mup, t.isg, t.sinv, priors,
                     flags, enc, normpart, numTerms, numKB, sched);
      return;
    }
  }
  VLHIP_LAUNCH("mfma_fisher", (mfma_fisher_kernel<float, H, NS, CT, WAVES>), grid, dim3(WAVES * 64), shmem, rec,
               offsets, gstart, b0, n, D, K, t.Whi, t.Wlo, t.bias, t.scp, t.mup, t.isg, t.sinv, priors, flags, enc,
               normpart, numTerms, numKB, gflag, list, M, 0.f, nullptr, sched);
}

// Dynamic LDS limits of the second-pass kernels of one tiling (shmem: that of
// mfma_fisher_kernel, which serves gated launches).
template <typename H, int NS, int CT, int WAVES>
static void second_pass_attrs(size_t shmem) {
  if constexpr (CT == 2 && WAVES == 4) {
    const int sh = (int)pair_lds_bytes<NS>();
    VLHIP_CHECK(hipFuncSetAttribute((const void*)mfma_fisher_pair_kernel<float, H, NS, true, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, sh));
    VLHIP_CHECK(hipFuncSetAttribute((const void*)mfma_fisher_pair_kernel<float, H, NS, false, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, sh));
  }
  if constexpr (pipe_fits<NS>()) {
    const int sh = (int)pipe_lds_bytes<NS>();
    VLHIP_CHECK(hipFuncSetAttribute((const void*)mfma_fisher_pipe_kernel<float, H, NS, CT, WAVES, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, sh));
    VLHIP_CHECK(hipFuncSetAttribute((const void*)mfma_fisher_pipe_kernel<float, H, NS, CT, WAVES, false>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, sh));
  }
  VLHIP_CHECK(hipFuncSetAttribute((const void*)mfma_fisher_kernel<float, H, NS, CT, WAVES>,
                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
}

template <typename H, int NS, int CT, int WAVES>
static void mfma_batch_launch(const MfmaTables<H>& t, const float* xb, int D, int K, int Kp, int flags,
                              unsigned char* rec, size_t nG, const int4* gtab, const uint64_t* offsets,
                              const long long* gstart, int b0, int n, const float* priors, float* enc,
                              double* normpart, unsigned long long* numTerms, int numKB, size_t shmem,
                              int sched = 0) {
  mfma_first_pass<H, NS>(t, xb, D, Kp, flags, rec, nG, gtab);
  mfma_second_pass<H, NS, CT, WAVES>(t, D, K, flags, rec, offsets, gstart, b0, n, priors, enc, normpart, numTerms,
                                     numKB, nullptr, nullptr, 0, shmem, sched);
}

template <typename H, int NS, int CT, int WAVES>
static size_t mfma_shmem(bool gated) {
  return 2 * GroupRec<NS>::kBytes + (gated ? (size_t)GR * (WAVES * 16 * CT) * sizeof(float) : 0);
}

template <typename H, int NS, int CT, int WAVES>
static void run_mfma(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                     const float* x, const uint64_t* offsets, size_t numImages, int flags,
                     unsigned long long* numTerms, const std::vector<uint64_t>& off, const GateUse* gu) {
  constexpr int Dp = NS * 16;
  constexpr int BC = WAVES * 16 * CT;
  typedef GroupRec<NS> R;
  const int Kp = (int)((K + BC - 1) / BC) * BC;
  unsigned char* base = (unsigned char*)scratch(SLOT_PREP, mfma_tables_bytes<H>(Dp, Kp));
  const MfmaTables<H> t = mfma_tables_carve<H>(base, Dp, Kp);
  mfma_tables_launch<H>(t, means, covs, priors, (int)D, Dp, (int)K, Kp);

  const int numKB = Kp / BC;
  double* normpart = (double*)scratch(SLOT_PARTIAL, numImages * (size_t)numKB * sizeof(double));
  if (numTerms) zero_words(numTerms, 2 * numImages);
  // two group records, plus the posterior tile of gate-served groups
  const size_t shmem = mfma_shmem<H, NS, CT, WAVES>(gu != nullptr);
  second_pass_attrs<H, NS, CT, WAVES>(shmem);
  // image-aligned groups: gstart[i] = sum_{j<i} ceil(n_j / GR)
  std::vector<long long> gs(numImages + 1);
  gs[0] = 0;
  for (size_t i = 0; i < numImages; ++i) gs[i + 1] = gs[i] + (long long)((off[i + 1] - off[i] + GR - 1) / GR);
  long long* gstart = (long long*)scratch(SLOT_GROUPS, (numImages + 1) * sizeof(long long));
  upload(gstart, gs.data(), (numImages + 1) * sizeof(long long));
  sync();  // gs is a local buffer
  bool gateOff = false;  // a probe switched the gate off for the rest of this call
  size_t b0 = 0;
  while (b0 < numImages) {
    size_t b1 = b0 + 1;
    while (b1 < numImages && (size_t)(gs[b1 + 1] - gs[b0]) * R::kBytes <= kRecordBudget) ++b1;
    const uint64_t rowBase = off[b0], rows = off[b1] - off[b0];
    VLHIP_REQUIRE(rows < (uint64_t(1) << 31), "an image batch of the matrix-core encoder exceeds 2^31 descriptors");
    const size_t nG = (size_t)(gs[b1] - gs[b0]);
    unsigned char* rec = (unsigned char*)scratch(SLOT_FISHER_WORK, (nG ? nG : 1) * R::kBytes);
    const int n = (int)(b1 - b0);
    unsigned char* gflag = nullptr;
    int2* list = nullptr;
    int4* gtab = nullptr;
    if (nG > 0) {
      gtab = (int4*)scratch(SLOT_GROUPTAB, nG * sizeof(int4));
      VLHIP_LAUNCH("mfma_group_table", mfma_group_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                   offsets, gstart, (int)b0, n, gtab);
    }
    // first pass over every group: the exact row statistics, which the second
    // pass normalises with and which judge the gate's candidates
    mfma_first_pass<H, NS>(t, x + rowBase * D, (int)D, Kp, flags, rec, nG, gtab);
    if (nG > 0 && gu) {
      ctx().gate_groups += nG;
      if (!gateOff) {
        // gate: candidates and the network's pre-screen per row, then the
        // groups whose candidates carry an exact mass >= need in every row
        int* cand = (int*)scratch(SLOT_GATE, (size_t)rows * gu->M * sizeof(int));
        unsigned char* gb = (unsigned char*)scratch(SLOT_GATE2, al256(rows * sizeof(float)) + al256(rows) + al256(nG));
        float* mass = (float*)gb;
        unsigned char* conf = gb + al256(rows * sizeof(float));
        gflag = conf + al256(rows);
        list = (int2*)scratch(SLOT_GATELIST, nG * GR * gu->M * sizeof(int2));
        auto sparse = [&](unsigned long long* terms) {
          VLHIP_LAUNCH("mfma_sparse", mfma_sparse_kernel<NS>, dim3((unsigned)((nG + 3) / 4)), dim3(256), 0,
                       x + rowBase * D, (int)D, (long long)nG, gtab, conf, cand, gu->M, t.bias, t.mup, t.sinv, t.g,
                       t.scp, rec, gu->need, flags, gflag, list, terms);
        };
        // The gate first runs on a probe of the batch's first rows; when it
        // serves fewer than 1% of the groups that lie inside the probe, the
        // rest of the batch (and of the call) takes the dense path without
        // running the network: only the gate's cost changes.
        const size_t probe = std::min<size_t>(rows, kGateProbeRows);
        vlhip_gate_select(gu->gate, x + rowBase * D, probe, gu->M, gu->need, cand, mass, conf);
        bool on = true;
        if (probe < rows) {
          VLHIP_CHECK(hipMemsetAsync(conf + probe, 0, rows - probe, ctx().stream));
          sparse(nullptr);
          // only the groups of the images that start inside the probe are read back
          size_t iEnd = b0, gEnd = 0;
          while (iEnd < b1 && off[iEnd] - rowBase < probe) {
            gEnd += (size_t)(gs[iEnd + 1] - gs[iEnd]);
            ++iEnd;
          }
          std::vector<unsigned char> hf(gEnd ? gEnd : 1);
          if (gEnd) download(hf.data(), gflag, gEnd);
          size_t inProbe = 0, servedProbe = 0, gi = 0;
          for (size_t i = b0; i < iEnd; ++i) {
            const uint64_t a = off[i] - rowBase, e = off[i + 1] - rowBase;
            for (uint64_t r = a; r < e; r += GR, ++gi) {
              if (std::min<uint64_t>(e, r + GR) > probe) continue;
              ++inProbe;
              servedProbe += hf[gi];
            }
          }
          if (inProbe >= 64 && servedProbe * 100 < inProbe) {
            on = false;
            gateOff = true;
          } else {
            vlhip_gate_select(gu->gate, x + (rowBase + probe) * D, rows - probe, gu->M, gu->need,
                              cand + probe * gu->M, mass + probe, conf + probe);
          }
        }
        size_t served = 0;
        if (on) {
          sparse(numTerms);
          std::vector<unsigned char> hf(nG);
          download(hf.data(), gflag, nG);
          for (size_t i = 0; i < nG; ++i) served += hf[i];
        }
        ctx().gate_served += served;
        if (served == 0) {
          // every group of the batch is dense: the plain dense kernels, which give
          // the same encodings without reading a gate flag per group
          gflag = nullptr;
          list = nullptr;
        }
      }
    }
    mfma_second_pass<H, NS, CT, WAVES>(t, (int)D, (int)K, flags, rec, offsets, gstart, (int)b0, n, priors, enc,
                                       normpart, numTerms, numKB, gflag, list, gu ? gu->M : 0, shmem);
    b0 = b1;
  }
  if (flags & 2) {
    launch_fisher_normalize(enc, normpart, numKB, numImages, (size_t)2 * K * D);
  }
}

// Fused single-pass encoder (mfma_fused_kernel): one cooperative launch over
// all images.  Returns false (nothing launched) when the shape or the device
// does not allow it or, in the automatic mode, when there are fewer images
// than streams (the two-pass encoder spreads a few images over more of the
// device) or more than one component block per image: with partners, the
// per-group exchange and the fragments every partner builds cost more than
// the first pass saves (K = 1024, D = 80, 2048 images of 3000 descriptors:
// 21.2 ms fused against 15.7 ms in two passes; K = 256: 4.7 against 5.5 ms).
static int g_fused_mode = 1;  // 0: never, 1: automatic, 2: whenever the shape allows

template <typename H, int NS, int CT, int WAVES>
static bool run_fused(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                      const float* x, const uint64_t* offsets, size_t numImages, int flags,
                      unsigned long long* numTerms, const std::vector<uint64_t>& off) {
  constexpr int Dp = NS * 16;
  constexpr int BC = WAVES * 16 * CT;
  constexpr int NT = WAVES * 64;
  typedef GroupRec<NS> R;
  if (g_fused_mode == 0) return false;
  const int Kp = (int)((K + BC - 1) / BC) * BC;
  const int numKB = Kp / BC;
  if (K > 2048) return false;  // the finalisation keeps 3 K doubles in LDS
  const int spx = (ctx().num_cus / 8) / numKB;  // streams per XCD
  if (spx < 1) return false;
  const int S = 8 * spx, G = S * numKB;
  if (g_fused_mode == 1 && (numImages < (size_t)S || numKB > 1)) return false;
  if (numImages >= (size_t(1) << 31)) return false;
  const unsigned long long xBytes = (unsigned long long)off[numImages] * D * sizeof(float);
  const bool dma = (D % 4) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0 && xBytes >= 16;
  const bool fast = (flags & 4) != 0;
  const void* kern = fast ? (dma ? (const void*)mfma_fused_kernel<H, NS, CT, WAVES, true, true>
                                 : (const void*)mfma_fused_kernel<H, NS, CT, WAVES, true, false>)
                          : (dma ? (const void*)mfma_fused_kernel<H, NS, CT, WAVES, false, true>
                                 : (const void*)mfma_fused_kernel<H, NS, CT, WAVES, false, false>);
  const size_t shmem = 2 * (R::kA + R::kB);
  VLHIP_CHECK(hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  int occ = 0, coop = 0;
  VLHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, kern, NT, shmem));
  VLHIP_CHECK(hipDeviceGetAttribute(&coop, hipDeviceAttributeCooperativeLaunch, ctx().device));
  if (!coop || occ < 1 || G > occ * ctx().num_cus) return false;

  unsigned char* base = (unsigned char*)scratch(SLOT_PREP, mfma_tables_bytes<H>(Dp, Kp));
  const MfmaTables<H> t = mfma_tables_carve<H>(base, Dp, Kp);
  mfma_tables_launch<H>(t, means, covs, priors, (int)D, Dp, (int)K, Kp);
  VLHIP_LAUNCH("mfma_shift", mfma_shift_kernel, dim3(1), dim3(256), 0, t.bias, Kp, (int)D, t.scp);
  float* s0raw = (float*)scratch(SLOT_PARTIAL, numImages * K * sizeof(float));
  if (numTerms) zero_words(numTerms, 2 * numImages);

  // streams: images by size, largest first, each to the least-loaded stream
  std::vector<int> order(numImages);
  for (size_t i = 0; i < numImages; ++i) order[i] = (int)i;
  std::stable_sort(order.begin(), order.end(),
                   [&](int a, int b) { return off[a + 1] - off[a] > off[b + 1] - off[b]; });
  std::vector<std::vector<int>> lists(S);
  std::vector<unsigned long long> load(S, 0);
  for (int img : order) {
    int best = 0;
    for (int s = 1; s < S; ++s)
      if (load[s] < load[best]) best = s;
    lists[best].push_back(img);
    load[best] += (off[img + 1] - off[img] + GR - 1) / GR + 1;  // groups, plus the image's finalisation
  }
  std::vector<long long> hOff(S + 1);
  std::vector<int> hImg;
  hImg.reserve(numImages);
  hOff[0] = 0;
  for (int s = 0; s < S; ++s) {
    hImg.insert(hImg.end(), lists[s].begin(), lists[s].end());
    hOff[s + 1] = (long long)hImg.size();
  }
  unsigned char* sb = (unsigned char*)scratch(SLOT_STREAMS, al256((S + 1) * sizeof(long long)) +
                                                                al256(numImages * sizeof(int)));
  long long* dOff = (long long*)sb;
  int* dImg = (int*)(sb + al256((S + 1) * sizeof(long long)));
  upload(dOff, hOff.data(), (S + 1) * sizeof(long long));
  upload(dImg, hImg.data(), numImages * sizeof(int));
  // exchange slots (tags start at 1, so zero means unwritten) and the abort flag
  const size_t xw = (size_t)S * numKB * FX_RING * 3 * GR;
  unsigned char* xb = (unsigned char*)scratch(SLOT_FUSED, xw * 8 + 256);
  zero_words(xb, (xw * 8 + 256) / 4);
  unsigned long long* xch = (unsigned long long*)xb;
  unsigned* abortFlag = (unsigned*)(xb + xw * 8);
  sync();  // the host lists are local

  int Di = (int)D, Ki = (int)K;
  unsigned long long xb64 = xBytes;
  void* args[] = {(void*)&x,    (void*)&xb64,  (void*)&offsets, (void*)&dOff,  (void*)&dImg,
                  (void*)&numKB, (void*)&Di,   (void*)&Ki,      (void*)&t.Whi, (void*)&t.Wlo,
                  (void*)&t.bias, (void*)&t.scp, (void*)&t.g,   (void*)&enc,   (void*)&s0raw,
                  (void*)&numTerms, (void*)&xch, (void*)&abortFlag};
  hipEvent_t ts = nullptr;
  const int ti = timing_begin("mfma_fused", &ts);
  VLHIP_CHECK(hipLaunchCooperativeKernel(kern, dim3((unsigned)G), dim3(NT), args, (unsigned)shmem, ctx().stream));
  if (ti >= 0) timing_end(ti, ts);
  const size_t fsh = 3 * (size_t)K * sizeof(double);
  VLHIP_CHECK(hipFuncSetAttribute((const void*)mfma_fused_finalize_kernel, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  (int)fsh));
  VLHIP_LAUNCH("mfma_fused_finalize", mfma_fused_finalize_kernel, dim3((unsigned)numImages), dim3(FF_THREADS), fsh,
               enc, s0raw, t.scp, t.mup, t.isg, t.sinv, priors, offsets, (int)D, Dp, (int)K, flags);
  unsigned ab = 0;
  download(&ab, abortFlag, sizeof(unsigned));
  VLHIP_REQUIRE(ab == 0, "the fused matrix-core encoder's exchange between component blocks timed out");
  return true;
}

// Component tiling: 8 waves x 32 components for large mixtures, 4 waves x 32
// components when K is small (less padding) or D > 80 (registers).
template <typename H, int NS>
static void run_mfma_k(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                       const float* x, const uint64_t* offsets, size_t numImages, int flags,
                       unsigned long long* numTerms, const std::vector<uint64_t>& off, const GateUse* gu) {
  if constexpr (NS <= 5) {
    if (K > 128) {
      if (!gu && run_fused<H, NS, 2, 8>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off))
        return;
      run_mfma<H, NS, 2, 8>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off, gu);
      return;
    }
  }
  run_mfma<H, NS, 2, 4>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off, gu);
}

template <typename H>
static void mfma_dispatch(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                          const float* x, const uint64_t* offsets, size_t numImages, int flags,
                          unsigned long long* numTerms, const GateUse* gu) {
  std::vector<uint64_t> off(numImages + 1);
  download(off.data(), offsets, (numImages + 1) * sizeof(uint64_t));
  const size_t Dp = (D + 15) / 16 * 16;
  switch (Dp) {
    case 16: run_mfma_k<H, 1>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off, gu); break;
    case 32: run_mfma_k<H, 2>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off, gu); break;
    case 48: run_mfma_k<H, 3>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off, gu); break;
    case 64: run_mfma_k<H, 4>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off, gu); break;
    case 80: run_mfma_k<H, 5>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off, gu); break;
    case 96: run_mfma_k<H, 6>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off, gu); break;
    case 112: run_mfma_k<H, 7>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off, gu); break;
    case 128: run_mfma_k<H, 8>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off, gu); break;
    default: fatal("vlhip_fisher_encode_batch", __FILE__, __LINE__,
                   "the matrix-core Fisher encoder supports dimension <= 128");
  }
}

// ---------------------------------------------------------------------------
// Encode plans (batched uneven images, graph replay).  Everything that
// depends only on the image sizes is prepared once: the image-aligned group
// table of every batch, the batches (<= 8 GiB of records), device buffers,
// and the component tiling, chosen by timing the candidate tilings on the
// real shapes (autotuner, cached per shape).  A run is then a pure sequence of
// kernel launches -- zero fills, tables, both passes per batch, the
// normalisation and, optionally, the nearest-mean quantizer and the hard
// VLAD -- captured into a hipGraph on the first run and replayed while the
// argument pointers stay the same.

struct PlanBatch {
  size_t b0, n;       // images [b0, b0 + n)
  uint64_t rowBase;   // first descriptor of the batch
  size_t nG;          // groups of the batch
  long long gOff;     // first group of the batch in the plan's group table
};

struct EncodePlan {
  size_t D = 0, K = 0, numImages = 0, totalRows = 0;
  int flags = 0, vladFlags = -1, NS = 0;
  int tiling = -1;             // chosen component tiling (kPlanTilings index) ...
  int sched = 0;               // ... and component-block schedule (fisher_mfma_coords)
  int chosen = -1;             // index of the chosen candidate in tried (-1: cached choice)
  std::vector<std::array<double, 3>> tried;  // autotuner candidates {tiling, schedule, ms}
  std::vector<PlanBatch> batches;
  size_t maxGroups = 0;
  uint64_t* dOff = nullptr;
  long long* dGstart = nullptr;
  int4* dGtab = nullptr;
  unsigned char* dRec = nullptr;
  unsigned char* dTables = nullptr;
  double* dNorm = nullptr;
  uint32_t* dAssign = nullptr;
  unsigned char* dQtab = nullptr;  // centre table of the screened quantizer (hard VLAD), plan-owned
  hipStream_t stream = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  const void* key[7] = {};
  // optional load stage: descriptors of Din dimensions projected to D
  size_t Din = 0;
  int renorm = 0;
  float* dProjP = nullptr;  // D x Din, owned copy
  float* dProjC = nullptr;  // Din, owned copy
  float* dProjX = nullptr;  // totalRows x D, the projected descriptors
};

// Component tilings of the second pass the autotuner tries: CT tiles of 16
// components per wave, WAVES waves per workgroup (a workgroup owns 16 CT WAVES
// components).  {2, 8} needs D <= 80 (registers) and is tried only when
// K > 128; the others for every shape.
struct PlanTiling {
  int ct, waves;
};
constexpr PlanTiling kPlanTilings[4] = {{2, 8}, {2, 4}, {1, 8}, {1, 4}};
constexpr int kPlanSchedules = 3;
constexpr int kPlanMinBlock = 64;  // smallest 16 CT WAVES of the tilings

static bool plan_tiling_allowed(int tiling, int NS, size_t K) {
  return tiling != 0 || (NS <= 5 && K > 128);
}

// autotuner cache: (D, K, Fisher flags, VLAD flags, load-stage input dimension,
// renormalisation, rows bucket, images bucket) -> (tiling, schedule)
static std::vector<std::pair<std::vector<long long>, std::pair<int, int>>>& tuning_cache() {
  static std::vector<std::pair<std::vector<long long>, std::pair<int, int>>> c;
  return c;
}

template <typename H, int NS, int CT, int WAVES>
static void plan_enqueue(EncodePlan& P, int sched, const float* x, const float* means, const float* covs,
                         const float* priors, float* enc, float* vladEnc, unsigned long long* numTerms) {
  constexpr int Dp = NS * 16;
  constexpr int BC = WAVES * 16 * CT;
  const int Kp = (int)((P.K + BC - 1) / BC) * BC;
  const int numKB = Kp / BC;
  const MfmaTables<H> t = mfma_tables_carve<H>(P.dTables, Dp, Kp);
  mfma_tables_launch<H>(t, means, covs, priors, (int)P.D, Dp, (int)P.K, Kp);
  if (numTerms) zero_words(numTerms, 2 * P.numImages);
  const size_t shmem = mfma_shmem<H, NS, CT, WAVES>(false);
  for (const PlanBatch& b : P.batches)
    mfma_batch_launch<H, NS, CT, WAVES>(t, x + b.rowBase * P.D, (int)P.D, (int)P.K, Kp, P.flags, P.dRec, b.nG,
                                        P.dGtab + b.gOff, P.dOff, P.dGstart, (int)b.b0, (int)b.n, priors, enc,
                                        P.dNorm, numTerms, numKB, shmem, sched);
  if (P.flags & 2) launch_fisher_normalize(enc, P.dNorm, numKB, P.numImages, (size_t)2 * P.K * P.D);
  if (P.vladFlags >= 0 && vladEnc) {
    // the plan's own centre table: no library buffer is captured into the graph
    quantize_t<float>(P.dAssign, nullptr, x, means, P.totalRows, P.D, P.K, 8, P.dQtab);
    vlad_t<float>(vladEnc, means, P.D, P.K, x, P.dOff, P.numImages, nullptr, P.dAssign, P.vladFlags);
  }
}

template <typename H, int NS>
static void plan_enqueue_variant(EncodePlan& P, int tiling, int sched, const float* x, const float* means,
                                 const float* covs, const float* priors, float* enc, float* vladEnc,
                                 unsigned long long* numTerms) {
  switch (tiling) {
    case 0:
      if constexpr (NS <= 5) {
        plan_enqueue<H, NS, 2, 8>(P, sched, x, means, covs, priors, enc, vladEnc, numTerms);
        return;
      }
      break;
    case 2: plan_enqueue<H, NS, 1, 8>(P, sched, x, means, covs, priors, enc, vladEnc, numTerms); return;
    case 3: plan_enqueue<H, NS, 1, 4>(P, sched, x, means, covs, priors, enc, vladEnc, numTerms); return;
    default: break;
  }
  plan_enqueue<H, NS, 2, 4>(P, sched, x, means, covs, priors, enc, vladEnc, numTerms);
}

static void plan_enqueue_any(EncodePlan& P, int tiling, int sched, const float* x, const float* means,
                             const float* covs, const float* priors, float* enc, float* vladEnc,
                             unsigned long long* numTerms) {
  if (P.Din) {
    // load stage: the encoders below read the projected descriptors
    project_t(P.dProjX, x, P.totalRows, P.Din, P.D, P.dProjP, P.dProjC, P.renorm);
    x = P.dProjX;
  }
  switch (P.NS) {
    case 1: plan_enqueue_variant<_Float16, 1>(P, tiling, sched, x, means, covs, priors, enc, vladEnc, numTerms); break;
    case 2: plan_enqueue_variant<_Float16, 2>(P, tiling, sched, x, means, covs, priors, enc, vladEnc, numTerms); break;
    case 3: plan_enqueue_variant<_Float16, 3>(P, tiling, sched, x, means, covs, priors, enc, vladEnc, numTerms); break;
    case 4: plan_enqueue_variant<_Float16, 4>(P, tiling, sched, x, means, covs, priors, enc, vladEnc, numTerms); break;
    case 5: plan_enqueue_variant<_Float16, 5>(P, tiling, sched, x, means, covs, priors, enc, vladEnc, numTerms); break;
    case 6: plan_enqueue_variant<_Float16, 6>(P, tiling, sched, x, means, covs, priors, enc, vladEnc, numTerms); break;
    case 7: plan_enqueue_variant<_Float16, 7>(P, tiling, sched, x, means, covs, priors, enc, vladEnc, numTerms); break;
    default: plan_enqueue_variant<_Float16, 8>(P, tiling, sched, x, means, covs, priors, enc, vladEnc, numTerms); break;
  }
}

template <int NS, int CT, int WAVES>
static void plan_set_attr() {
  second_pass_attrs<_Float16, NS, CT, WAVES>(mfma_shmem<_Float16, NS, CT, WAVES>(false));
}
template <int NS>
static void plan_set_attrs_ns() {
  if constexpr (NS <= 5) plan_set_attr<NS, 2, 8>();
  plan_set_attr<NS, 2, 4>();
  plan_set_attr<NS, 1, 8>();
  plan_set_attr<NS, 1, 4>();
}
static void plan_set_attrs(int NS) {
  switch (NS) {
    case 1: plan_set_attrs_ns<1>(); break;
    case 2: plan_set_attrs_ns<2>(); break;
    case 3: plan_set_attrs_ns<3>(); break;
    case 4: plan_set_attrs_ns<4>(); break;
    case 5: plan_set_attrs_ns<5>(); break;
    case 6: plan_set_attrs_ns<6>(); break;
    case 7: plan_set_attrs_ns<7>(); break;
    default: plan_set_attrs_ns<8>(); break;
  }
}

static size_t plan_record_bytes(int NS) {
  switch (NS) {
    case 1: return GroupRec<1>::kBytes;
    case 2: return GroupRec<2>::kBytes;
    case 3: return GroupRec<3>::kBytes;
    case 4: return GroupRec<4>::kBytes;
    case 5: return GroupRec<5>::kBytes;
    case 6: return GroupRec<6>::kBytes;
    case 7: return GroupRec<7>::kBytes;
    default: return GroupRec<8>::kBytes;
  }
}

void fisher_split16(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                    const float* x, const uint64_t* offsets, size_t numImages, int flags,
                    unsigned long long* numTerms, bool fp16, const GateUse* gu) {
  VLHIP_REQUIRE(K >= 1 && D >= 1, "numClusters and dimension must be positive");
  if (numImages == 0) return;
  if (gu) VLHIP_REQUIRE(gu->M >= 1 && gu->M <= 16 && (size_t)gu->M <= K, "gate candidates must be in [1, min(16, K)]");
  if (fp16) mfma_dispatch<_Float16>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, gu);
  else mfma_dispatch<__bf16>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, gu);
}

// ---------------------------------------------------------------------------
// FP32 matrix-core encoder (VLHIP_MODE_FP32_MFMA).  The two contractions of
// this file with plain FP32 operands on v_mfma_f32_16x16x4_f32 (exact FP32
// products, FP32 accumulation; no split, no scaling: the tables are those of
// the compensated encoder built with float in place of the 16-bit type, whose
// low parts are zero).  Two passes as in the compensated encoder, but without
// operand records: the first pass writes each group's row statistics (-max
// logit, 1 / sum, argmax: 384 bytes per group of 32 descriptors) and the
// second rebuilds its operands from the descriptors, staged in LDS as
// x' = x - g.  Nothing of size N x K is written.
//
// 16x16x4 operands: lane l supplies A[l & 15][l >> 4] and B[l >> 4][l & 15].
// In the logit contraction, step j of a row tile takes feature 8NS q + j of
// [x'^2 | x'] (2Dp = 32NS features) from the lanes of quarter q = l >> 4, so
// that a lane's weights W[comp][8NS q, 8NS q + 8NS) are one contiguous run of
// the table and its descriptor values are x'[8NS (q & 1), +8NS), squared for
// q < 2.  In the moment contraction, element r of the logit accumulator of row
// tile t (rows 16t + 4q + r, the lane's component) is the A operand as it
// stands, and B is x' or x'^2 of the same rows: one product per 4 rows.

constexpr int F32_STATS = 3 * GR;  // floats of row statistics per group
template <int NS> struct F32LseWaves { static constexpr int value = NS <= 5 ? 8 : 4; };

__device__ __forceinline__ f32x4 mfma4(float a, float b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
}

// Logits of 32 rows x 64 staged components (LDS rows of 32NS + 4 floats):
// 2 row tiles x 4 component tiles x 8NS steps.
template <int NS>
__device__ __forceinline__ void f32_lse_products(f32x4 (&acc)[4][2], const float (&a)[2][8 * NS], const float* w,
                                                 int lr, int lg) {
  constexpr int WS = 32 * NS + 4;
#pragma unroll
  for (int c = 0; c < 4; ++c) acc[c][0] = acc[c][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int j4 = 0; j4 < 2 * NS; ++j4) {
    f32x4 bw[4];
#pragma unroll
    for (int c = 0; c < 4; ++c) bw[c] = *reinterpret_cast<const f32x4*>(w + (16 * c + lr) * WS + 8 * NS * lg + 4 * j4);
#pragma unroll
    for (int e = 0; e < 4; ++e)
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[c][t] = mfma4(a[t][4 * j4 + e], bw[c][e], acc[c][t]);
  }
}

// First pass: wave = one group of 32 rows; the weights of 64 components at a
// time are double-buffered in LDS (the next block fetched into registers while
// the current one is multiplied); a lane-local online log-sum-exp per row,
// merged over the 16 lanes that share the rows at the end.
template <int NS, bool ARG>
__global__ __launch_bounds__(F32LseWaves<NS>::value * 64, 1) void f32_lse_kernel(
    const float* __restrict__ x, int D, int Kp, const float* __restrict__ W, const float* __restrict__ bias,
    const float* __restrict__ g, float* __restrict__ stats, long long numGroups, const int4* __restrict__ gtab) {
  constexpr int Dp = NS * 16;
  constexpr int F = 2 * Dp;
  constexpr int WS = F + 4;
  constexpr int LG = F32LseWaves<NS>::value;
  constexpr int NT = LG * 64;
  constexpr int WORDS = MB_COMPS * F / 4;  // 16-byte words of one weight block
  constexpr int STG = WORDS / NT;
  static_assert(STG * NT == WORDS, "staging must divide evenly");
  __shared__ __align__(16) float wsh[2][MB_COMPS][WS];
  __shared__ float bsh[4][MB_COMPS];
  static_assert(LG * GR * Dp <= 2 * MB_COMPS * WS, "row staging must fit the weight buffers");

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const long long grp = (long long)blockIdx.x * LG + w;
  const bool live = grp < numGroups;  // uniform per wave; dead waves still take part in barriers
  long long row0 = 0, rowEnd = 0;
  if (live) {
    const int4 gr = gtab[grp];
    row0 = gr.x;
    rowEnd = gr.y;
  }
  int srcOff[STG], dstOff[STG];
#pragma unroll
  for (int q = 0; q < STG; ++q) {
    const int i = tid + NT * q;
    const int c = i / (F / 4), q4 = i - c * (F / 4);
    srcOff[q] = c * F + 4 * q4;
    dstOff[q] = c * WS + 4 * q4;
  }
  f32x4 wpre[STG];
#define F32_FETCH(B)                                                                                  \
  {                                                                                                   \
    _Pragma("unroll") for (int q = 0; q < STG; ++q) wpre[q] =                                         \
        *reinterpret_cast<const f32x4*>(W + (size_t)(B) * MB_COMPS * F + srcOff[q]);                  \
    if (tid < MB_COMPS) bpre = bias[(B) * MB_COMPS + tid];                                            \
  }
#define F32_STAGE(B)                                                                                  \
  {                                                                                                   \
    _Pragma("unroll") for (int q = 0; q < STG; ++q)                                                   \
        *reinterpret_cast<f32x4*>(&wsh[(B) & 1][0][0] + dstOff[q]) = wpre[q];                         \
    if (tid < MB_COMPS) bsh[(B) & 3][tid] = bpre;                                                     \
  }
  float bpre = 0.f;
  F32_FETCH(0)

  // the wave's 32 centred rows x' = x - g (zero past D and past the image) in
  // LDS, row stride Dp, inside the weight buffers (free until the sweep)
  float* xsh = &wsh[0][0][0] + w * GR * Dp;
  {
    constexpr int XW = GR * (Dp / 4);
    constexpr int XN = (XW + 63) / 64;
    const bool vecx = (D & 3) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
    f32x4 xv[XN], gv[XN];
#pragma unroll
    for (int u = 0; u < XN; ++u) {
      const int i = lane + 64 * u;
      const int c4 = (i - (i / (Dp / 4)) * (Dp / 4)) * 4;
      gv[u] = i < XW ? *reinterpret_cast<const f32x4*>(g + c4) : f32x4{0.f, 0.f, 0.f, 0.f};
    }
#pragma unroll
    for (int u = 0; u < XN; ++u) {
      const int i = lane + 64 * u;
      const int r = i / (Dp / 4), c4 = (i - r * (Dp / 4)) * 4;
      const long long row = row0 + r;
      f32x4 v = f32x4{0.f, 0.f, 0.f, 0.f};
      if (i < XW && row < rowEnd) {
        if (vecx) {
          if (c4 < D) v = *reinterpret_cast<const f32x4*>(x + row * D + c4);
        } else {
#pragma unroll
          for (int e = 0; e < 4; ++e) v[e] = c4 + e < D ? x[row * D + c4 + e] : 0.f;
        }
      }
      xv[u] = v;
    }
#pragma unroll
    for (int u = 0; u < XN; ++u) {
      const int i = lane + 64 * u;
      if (i >= XW) continue;
      const int r = i / (Dp / 4), c4 = (i - r * (Dp / 4)) * 4;
      f32x4 v = xv[u];
      if (row0 + r < rowEnd) v = v - gv[u];
      *reinterpret_cast<f32x4*>(xsh + r * Dp + c4) = v;
    }
  }
  __syncthreads();
  // A operands of the 2 row tiles, resident for the whole sweep
  float a[2][8 * NS];
  {
    const bool sq = lg < 2;
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int j4 = 0; j4 < 2 * NS; ++j4) {
        const f32x4 v = *reinterpret_cast<const f32x4*>(xsh + (16 * t + lr) * Dp + 8 * NS * (lg & 1) + 4 * j4);
#pragma unroll
        for (int e = 0; e < 4; ++e) a[t][4 * j4 + e] = sq ? v[e] * v[e] : v[e];
      }
  }
  __syncthreads();  // the staging area becomes the weight buffers

  LseState st;
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) { st.m[t][r] = -INFINITY; st.s[t][r] = 0.f; st.bv[t][r] = -INFINITY; st.bi[t][r] = 0; }
  if (!__syncthreads_or(live)) return;
  const int nb = Kp / MB_COMPS;
  F32_STAGE(0)
  if (nb > 1) F32_FETCH(1)
  __syncthreads();
  for (int b = 0; b < nb; ++b) {
    f32x4 acc[4][2];
    if (live) f32_lse_products<NS>(acc, a, &wsh[b & 1][0][0], lr, lg);
    // the other buffer held block b - 1, read before the last barrier
    if (b + 1 < nb) {
      F32_STAGE(b + 1)
      if (b + 2 < nb) F32_FETCH(b + 2)
    }
    if (live) lse_update<ARG>(acc, &bsh[b & 3][0], b * MB_COMPS, lr, st);
    __syncthreads();
  }
#undef F32_FETCH
#undef F32_STAGE

  float* Sf = stats + (size_t)grp * F32_STATS;
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float m = row16_max(st.m[t][r]);
      const float ms = (m == -INFINITY) ? 0.f : m;
      const float s = row16_sum(st.s[t][r] * __builtin_amdgcn_exp2f(st.m[t][r] - ms));
      float bv = st.bv[t][r];
      int bi = st.bi[t][r];
      if (ARG) row16_argmax(bv, bi);
      const int row = 16 * t + 4 * lg + r;
      if (live && lr == 0) {
        const bool pad = row0 + row >= rowEnd;
        Sf[row] = pad ? 0.f : -m;
        Sf[GR + row] = pad ? 0.f : 1.0f / s;
        Sf[2 * GR + row] = __int_as_float(pad ? -1 : bi);
      }
    }
}

// Second pass: workgroup = (image, WAVES x 16CT components); wave w owns CT
// component tiles with their weights in registers and walks the image's
// groups: the group's rows x' and row statistics are staged in LDS (double
// buffered, the next group prefetched into registers), then
//   logits   2 row tiles x CT tiles x 8NS steps, accumulators starting at
//            bias - max;
//   q        2^(logit - max) / sum, FAST one-hot, < 1e-6 dropped
//            (fisher.c:436-441); padding rows have 1 / sum = 0;
//   moments  S0 (q against ones), S1 += q x', S2 += q x'^2, one product per
//            4 rows and 16 features.
// The finalisation is that of the compensated encoder (scales 1).
template <int NS, int CT, int WAVES>
__global__ __launch_bounds__(WAVES * 64, 1) void f32_fisher_kernel(
    const float* __restrict__ x, int D, const float* __restrict__ g, const float* __restrict__ stats,
    const uint64_t* __restrict__ offsets, const long long* __restrict__ gstart, int imgBegin, int numImages, int K,
    const float* __restrict__ W, const float* __restrict__ bias, const MfmaScales* __restrict__ scp,
    const float* __restrict__ mup, const float* __restrict__ isg, const float* __restrict__ sinv,
    const float* __restrict__ priors, int flags, float* __restrict__ enc, double* __restrict__ normpart,
    unsigned long long* __restrict__ numTerms, int numKB) {
  constexpr int Dp = NS * 16;
  constexpr int F = 2 * Dp;
  constexpr int NT = WAVES * 64;
  constexpr int BC = WAVES * 16 * CT;
  constexpr int XS = Dp + 4;              // LDS row stride of a staged group
  constexpr int XWORDS = GR * Dp / 4;     // 16-byte words of a group's rows
  constexpr int NV = XWORDS + F32_STATS / 4;
  constexpr int PER = (NV + NT - 1) / NT;
  constexpr int BUF = GR * XS + F32_STATS;
  __shared__ __align__(16) float xs[2][BUF];
  __shared__ double red[WAVES];

  int li, kb;
  fisher_mfma_coords(numKB, numImages, 0, li, kb);
  if (li >= numImages) return;
  const int img = imgBegin + li;
  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int kw = kb * BC + 16 * CT * w;
  const long long r0 = (long long)offsets[img], r1 = (long long)offsets[img + 1];
  const long long nImg = r1 - r0;
  const long long gBase = gstart[imgBegin];
  const long long gB = gstart[img] - gBase, gE = gstart[img + 1] - gBase;
  const bool fast = (flags & 4) != 0;
  const MfmaScales sc = *scp;
  const bool vecx = (D & 3) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;

  float wr[CT][8 * NS];
  float bc[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    const int comp = kw + 16 * c + lr;
    bc[c] = bias[comp];
#pragma unroll
    for (int j4 = 0; j4 < 2 * NS; ++j4) {
      const f32x4 v = *reinterpret_cast<const f32x4*>(W + (size_t)comp * F + 8 * NS * lg + 4 * j4);
#pragma unroll
      for (int e = 0; e < 4; ++e) wr[c][4 * j4 + e] = v[e];
    }
  }
  // staging words of this thread: row words (minus g on real rows, g read
  // from LDS) and the group's statistics
  __shared__ __align__(16) float gsh[Dp];
  for (int d = tid; d < Dp; d += NT) gsh[d] = g[d];
  auto load_group = [&](long long gi, f32x4 (&pre)[PER]) {
    const long long rg = r0 + (long long)GR * (gi - gB);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = tid + NT * q;
      f32x4 v = f32x4{0.f, 0.f, 0.f, 0.f};
      if (i < XWORDS) {
        const int r = i / (Dp / 4), c4 = (i - r * (Dp / 4)) * 4;
        const long long row = rg + r;
        if (row < r1) {
          if (vecx) {
            if (c4 < D) v = *reinterpret_cast<const f32x4*>(x + row * D + c4);
          } else {
#pragma unroll
            for (int e = 0; e < 4; ++e) v[e] = c4 + e < D ? x[row * D + c4 + e] : 0.f;
          }
        }
      } else if (i < NV) {
        v = *reinterpret_cast<const f32x4*>(stats + (size_t)gi * F32_STATS + 4 * (i - XWORDS));
      }
      pre[q] = v;
    }
  };
  auto store_group = [&](const f32x4 (&pre)[PER], float* dst, long long gi) {
    const long long rg = r0 + (long long)GR * (gi - gB);
#pragma unroll
    for (int q = 0; q < PER; ++q) {
      const int i = tid + NT * q;
      if (i < XWORDS) {
        const int r = i / (Dp / 4), c4 = (i - r * (Dp / 4)) * 4;
        *reinterpret_cast<f32x4*>(dst + r * XS + c4) =
            rg + r < r1 ? pre[q] - *reinterpret_cast<const f32x4*>(gsh + c4) : pre[q];
      } else if (i < NV) {
        *reinterpret_cast<f32x4*>(dst + GR * XS + 4 * (i - XWORDS)) = pre[q];
      }
    }
  };

  f32x4 macc[CT][2][NS];
  f32x4 macc0[CT];
#pragma unroll
  for (int c = 0; c < CT; ++c) {
    macc0[c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int s = 0; s < 2; ++s)
#pragma unroll
      for (int ft = 0; ft < NS; ++ft) macc[c][s][ft] = f32x4{0.f, 0.f, 0.f, 0.f};
  }
  unsigned terms = 0;
  const float qmin = __uint_as_float(0x358637beu) * sc.sq;  // (double)q < 1e-6 for a float q
  const bool sqLane = lg < 2;

  f32x4 pre[PER];
  __syncthreads();  // gsh
  if (gB < gE) {
    load_group(gB, pre);
    store_group(pre, xs[0], gB);
    if (gB + 1 < gE) load_group(gB + 1, pre);
  }
  __syncthreads();
  for (long long gi = gB; gi < gE; ++gi) {
    const int buf = (int)((gi - gB) & 1);
    if (gi + 1 < gE) {
      store_group(pre, xs[buf ^ 1], gi + 1);  // that buffer (group gi - 1) was read before the last barrier
      if (gi + 2 < gE) load_group(gi + 2, pre);
    }
    const float* cur = xs[buf];
    const float* Sf = cur + GR * XS;
    f32x4 nm4[2], rs4[2];
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      nm4[t] = *reinterpret_cast<const f32x4*>(&Sf[16 * t + 4 * lg]);
      rs4[t] = *reinterpret_cast<const f32x4*>(&Sf[GR + 16 * t + 4 * lg]);
    }
    f32x4 acc[2][CT];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < CT; ++c)
        acc[t][c] = f32x4{bc[c] + nm4[t][0], bc[c] + nm4[t][1], bc[c] + nm4[t][2], bc[c] + nm4[t][3]};
#pragma unroll
    for (int j4 = 0; j4 < 2 * NS; ++j4) {
      f32x4 av[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const f32x4 v = *reinterpret_cast<const f32x4*>(cur + (16 * t + lr) * XS + 8 * NS * (lg & 1) + 4 * j4);
        av[t] = sqLane ? v * v : v;
      }
#pragma unroll
      for (int e = 0; e < 4; ++e)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int c = 0; c < CT; ++c) acc[t][c] = mfma4(av[t][e], wr[c][4 * j4 + e], acc[t][c]);
    }
    float qv[CT][2][4];
    if (fast) {
      f32x4 ar4[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) ar4[t] = *reinterpret_cast<const f32x4*>(&Sf[2 * GR + 16 * t + 4 * lg]);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            const bool hit = kw + 16 * c + lr == __float_as_int(ar4[t][r]) && bc[c] != -INFINITY;
            terms += (unsigned)__popcll(__ballot(hit));
            qv[c][t][r] = hit ? sc.sq : 0.f;
          }
    } else {
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r)
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            const float v = __builtin_amdgcn_exp2f(acc[t][c][r]) * rs4[t][r];
            const bool keep = v >= qmin;  // false for NaN (padding rows: inf * 0)
            terms += (unsigned)__popcll(__ballot(keep));
            qv[c][t][r] = keep ? v : 0.f;
          }
    }
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
#pragma unroll
        for (int c = 0; c < CT; ++c) macc0[c] = mfma4(qv[c][t][r], 1.0f, macc0[c]);
#pragma unroll
        for (int ft = 0; ft < NS; ++ft) {
          const float b1 = cur[(16 * t + 4 * lg + r) * XS + 16 * ft + lr];
          const float b2 = b1 * b1;
#pragma unroll
          for (int c = 0; c < CT; ++c) {
            macc[c][0][ft] = mfma4(qv[c][t][r], b1, macc[c][0][ft]);
            macc[c][1][ft] = mfma4(qv[c][t][r], b2, macc[c][1][ft]);
          }
        }
      }
    __syncthreads();
  }
  fisher_slice_finalize<float, NS, CT, WAVES>(macc0, macc, sc, kw, lr, lg, tid, lane, img, kb, numKB, nImg, D, K, mup,
                                              isg, sinv, priors, flags, enc, normpart, numTerms, terms, red);
}

// Bytes of row statistics materialised at once by the FP32 matrix-core encoder.
constexpr size_t kF32StatsBudget = size_t(1) << 30;

template <int NS, int CT, int WAVES>
static void run_f32mfma(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                        const float* x, const uint64_t* offsets, size_t numImages, int flags,
                        unsigned long long* numTerms, const std::vector<uint64_t>& off) {
  constexpr int Dp = NS * 16;
  constexpr int BC = WAVES * 16 * CT;
  static_assert(BC % MB_COMPS == 0, "component blocks of both passes must agree");
  constexpr int LG = F32LseWaves<NS>::value;
  const int Kp = (int)((K + BC - 1) / BC) * BC;
  unsigned char* base = (unsigned char*)scratch(SLOT_PREP, mfma_tables_bytes<float>(Dp, Kp));
  const MfmaTables<float> t = mfma_tables_carve<float>(base, Dp, Kp);
  mfma_tables_launch<float>(t, means, covs, priors, (int)D, Dp, (int)K, Kp);
  const int numKB = Kp / BC;
  double* normpart = (double*)scratch(SLOT_PARTIAL, numImages * (size_t)numKB * sizeof(double));
  if (numTerms) zero_words(numTerms, 2 * numImages);
  std::vector<long long> gs(numImages + 1);
  gs[0] = 0;
  for (size_t i = 0; i < numImages; ++i) gs[i + 1] = gs[i] + (long long)((off[i + 1] - off[i] + GR - 1) / GR);
  long long* gstart = (long long*)scratch(SLOT_GROUPS, (numImages + 1) * sizeof(long long));
  upload(gstart, gs.data(), (numImages + 1) * sizeof(long long));
  sync();  // gs is a local buffer
  constexpr size_t kStatBytes = F32_STATS * sizeof(float);
  size_t b0 = 0;
  while (b0 < numImages) {
    size_t b1 = b0 + 1;
    while (b1 < numImages && (size_t)(gs[b1 + 1] - gs[b0]) * kStatBytes <= kF32StatsBudget &&
           off[b1 + 1] - off[b0] < (uint64_t(1) << 31))
      ++b1;
    const uint64_t rowBase = off[b0], rows = off[b1] - off[b0];
    VLHIP_REQUIRE(rows < (uint64_t(1) << 31), "an image of the FP32 matrix-core encoder exceeds 2^31 descriptors");
    const size_t nG = (size_t)(gs[b1] - gs[b0]);
    const int n = (int)(b1 - b0);
    float* stats = (float*)scratch(SLOT_FISHER_WORK, (nG ? nG : 1) * kStatBytes);
    if (nG > 0) {
      int4* gtab = (int4*)scratch(SLOT_GROUPTAB, nG * sizeof(int4));
      VLHIP_LAUNCH("mfma_group_table", mfma_group_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0,
                   offsets, gstart, (int)b0, n, gtab);
      const unsigned lb = (unsigned)((nG + LG - 1) / LG);
      if (flags & 4)
        VLHIP_LAUNCH("f32_lse", (f32_lse_kernel<NS, true>), dim3(lb), dim3(LG * 64), 0, x + rowBase * D, (int)D, Kp,
                     (const float*)t.Whi, t.bias, t.g, stats, (long long)nG, gtab);
      else
        VLHIP_LAUNCH("f32_lse", (f32_lse_kernel<NS, false>), dim3(lb), dim3(LG * 64), 0, x + rowBase * D, (int)D, Kp,
                     (const float*)t.Whi, t.bias, t.g, stats, (long long)nG, gtab);
    }
    const int padImages = n >= 8 ? ((n + 7) / 8) * 8 : n;
    VLHIP_LAUNCH("f32_fisher", (f32_fisher_kernel<NS, CT, WAVES>), dim3((unsigned)((size_t)padImages * numKB)),
                 dim3(WAVES * 64), 0, x, (int)D, t.g, stats, offsets, gstart, (int)b0, n, (int)K,
                 (const float*)t.Whi, t.bias, t.scp, t.mup, t.isg, t.sinv, priors, flags, enc, normpart, numTerms,
                 numKB);
    b0 = b1;
  }
  if (flags & 2) launch_fisher_normalize(enc, normpart, numKB, numImages, (size_t)2 * K * D);
}

template <int NS>
static void run_f32mfma_k(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                          const float* x, const uint64_t* offsets, size_t numImages, int flags,
                          unsigned long long* numTerms, const std::vector<uint64_t>& off) {
  // wider component blocks when there are components enough (the weights of
  // 2 tiles fit a wave's registers up to D = 80)
  if constexpr (NS <= 5) {
    if (K > 128) {
      run_f32mfma<NS, 2, 8>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off);
      return;
    }
    run_f32mfma<NS, 2, 4>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off);
  } else {
    run_f32mfma<NS, 1, 8>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off);
  }
}

void fisher_f32mfma(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                    const float* x, const uint64_t* offsets, size_t numImages, int flags,
                    unsigned long long* numTerms) {
  VLHIP_REQUIRE(K >= 1 && D >= 1, "numClusters and dimension must be positive");
  if (numImages == 0) return;
  VLHIP_REQUIRE(K <= (size_t)1 << 20, "the FP32 matrix-core encoder supports at most 2^20 components");
  std::vector<uint64_t> off(numImages + 1);
  download(off.data(), offsets, (numImages + 1) * sizeof(uint64_t));
  switch ((D + 15) / 16) {
    case 1: run_f32mfma_k<1>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off); break;
    case 2: run_f32mfma_k<2>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off); break;
    case 3: run_f32mfma_k<3>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off); break;
    case 4: run_f32mfma_k<4>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off); break;
    case 5: run_f32mfma_k<5>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off); break;
    case 6: run_f32mfma_k<6>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off); break;
    case 7: run_f32mfma_k<7>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off); break;
    case 8: run_f32mfma_k<8>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, off); break;
    default: fatal("vlhip_fisher_encode_batch", __FILE__, __LINE__,
                   "the FP32 matrix-core Fisher encoder supports dimension <= 128");
  }
}


// ---------------------------------------------------------------------------
// EM on the matrix cores (vl_gmm_em with a compensated precision).  One EM
// step over N descriptors: the data set is cut into chunks of C rows that play
// the part of images (groups of 32 rows, per-chunk partials), the first pass
// writes the group records and row statistics, the log-likelihood
//   LL = sum_i log sum_k pi_k N(x_i; mu_k, Sigma_k)          (gmm.c:786)
// is summed from the row statistics in FP64, and the second pass (EM mode of
// mfma_fisher_kernel) writes per-chunk FP64 moments about g that one kernel
// reduces over the chunks in a fixed order and converts to the moments the
// M-step finalisation takes (gmm.c:1093-1149):
//   S1 = sum q x = S1' + g S0,   S2 = sum q (x - mu_old)^2 = S2' - 2 a S1' + a^2 S0,
// a = mu_old - g, S1' = sum q x', S2' = sum q x'^2.

// ln of the row sums: log2 sum_k 2^l_k = -stat0 + log2(sq) - log2(stat1),
// summed over the real rows (stat1 = 0 marks padding rows).
template <int NS>
__global__ __launch_bounds__(256) void mfma_em_ll_kernel(const unsigned char* __restrict__ rec, long long numGroups,
                                                         const MfmaScales* __restrict__ scp, double* __restrict__ part) {
  typedef GroupRec<NS> R;
  __shared__ double red[256];
  const double lsq = log2((double)scp->sq);
  double s = 0.0;
  const long long rows = numGroups * GR;
  for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < rows; i += (long long)gridDim.x * 256) {
    const long long grp = i / GR;
    const int r = (int)(i - grp * GR);
    const float* Sf = reinterpret_cast<const float*>(rec + (size_t)grp * R::kBytes + R::kA + R::kB);
    const float rs = Sf[GR + r];
    if (rs > 0.f) s += lsq - (double)Sf[r] - log2((double)rs);
  }
  const double b = block_sum_256(s, red);
  if (threadIdx.x == 0) part[blockIdx.x] = b;
}

// Fixed-order FP64 sum of the per-chunk moments and the change of origin;
// one thread per (component, dimension).  out = [S0 K | mass K | S1 KD | S2 KD].
__global__ __launch_bounds__(256) void mfma_em_reduce_kernel(const double* __restrict__ part, int chunks, int D, int K,
                                                             const float* __restrict__ oldMeans,
                                                             const float* __restrict__ g, double* __restrict__ out) {
  const size_t KD = (size_t)K * D;
  const size_t i = (size_t)blockIdx.x * 256 + threadIdx.x;
  if (i >= KD) return;
  const size_t k = i / D, d = i - k * D;
  const size_t PS = 2 * (size_t)K + 2 * KD;
  double s0 = 0.0, s1 = 0.0, s2 = 0.0, ms = 0.0;
  for (int c = 0; c < chunks; ++c) {
    const double* P = part + (size_t)c * PS;
    s0 += P[k];
    s1 += P[2 * K + i];
    s2 += P[2 * K + KD + i];
    if (d == 0) ms += P[K + k];
  }
  const double gd = (double)g[d];
  const double a = (double)oldMeans[i] - gd;
  out[2 * K + i] = s1 + gd * s0;
  out[2 * K + KD + i] = s2 - 2.0 * a * s1 + a * a * s0;
  if (d == 0) {
    out[k] = s0;
    out[K + k] = ms;
  }
}

// Rows per chunk: 4096 from N = 2^18 rows on (short enough that the FP32
// matrix accumulators stay accurate), fewer below so that there are at least
// 64 chunks, more only when N needs more than 2048 chunks.
static size_t em_chunk_rows(size_t N) {
  size_t c = (N + 63) / 64;
  c = std::max<size_t>(256, std::min<size_t>(4096, c));
  c = std::max(c, (N + 2047) / 2048);
  return (c + GR - 1) / GR * GR;
}

template <typename H, int NS, int CT, int WAVES>
static double run_em_step(double* out_host, const float* means, const float* covs, const float* priors, size_t D,
                          size_t K, const float* x, size_t N) {
  constexpr int Dp = NS * 16;
  constexpr int BC = WAVES * 16 * CT;
  typedef GroupRec<NS> R;
  const int Kp = (int)((K + BC - 1) / BC) * BC;
  unsigned char* base = (unsigned char*)scratch(SLOT_PREP, mfma_tables_bytes<H>(Dp, Kp));
  const MfmaTables<H> t = mfma_tables_carve<H>(base, Dp, Kp);
  mfma_tables_launch<H>(t, means, covs, priors, (int)D, Dp, (int)K, Kp);
  const int numKB = Kp / BC;

  const size_t C = em_chunk_rows(N);
  const size_t numChunks = (N + C - 1) / C;
  std::vector<uint64_t> off(numChunks + 1);
  std::vector<long long> gs(numChunks + 1);
  gs[0] = 0;
  for (size_t i = 0; i <= numChunks; ++i) off[i] = std::min<uint64_t>(N, (uint64_t)(i * C));
  for (size_t i = 0; i < numChunks; ++i) gs[i + 1] = gs[i] + (long long)((off[i + 1] - off[i] + GR - 1) / GR);
  uint64_t* offsets = (uint64_t*)scratch(SLOT_EMCHUNKS, (numChunks + 1) * sizeof(uint64_t));
  long long* gstart = (long long*)scratch(SLOT_GROUPS, (numChunks + 1) * sizeof(long long));
  upload(offsets, off.data(), (numChunks + 1) * sizeof(uint64_t));
  upload(gstart, gs.data(), (numChunks + 1) * sizeof(long long));

  const size_t KD = K * D;
  const size_t PS = 2 * K + 2 * KD;
  double* part = out_host ? (double*)scratch(SLOT_PARTIAL, numChunks * PS * sizeof(double)) : nullptr;
  float qthr = (float)(0.01 / (double)K);  // smallest float q with !((double)q < 0.01 / K)
  if ((double)qthr < 0.01 / (double)K) qthr = std::nextafter(qthr, INFINITY);
  // 128-component blocks: the paired kernel on eight waves
  constexpr bool kPair = CT == 2 && WAVES == 4;
  const size_t shmem = kPair ? pair_lds_bytes<NS>() : mfma_shmem<H, NS, CT, WAVES>(false);
  if constexpr (kPair)
    VLHIP_CHECK(hipFuncSetAttribute((const void*)mfma_fisher_pair_kernel<float, H, NS, false, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  else
    VLHIP_CHECK(hipFuncSetAttribute((const void*)mfma_fisher_kernel<float, H, NS, CT, WAVES, true>,
                                    hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  // the scaled threshold needs the device scale sq: read it once
  MfmaScales hs;
  download(&hs, t.scp, sizeof(MfmaScales));
  const float emThr = qthr * hs.sq;

  constexpr unsigned kLLBlocks = 512;
  std::vector<double> llh;
  size_t b0 = 0;
  while (b0 < numChunks) {
    size_t b1 = b0 + 1;
    while (b1 < numChunks && (size_t)(gs[b1 + 1] - gs[b0]) * R::kBytes <= kRecordBudget) ++b1;
    const uint64_t rowBase = off[b0];
    const size_t nG = (size_t)(gs[b1] - gs[b0]);
    const int n = (int)(b1 - b0);
    unsigned char* rec = (unsigned char*)scratch(SLOT_FISHER_WORK, nG * R::kBytes);
    int4* gtab = (int4*)scratch(SLOT_GROUPTAB, nG * sizeof(int4));
    double* llp = (double*)scratch(SLOT_EMLL, kLLBlocks * sizeof(double));
    VLHIP_LAUNCH("mfma_group_table", mfma_group_table_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, offsets,
                 gstart, (int)b0, n, gtab);
    constexpr int LG = LseGroups<NS>::value;
    VLHIP_LAUNCH("mfma_lse", (mfma_lse_kernel<H, NS, false>), dim3((unsigned)((nG + LG - 1) / LG)), dim3(LG * 64), 0,
                 x + rowBase * D, (int)D, Kp, t.Wst, t.bias, t.scp, t.g, rec, (long long)nG, gtab,
                 (const unsigned char*)nullptr);
    VLHIP_LAUNCH("mfma_em_ll", mfma_em_ll_kernel<NS>, dim3(kLLBlocks), dim3(256), 0, rec, (long long)nG, t.scp, llp);
    if (out_host) {
      const int padImages = n >= 8 ? ((n + 7) / 8) * 8 : n;
      const dim3 grid((unsigned)((size_t)padImages * numKB));
      if constexpr (kPair)
        VLHIP_LAUNCH("mfma_em_moments", (mfma_fisher_pair_kernel<float, H, NS, false, true>), grid, dim3(512),
                     shmem, rec, offsets, gstart, (int)b0, n, (int)D, (int)K, t.Whi, t.Wlo, t.bias, t.scp, t.mup,
                     t.isg, t.sinv, priors, 0, (float*)nullptr, (double*)nullptr, (unsigned long long*)nullptr,
                     numKB, emThr, part, 0);
      else
        VLHIP_LAUNCH("mfma_em_moments", (mfma_fisher_kernel<float, H, NS, CT, WAVES, true>), grid,
                     dim3(WAVES * 64), shmem, rec, offsets, gstart, (int)b0, n, (int)D, (int)K, t.Whi, t.Wlo, t.bias,
                     t.scp, t.mup, t.isg, t.sinv, priors, 0, (float*)nullptr, (double*)nullptr,
                     (unsigned long long*)nullptr, numKB, (const unsigned char*)nullptr, (const int2*)nullptr, 0,
                     emThr, part, 0);
    }
    const size_t o = llh.size();
    llh.resize(o + kLLBlocks);
    download(llh.data() + o, llp, kLLBlocks * sizeof(double));  // synchronous: llp is reused by the next batch
    b0 = b1;
  }
  double ll2 = 0.0;
  for (double v : llh) ll2 += v;
  if (out_host) {
    double* out = (double*)scratch(SLOT_REDUCED, PS * sizeof(double));
    VLHIP_LAUNCH("mfma_em_reduce", mfma_em_reduce_kernel, dim3((unsigned)((KD + 255) / 256)), dim3(256), 0, part,
                 (int)numChunks, (int)D, (int)K, means, t.g, out);
    download(out_host, out, PS * sizeof(double));
  }
  return ll2 * 0.6931471805599453;  // ln 2
}

template <typename H>
static double em_dispatch(double* out, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                          const float* x, size_t N) {
  switch ((D + 15) / 16) {
#define EM_CASE(NS)                                                                      \
  case NS:                                                                               \
    if constexpr (NS <= 4) {                                                             \
      if (K > 128) return run_em_step<H, NS, 2, 8>(out, means, covs, priors, D, K, x, N); \
    }                                                                                    \
    return run_em_step<H, NS, 2, 4>(out, means, covs, priors, D, K, x, N);
    EM_CASE(1) EM_CASE(2) EM_CASE(3) EM_CASE(4) EM_CASE(5) EM_CASE(6) EM_CASE(7) EM_CASE(8)
#undef EM_CASE
    default: fatal("vlhip_gmm_em_step", __FILE__, __LINE__, "the matrix-core EM step supports dimension <= 128");
  }
}

}  // namespace vlhip

using namespace vlhip;

extern "C" {

double vlhip_gmm_em_step(int mode, double* out_host, void const* priors, void const* means, void const* covs,
                         size_t D, size_t K, void const* data, size_t N) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  VLHIP_REQUIRE(mode == VLHIP_MODE_BF16X3 || mode == VLHIP_MODE_F16X3,
                "the matrix-core EM step takes VLHIP_MODE_BF16X3 or VLHIP_MODE_F16X3");
  VLHIP_REQUIRE(N >= 1 && K >= 1 && D >= 1 && D <= 128,
                "the matrix-core EM step needs numData >= 1, numClusters >= 1 and 1 <= dimension <= 128");
  if (mode == VLHIP_MODE_F16X3)
    return em_dispatch<_Float16>(out_host, (const float*)means, (const float*)covs, (const float*)priors, D, K,
                                 (const float*)data, N);
  return em_dispatch<__bf16>(out_host, (const float*)means, (const float*)covs, (const float*)priors, D, K,
                             (const float*)data, N);
}

void vlhip_fisher_set_fused(int mode) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  VLHIP_REQUIRE(mode >= 0 && mode <= 2, "the fused-encoder mode is 0 (never), 1 (automatic) or 2 (whenever possible)");
  g_fused_mode = mode;
}

void* vlhip_encode_plan_new(size_t D, size_t K, uint64_t const* offsets_host, size_t numImages, int fisherFlags,
                            int vladFlags) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  VLHIP_REQUIRE(D >= 1 && D <= 128 && K >= 1, "encode plans need 1 <= dimension <= 128 and numClusters >= 1");
  VLHIP_REQUIRE(numImages >= 1, "encode plans need at least one image");
  EncodePlan* P = new EncodePlan();
  P->D = D;
  P->K = K;
  P->numImages = numImages;
  P->flags = fisherFlags;
  P->vladFlags = vladFlags;
  P->NS = (int)((D + 15) / 16);
  P->totalRows = offsets_host[numImages] - offsets_host[0];
  VLHIP_REQUIRE(offsets_host[0] == 0, "offsets must start at 0");
  const size_t recBytes = plan_record_bytes(P->NS);
  // image-aligned groups and batches
  std::vector<long long> gs(numImages + 1);
  gs[0] = 0;
  for (size_t i = 0; i < numImages; ++i) {
    VLHIP_REQUIRE(offsets_host[i + 1] >= offsets_host[i], "offsets must be non-decreasing");
    gs[i + 1] = gs[i] + (long long)((offsets_host[i + 1] - offsets_host[i] + GR - 1) / GR);
  }
  std::vector<int4> tab((size_t)gs[numImages] ? (size_t)gs[numImages] : 1);
  size_t b0 = 0;
  while (b0 < numImages) {
    size_t b1 = b0 + 1;
    while (b1 < numImages && (size_t)(gs[b1 + 1] - gs[b0]) * recBytes <= kRecordBudget) ++b1;
    PlanBatch b;
    b.b0 = b0;
    b.n = b1 - b0;
    b.rowBase = offsets_host[b0];
    b.nG = (size_t)(gs[b1] - gs[b0]);
    b.gOff = gs[b0];
    VLHIP_REQUIRE(offsets_host[b1] - offsets_host[b0] < (uint64_t(1) << 31), "image batch exceeds 2^31 descriptors");
    for (size_t i = b0; i < b1; ++i)
      for (long long k = 0; k < gs[i + 1] - gs[i]; ++k)
        tab[(size_t)(gs[i] + k)] = make_int4((int)(offsets_host[i] - b.rowBase + (uint64_t)(GR * k)),
                                             (int)(offsets_host[i + 1] - b.rowBase), (int)i, 0);
    P->maxGroups = std::max(P->maxGroups, b.nG);
    P->batches.push_back(b);
    b0 = b1;
  }
  VLHIP_CHECK(hipMalloc(&P->dOff, (numImages + 1) * sizeof(uint64_t)));
  VLHIP_CHECK(hipMalloc(&P->dGstart, (numImages + 1) * sizeof(long long)));
  VLHIP_CHECK(hipMalloc(&P->dGtab, tab.size() * sizeof(int4)));
  VLHIP_CHECK(hipMalloc(&P->dRec, (P->maxGroups ? P->maxGroups : 1) * recBytes));
  const int Kp = (int)((K + 127) / 128) * 128 > (int)((K + 255) / 256) * 256 ? (int)((K + 127) / 128) * 128
                                                                              : (int)((K + 255) / 256) * 256;
  VLHIP_CHECK(hipMalloc(&P->dTables, mfma_tables_bytes<_Float16>(P->NS * 16, Kp)));
  VLHIP_CHECK(hipMalloc(&P->dNorm, numImages * (size_t)(Kp / kPlanMinBlock) * sizeof(double)));
  if (vladFlags >= 0) {
    VLHIP_CHECK(hipMalloc(&P->dAssign, (P->totalRows ? P->totalRows : 1) * sizeof(uint32_t)));
    const size_t qb = quantize_screened_bytes(D, K);
    if (qb) VLHIP_CHECK(hipMalloc(&P->dQtab, qb));
  }
  VLHIP_CHECK(hipMemcpy(P->dOff, offsets_host, (numImages + 1) * sizeof(uint64_t), hipMemcpyHostToDevice));
  VLHIP_CHECK(hipMemcpy(P->dGstart, gs.data(), (numImages + 1) * sizeof(long long), hipMemcpyHostToDevice));
  VLHIP_CHECK(hipMemcpy(P->dGtab, tab.data(), tab.size() * sizeof(int4), hipMemcpyHostToDevice));
  VLHIP_CHECK(hipStreamCreateWithFlags(&P->stream, hipStreamNonBlocking));
  plan_set_attrs(P->NS);
  return P;
}

void vlhip_encode_plan_delete(void* plan) {
  if (!plan) return;
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  EncodePlan* P = (EncodePlan*)plan;
  if (P->stream) VLHIP_CHECK(hipStreamSynchronize(P->stream));
  if (P->exec) VLHIP_CHECK(hipGraphExecDestroy(P->exec));
  if (P->graph) VLHIP_CHECK(hipGraphDestroy(P->graph));
  void* bufs[] = {P->dOff, P->dGstart, P->dGtab, P->dRec, P->dTables, P->dNorm, P->dAssign, P->dQtab,
                  P->dProjP, P->dProjC, P->dProjX};
  for (void* b : bufs)
    if (b) VLHIP_CHECK(hipFree(b));
  if (P->stream) VLHIP_CHECK(hipStreamDestroy(P->stream));
  delete P;
}

// One encode of all the plan's images (FP16 compensated Fisher vectors and,
// when the plan has VLAD flags, hard-assignment VLAD against the means).
// Returns when the results are complete.
void vlhip_encode_plan_run(void* plan, void const* data, void const* means, void const* covs, void const* priors,
                           void* fisherEnc, void* vladEnc, unsigned long long* numTerms) {
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  EncodePlan* P = (EncodePlan*)plan;
  sync();  // inputs produced on the library stream are complete
  const float* x = (const float*)data;
  const void* key[7] = {data, means, covs, priors, fisherEnc, vladEnc, numTerms};
  hipStream_t saved = ctx().stream;
  const bool savedTiming = ctx().timing;
  ctx().stream = P->stream;
  if (P->tiling < 0) {
    // Autotuner, on the plan's own shapes: every allowed component tiling with
    // the default schedule, then the other component-block schedules with the
    // fastest tiling; each candidate is run once to warm up and once timed
    // (hipEvents on the plan's stream).  Choices are cached per shape.
    std::vector<long long> shape = {(long long)P->D, (long long)P->K, (long long)P->flags,
                                    (long long)P->vladFlags, (long long)P->Din, (long long)P->renorm,
                                    (long long)(64 - __builtin_clzll((unsigned long long)P->totalRows | 1)),
                                    (long long)(64 - __builtin_clzll((unsigned long long)P->numImages | 1))};
    for (auto& e : tuning_cache())
      if (e.first == shape) {
        P->tiling = e.second.first;
        P->sched = e.second.second;
      }
    if (P->tiling < 0) {
      hipEvent_t e0, e1;
      VLHIP_CHECK(hipEventCreate(&e0));
      VLHIP_CHECK(hipEventCreate(&e1));
      ctx().timing = false;
      auto timed = [&](int tiling, int sched) {
        plan_enqueue_any(*P, tiling, sched, x, (const float*)means, (const float*)covs, (const float*)priors,
                         (float*)fisherEnc, (float*)vladEnc, numTerms);  // warm-up
        VLHIP_CHECK(hipEventRecord(e0, P->stream));
        plan_enqueue_any(*P, tiling, sched, x, (const float*)means, (const float*)covs, (const float*)priors,
                         (float*)fisherEnc, (float*)vladEnc, numTerms);
        VLHIP_CHECK(hipEventRecord(e1, P->stream));
        VLHIP_CHECK(hipEventSynchronize(e1));
        float ms = 0.f;
        VLHIP_CHECK(hipEventElapsedTime(&ms, e0, e1));
        P->tried.push_back({(double)tiling, (double)sched, (double)ms});
      };
      for (int ti = 0; ti < 4; ++ti)
        if (plan_tiling_allowed(ti, P->NS, P->K)) timed(ti, 0);
      auto best = [&]() {
        size_t b = 0;
        for (size_t i = 1; i < P->tried.size(); ++i)
          if (P->tried[i][2] < P->tried[b][2]) b = i;
        return b;
      };
      const int bestTiling = (int)P->tried[best()][0];
      if (P->numImages >= 8)  // with fewer images every schedule is the linear one
        for (int sc = 1; sc < kPlanSchedules; ++sc) timed(bestTiling, sc);
      VLHIP_CHECK(hipEventDestroy(e0));
      VLHIP_CHECK(hipEventDestroy(e1));
      P->chosen = (int)best();
      P->tiling = (int)P->tried[P->chosen][0];
      P->sched = (int)P->tried[P->chosen][1];
      tuning_cache().push_back({shape, {P->tiling, P->sched}});
    }
  }
  bool same = P->exec != nullptr;
  for (int i = 0; i < 7; ++i) same = same && P->key[i] == key[i];
  if (!same) {
    if (P->exec) VLHIP_CHECK(hipGraphExecDestroy(P->exec));
    if (P->graph) VLHIP_CHECK(hipGraphDestroy(P->graph));
    P->exec = nullptr;
    P->graph = nullptr;
    ctx().timing = false;  // no event records inside the capture
    VLHIP_CHECK(hipStreamBeginCapture(P->stream, hipStreamCaptureModeRelaxed));
    ctx().capturing = true;  // scratch() refuses to run: nothing library-owned enters the graph
    plan_enqueue_any(*P, P->tiling, P->sched, x, (const float*)means, (const float*)covs, (const float*)priors,
                     (float*)fisherEnc, (float*)vladEnc, numTerms);
    ctx().capturing = false;
    VLHIP_CHECK(hipStreamEndCapture(P->stream, &P->graph));
    VLHIP_CHECK(hipGraphInstantiate(&P->exec, P->graph, nullptr, nullptr, 0));
    for (int i = 0; i < 7; ++i) P->key[i] = key[i];
  }
  ctx().timing = savedTiming;
  ctx().stream = saved;
  VLHIP_CHECK(hipGraphLaunch(P->exec, P->stream));
  VLHIP_CHECK(hipStreamSynchronize(P->stream));
}

// Load stage of a plan: from the next run on, the data passed to
// vlhip_encode_plan_run have inDimension columns and are projected (and
// renormalised when asked) to the plan's dimension before encoding, as
// apps/recognition/encodeImage.m:91-97 does before vl_fisher / vl_vlad.
// projection (D x inDimension) and center (inDimension) are device arrays,
// copied into the plan.
void vlhip_encode_plan_set_projection(void* plan, size_t inDimension, void const* projection, void const* center,
                                      int renormalize) {
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  EncodePlan* P = (EncodePlan*)plan;
  VLHIP_REQUIRE(inDimension >= 1 && inDimension < (1u << 20), "projection inDimension out of range");
  if (P->stream) VLHIP_CHECK(hipStreamSynchronize(P->stream));
  if (P->exec) VLHIP_CHECK(hipGraphExecDestroy(P->exec));
  if (P->graph) VLHIP_CHECK(hipGraphDestroy(P->graph));
  P->exec = nullptr;
  P->graph = nullptr;
  for (float** b : {&P->dProjP, &P->dProjC}) {
    if (*b) VLHIP_CHECK(hipFree(*b));
    *b = nullptr;
  }
  VLHIP_CHECK(hipMalloc(&P->dProjP, P->D * inDimension * sizeof(float)));
  VLHIP_CHECK(hipMalloc(&P->dProjC, inDimension * sizeof(float)));
  if (!P->dProjX) VLHIP_CHECK(hipMalloc(&P->dProjX, (P->totalRows ? P->totalRows : 1) * P->D * sizeof(float)));
  sync();  // inputs produced on the library stream are complete
  VLHIP_CHECK(hipMemcpy(P->dProjP, projection, P->D * inDimension * sizeof(float), hipMemcpyDeviceToDevice));
  VLHIP_CHECK(hipMemcpy(P->dProjC, center, inDimension * sizeof(float), hipMemcpyDeviceToDevice));
  P->Din = inDimension;
  P->renorm = renormalize ? 1 : 0;
}

void vlhip_encode_plan_info(void* plan, int* variant, double* variantMs, size_t* numBatches) {
  EncodePlan* P = (EncodePlan*)plan;
  if (variant) *variant = P->tiling;
  if (variantMs) {
    // time of the chosen candidate, and of the default (tiling 0 or 1 with
    // schedule 0) when it was tried
    variantMs[0] = variantMs[1] = -1.0;
    if (P->chosen >= 0) variantMs[0] = P->tried[(size_t)P->chosen][2];
    for (const auto& c : P->tried)
      if (c[1] == 0.0 && c[0] == (plan_tiling_allowed(0, P->NS, P->K) ? 0.0 : 1.0)) variantMs[1] = c[2];
  }
  if (numBatches) *numBatches = P->batches.size();
}

int vlhip_encode_plan_candidates(void* plan, int index, int* componentsPerWave, int* waves, int* schedule,
                                 double* ms) {
  EncodePlan* P = (EncodePlan*)plan;
  const int n = (int)P->tried.size();
  if (index >= 0 && index < n) {
    const auto& c = P->tried[(size_t)index];
    const PlanTiling& t = kPlanTilings[(int)c[0]];
    if (componentsPerWave) *componentsPerWave = 16 * t.ct;
    if (waves) *waves = t.waves;
    if (schedule) *schedule = (int)c[1];
    if (ms) *ms = c[2];
  }
  return n;
}

}  // extern "C"
