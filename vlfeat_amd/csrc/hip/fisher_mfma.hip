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
// Kernels: mfma_prep (tables), mfma_lse (row max / sum / argmax over all
// components, 128 rows per workgroup), mfma_fisher (workgroup per image x 64
// components: logits -> posteriors -> moments -> finalised slice of the FV).
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

#include <algorithm>
#include <cmath>
#include <type_traits>

namespace vlhip {

typedef float f32x4 __attribute__((ext_vector_type(4)));
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

constexpr float kLog2e = 1.4426950408889634f;

// Power-of-two operand scales and the feature centre (device resident).
struct MfmaScales {
  float s1, s2;     // x' and x'^2 multipliers
  float sq;         // posterior multiplier
  float pad;
};

// ---------------------------------------------------------------------------
// Tables.  Kp = K rounded up to 64, Dp = D rounded up to 16.
//   Whi/Wlo [Kp][2Dp]   features [0,Dp) = w2 * 2^-e2, [Dp,2Dp) = w1 * 2^-e1
//   bias    [Kp]        b_k (FP32), -inf for skipped / padding components
//   mup, isg, sinv [Kp][Dp]  mu', sigma^-1, sigma^-2 (FP32) for finalisation

// g = sum_k pi_k mu_k / sum_k pi_k, and the range bound used to pick scales.
template <typename T, typename H>
__global__ void mfma_center_kernel(const T* __restrict__ means, const T* __restrict__ covs,
                                   const T* __restrict__ priors, int D, int K, float* __restrict__ g,
                                   MfmaScales* __restrict__ sc, int isHalf) {
  __shared__ double red[256];
  __shared__ float gsh[512];
  const int t = threadIdx.x;
  for (int d = 0; d < D; ++d) {
    double a = 0.0, w = 0.0;
    for (int k = t; k < K; k += 256) {
      a += (double)priors[k] * (double)means[(size_t)k * D + d];
      w += (double)priors[k];
    }
    const double sa = block_sum_256(a, red);
    const double sw = block_sum_256(w, red);
    if (t == 0) {
      const float gd = (float)(sw > 0 ? sa / sw : 0.0);
      g[d] = gd;
      if (d < 512) gsh[d] = gd;
    }
  }
  __syncthreads();
  // bound on |x'|: max over components of |mu'| + 6 sigma
  double mx = 0.0;
  for (int i = t; i < K * D; i += 256) {
    const int d = i % D;
    const double gd = (d < 512) ? (double)gsh[d] : (double)g[d];
    const double v = fabs((double)means[i] - gd) + 6.0 * sqrt((double)covs[i]);
    mx = v > mx ? v : mx;
  }
  red[t] = mx;
  __syncthreads();
  for (int s = 128; s > 0; s >>= 1) {
    if (t < s) red[t] = red[t] > red[t + s] ? red[t] : red[t + s];
    __syncthreads();
  }
  if (t == 0) {
    const double xm = red[0] > 1e-30 ? red[0] : 1.0;
    MfmaScales s;
    if (isHalf) {
      // put |x'| and |x'|^2 near 2^8 and the posterior (<= 1) at 2^14
      s.s1 = (float)exp2(floor(log2(256.0 / xm)));
      s.s2 = (float)exp2(floor(log2(256.0 / (xm * xm))));
      s.sq = 16384.0f;
    } else {
      s.s1 = 1.0f;
      s.s2 = 1.0f;
      s.sq = 1.0f;
    }
    s.pad = 0.f;
    *sc = s;
  }
}

template <typename T, typename H>
__global__ void mfma_tables_kernel(const T* __restrict__ means, const T* __restrict__ covs,
                                   const T* __restrict__ priors, int D, int Dp, int K, int Kp,
                                   const float* __restrict__ g, const MfmaScales* __restrict__ sc,
                                   H* __restrict__ Whi, H* __restrict__ Wlo, float* __restrict__ bias,
                                   float* __restrict__ mup, float* __restrict__ isg, float* __restrict__ sinv) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= Kp) return;
  const float e1 = 1.0f / sc->s1, e2 = 1.0f / sc->s2;
  H* wh = Whi + (size_t)k * 2 * Dp;
  H* wl = Wlo + (size_t)k * 2 * Dp;
  if (k >= K) {
    for (int f = 0; f < 2 * Dp; ++f) { wh[f] = H(0.f); wl[f] = H(0.f); }
    for (int d = 0; d < Dp; ++d) {
      mup[(size_t)k * Dp + d] = 0.f; isg[(size_t)k * Dp + d] = 0.f; sinv[(size_t)k * Dp + d] = 0.f;
    }
    bias[k] = -INFINITY;
    return;
  }
  double logdet = 0.0, quad = 0.0;
  for (int d = 0; d < Dp; ++d) {
    float w2 = 0.f, w1 = 0.f, mu = 0.f, is = 0.f, si = 0.f;
    if (d < D) {
      const double c = (double)covs[(size_t)k * D + d];
      const double s = 1.0 / c;
      const double m = (double)means[(size_t)k * D + d] - (double)g[d];
      logdet += log(c);
      quad += s * m * m;
      w2 = (float)(-0.5 * s) * e2;
      w1 = (float)(s * m) * e1;
      mu = (float)m;
      is = (float)sqrt(s);
      si = (float)s;
    }
    const H h2 = H(w2), h1 = H(w1);
    wh[d] = h2;
    wl[d] = H(w2 - (float)h2);
    wh[Dp + d] = h1;
    wl[Dp + d] = H(w1 - (float)h1);
    mup[(size_t)k * Dp + d] = mu;
    isg[(size_t)k * Dp + d] = is;
    sinv[(size_t)k * Dp + d] = si;
  }
  const double p = (double)priors[k];
  const double logw = (p < 1e-6) ? -(double)INFINITY : log(p);
  bias[k] = (float)(logw - 0.5 * (double)D * kLog2Pi - 0.5 * logdet - 0.5 * quad);
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
// Row statistics over all components (first pass).  Workgroup: 4 waves x 32
// rows; per 64-component block the split weights are staged in LDS and every
// wave runs 2 row tiles x 4 component tiles x NS steps x 3 products.
constexpr int ML_ROWS = 128;
constexpr int MB_COMPS = 64;

template <typename H, int NS>
__global__ __launch_bounds__(256, 2) void mfma_lse_kernel(
    const float* __restrict__ x, long long N, int D, int Kp,
    const H* __restrict__ Whi, const H* __restrict__ Wlo, const float* __restrict__ bias,
    const MfmaScales* __restrict__ scp, const float* __restrict__ g,
    float* __restrict__ row_m, float* __restrict__ row_s, int* __restrict__ row_arg) {
  typedef typename VecOf<H>::type V8;
  constexpr int Dp = NS * 16;
  constexpr int WS = 2 * Dp + 8;  // LDS row stride (halves) of the staged weights
  __shared__ __align__(16) H wsh[2][MB_COMPS][WS];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const long long row0 = (long long)blockIdx.x * ML_ROWS + 32 * w;
  const float s1 = scp->s1, s2 = scp->s2;

  // A fragments of this wave's 32 rows, resident for the whole sweep
  V8 ahi[2][NS], alo[2][NS];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const long long row = row0 + 16 * t + lr;
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int f0 = 32 * j + 8 * lg;
      const bool sq = f0 < Dp;
      const int d0 = sq ? f0 : f0 - Dp;
      float xc[8];
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int d = d0 + v;
        xc[v] = (row < N && d < D) ? x[row * D + d] - g[d] : 0.f;
      }
      make_frag<H>(xc, sq, s1, s2, ahi[t][j], alo[t][j]);
    }
  }

  float mrun[2][4], srun[2][4], bval[2][4];
  int bidx[2][4];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) { mrun[t][r] = -INFINITY; srun[t][r] = 0.f; bval[t][r] = -INFINITY; bidx[t][r] = 0; }

  for (int cb = 0; cb < Kp; cb += MB_COMPS) {
    __syncthreads();
    // stage the block's weights: 64 comps x 2Dp halves, hi and lo
    for (int i = tid; i < 2 * MB_COMPS * (2 * Dp / 8); i += 256) {
      const int part = i / (MB_COMPS * (2 * Dp / 8));
      const int rem = i % (MB_COMPS * (2 * Dp / 8));
      const int c = rem / (2 * Dp / 8), q8 = rem % (2 * Dp / 8);
      const H* src = (part ? Wlo : Whi) + (size_t)(cb + c) * 2 * Dp + 8 * q8;
      *reinterpret_cast<uint4*>(&wsh[part][c][8 * q8]) = *reinterpret_cast<const uint4*>(src);
    }
    __syncthreads();
    f32x4 acc4[4][2];
#pragma unroll
    for (int c = 0; c < 4; ++c) acc4[c][0] = acc4[c][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      V8 bh[4], bl[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        bh[c] = *reinterpret_cast<const V8*>(&wsh[0][16 * c + lr][32 * j + 8 * lg]);
        bl[c] = *reinterpret_cast<const V8*>(&wsh[1][16 * c + lr][32 * j + 8 * lg]);
      }
      // eight independent accumulators between dependent matrix ops
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc4[c][t] = mfma16(ahi[t][j], bh[c], acc4[c][t]);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc4[c][t] = mfma16(alo[t][j], bh[c], acc4[c][t]);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc4[c][t] = mfma16(ahi[t][j], bl[c], acc4[c][t]);
    }
#pragma unroll
    for (int c = 0; c < 4; ++c) {
      const f32x4* acc = acc4[c];
      const int comp = cb + 16 * c + lr;
      const float b = bias[comp];
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const float l = acc[t][r] + b;
          if (l > bval[t][r]) { bval[t][r] = l; bidx[t][r] = comp; }
          const float mn = l > mrun[t][r] ? l : mrun[t][r];
          if (mn != -INFINITY) {
            srun[t][r] = srun[t][r] * __builtin_amdgcn_exp2f((mrun[t][r] - mn) * kLog2e)
                       + __builtin_amdgcn_exp2f((l - mn) * kLog2e);
            mrun[t][r] = mn;
          }
        }
    }
  }

  // merge the 16 lanes that hold the same rows (lanes lg*16 .. lg*16+15)
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      float m = mrun[t][r];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) m = fmaxf(m, __shfl_xor(m, o, 64));
      float s = (mrun[t][r] == -INFINITY) ? 0.f : srun[t][r] * __builtin_amdgcn_exp2f((mrun[t][r] - m) * kLog2e);
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) s += __shfl_xor(s, o, 64);
      float bv = bval[t][r];
      int bi = bidx[t][r];
#pragma unroll
      for (int o = 1; o < 16; o <<= 1) {
        const float ov = __shfl_xor(bv, o, 64);
        const int oi = __shfl_xor(bi, o, 64);
        if (ov > bv || (ov == bv && oi < bi)) { bv = ov; bi = oi; }
      }
      const long long row = row0 + 16 * t + 4 * lg + r;
      if (lr == 0 && row < N) {
        row_m[row] = m;
        row_s[row] = s;
        row_arg[row] = bi;
      }
    }
}

// ---------------------------------------------------------------------------
// Moments and encoding (second pass).  Workgroup = (image, 128 components),
// 8 waves.  Per 64-row chunk:
//   stage   x' (fp32, row-major) and the split [x'_hi, x'_lo, x'^2_hi, x'^2_lo]
//           (16-bit, feature-major) in LDS, once for all 128 components;
//   logits  wave (wr, wc) = rows 32wr.., comps 32wc.. (weights in registers);
//   q       exp(logit - M) / S, FAST one-hot, < 1e-6 dropped (fisher.c:436-441),
//           rounded to 16 bits (scaled), written component-major to LDS;
//   moments wave (plane, half) accumulates one feature plane (x'_hi, x'_lo,
//           x'^2_hi or x'^2_lo) for 64 components over the chunk's rows.
constexpr int MF_ROWS = 64;
constexpr int MF_COMPS = 128;

template <typename H, int NS>
struct MfLds {
  static constexpr int Dp = NS * 16;
  static constexpr int RS = MF_ROWS + 8;  // row stride (halves) of feature-major tiles
  static constexpr int RW = 4 * Dp + 8;   // row stride (halves) of row-major tiles
  // row-major split features [x'^2_hi | x'_hi | x'^2_lo | x'_lo] (logit A operand)
  H xr[MF_ROWS][RW];
  // feature-major split features, planes x'_hi, x'_lo, x'^2_hi, x'^2_lo (moment B operand)
  H xct[4 * Dp][RS];
  H qt[MF_COMPS][RS];                      // component-major rounded posteriors
  float rm[MF_ROWS], rinv[MF_ROWS];
  int ra[MF_ROWS];
  float s0[2][MF_COMPS];
  // the finalisation reuses xr, xct and qt (contiguous) as [2][MF_COMPS][Dp] floats
  static constexpr bool kScratchFits =
      sizeof(H) * MF_ROWS * RW + sizeof(H) * 4 * Dp * RS + sizeof(H) * MF_COMPS * RS >=
      sizeof(float) * 2 * MF_COMPS * Dp;
  static_assert(kScratchFits, "finalisation scratch does not fit");
};

__device__ __forceinline__ void fisher_mfma_coords(int numKB, int numImages, int& img, int& kb) {
  const int b = blockIdx.x;
  if (numImages >= 8) {
    const int xcd = b & 7, slot = b >> 3;
    img = (slot / numKB) * 8 + xcd;
    kb = slot % numKB;
  } else {
    img = b / numKB;
    kb = b % numKB;
  }
}

template <typename H>
__device__ __forceinline__ void store4(H* dst, H a, H b, H c, H d) {
  union { H h[4]; uint2 u; } pk;
  pk.h[0] = a; pk.h[1] = b; pk.h[2] = c; pk.h[3] = d;
  *reinterpret_cast<uint2*>(dst) = pk.u;
}

template <typename T, typename H, int NS>
__global__ __launch_bounds__(256, 1) void mfma_fisher_kernel(
    const float* __restrict__ x, const uint64_t* __restrict__ offsets, int numImages, int D, int K,
    const H* __restrict__ Whi, const H* __restrict__ Wlo, const float* __restrict__ bias,
    const MfmaScales* __restrict__ scp, const float* __restrict__ g,
    const float* __restrict__ mup, const float* __restrict__ isg, const float* __restrict__ sinv,
    const T* __restrict__ priors, const float* __restrict__ row_m, const float* __restrict__ row_s,
    const int* __restrict__ row_arg, int flags, T* __restrict__ enc, double* __restrict__ normpart,
    unsigned long long* __restrict__ numTerms, int numKB) {
  typedef typename VecOf<H>::type V8;
  typedef MfLds<H, NS> L;
  constexpr int Dp = NS * 16;
  extern __shared__ __align__(16) unsigned char smem_raw[];
  L& sm = *reinterpret_cast<L*>(smem_raw);
  __shared__ double red[256];

  int img, kb;
  fisher_mfma_coords(numKB, numImages, img, kb);
  if (img >= numImages) return;

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const int wr = w & 1, wc = w >> 1;  // logits: rows 32wr.., comps 64wc..
  const int pp = w & 1, ch = w >> 1;  // moments: planes (x'_hi,x'_lo) or (x'^2_hi,x'^2_lo), comps 64ch..
  const int k0 = kb * MF_COMPS;
  const long long rBegin = (long long)offsets[img], rEnd = (long long)offsets[img + 1];
  const long long nImg = rEnd - rBegin;
  const bool fast = (flags & 4) != 0;
  const MfmaScales sc = *scp;

  V8 whi[4][NS], wlo[4][NS];
  float bcomp[4];
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int comp = k0 + 64 * wc + 16 * c + lr;
    bcomp[c] = bias[comp];
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      whi[c][j] = *reinterpret_cast<const V8*>(Whi + (size_t)comp * 2 * Dp + 32 * j + 8 * lg);
      wlo[c][j] = *reinterpret_cast<const V8*>(Wlo + (size_t)comp * 2 * Dp + 32 * j + 8 * lg);
    }
  }
  f32x4 macc[4][NS];  // [comp tile][feature tile]; hi and lo planes accumulate together
#pragma unroll
  for (int ct = 0; ct < 4; ++ct)
#pragma unroll
    for (int ft = 0; ft < NS; ++ft) macc[ct][ft] = f32x4{0.f, 0.f, 0.f, 0.f};
  float s0l[4] = {0.f, 0.f, 0.f, 0.f};
  unsigned terms = 0;
  // (double)q < 1e-6 for a float q is q < the next float above float(1e-6)
  const float qmin = __uint_as_float(0x358637beu);  // 1.0000001e-6

  // staging items: (4-row group, dimension); the next chunk's coordinates are
  // prefetched into registers while the current chunk computes.
  constexpr int ITEMS = ((MF_ROWS / 4) * Dp + 255) / 256;
  float xpre[ITEMS][4];
  auto prefetch = [&](long long c0) {
    const int n = (int)min((long long)MF_ROWS, rEnd - c0);
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
      const int it = tid + 256 * q;
      const int rq = it / Dp, d = it - rq * Dp;
#pragma unroll
      for (int i = 0; i < 4; ++i) {
        const int r = 4 * rq + i;
        xpre[q][i] = (it < (MF_ROWS / 4) * Dp && r < n && d < D) ? x[(c0 + r) * D + d] : 0.f;
      }
    }
  };
  if (rBegin < rEnd) prefetch(rBegin);

  for (long long r0 = rBegin; r0 < rEnd; r0 += MF_ROWS) {
    const int nr = (int)min((long long)MF_ROWS, rEnd - r0);
    __syncthreads();
    // ---- stage the prefetched chunk
#pragma unroll
    for (int q = 0; q < ITEMS; ++q) {
      const int it = tid + 256 * q;
      if (it < (MF_ROWS / 4) * Dp) {
        const int rq = it / Dp, d = it - rq * Dp;
        const float gd = d < D ? g[d] : 0.f;
        H h1[4], l1[4], h2[4], l2[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) {
          const int r = 4 * rq + i;
          const float xc = (r < nr && d < D) ? xpre[q][i] - gd : 0.f;
          split16<H>(xc * sc.s1, h1[i], l1[i]);
          split16<H>(xc * xc * sc.s2, h2[i], l2[i]);
          sm.xr[r][d] = h2[i];
          sm.xr[r][Dp + d] = h1[i];
          sm.xr[r][2 * Dp + d] = l2[i];
          sm.xr[r][3 * Dp + d] = l1[i];
        }
        store4<H>(&sm.xct[d][4 * rq], h1[0], h1[1], h1[2], h1[3]);
        store4<H>(&sm.xct[Dp + d][4 * rq], l1[0], l1[1], l1[2], l1[3]);
        store4<H>(&sm.xct[2 * Dp + d][4 * rq], h2[0], h2[1], h2[2], h2[3]);
        store4<H>(&sm.xct[3 * Dp + d][4 * rq], l2[0], l2[1], l2[2], l2[3]);
      }
    }
    if (tid < MF_ROWS) {
      const bool ok = tid < nr;
      sm.rm[tid] = ok ? row_m[r0 + tid] : 0.f;
      sm.rinv[tid] = ok ? 1.0f / row_s[r0 + tid] : 0.f;
      sm.ra[tid] = ok ? row_arg[r0 + tid] : -1;
    }
    if (r0 + MF_ROWS < rEnd) prefetch(r0 + MF_ROWS);
    __syncthreads();

    // ---- logits of rows 32wr.. x comps 64wc..
    f32x4 acc[2][4];
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[t][c] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int j = 0; j < NS; ++j) {
      const int f0 = 32 * j + 8 * lg;
      V8 ah[2], al[2];
#pragma unroll
      for (int t = 0; t < 2; ++t) {
        const int rl = 32 * wr + 16 * t + lr;
        ah[t] = *reinterpret_cast<const V8*>(&sm.xr[rl][f0]);
        al[t] = *reinterpret_cast<const V8*>(&sm.xr[rl][2 * Dp + f0]);
      }
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[t][c] = mfma16(ah[t], whi[c][j], acc[t][c]);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[t][c] = mfma16(al[t], whi[c][j], acc[t][c]);
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int c = 0; c < 4; ++c) acc[t][c] = mfma16(ah[t], wlo[c][j], acc[t][c]);
    }
    // ---- posteriors
#pragma unroll
    for (int t = 0; t < 2; ++t)
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int cl = 64 * wc + 16 * c + lr;
        const int comp = k0 + cl;
        H hq[4];
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const int rl = 32 * wr + 16 * t + 4 * lg + r;
          float q = 0.f;
          if (rl < nr && comp < K && bcomp[c] != -INFINITY) {
            if (fast) {
              q = (comp == sm.ra[rl]) ? 1.f : 0.f;
            } else {
              const float l = acc[t][c][r] + bcomp[c];
              q = __builtin_amdgcn_exp2f((l - sm.rm[rl]) * kLog2e) * sm.rinv[rl];
            }
            if (q < qmin) q = 0.f; else terms += 1;
          }
          hq[r] = H(q * sc.sq);
          s0l[c] += (float)hq[r];
        }
        store4<H>(&sm.qt[cl][32 * wr + 16 * t + 4 * lg], hq[0], hq[1], hq[2], hq[3]);
      }
    __syncthreads();

    // ---- moments: planes 2pp, 2pp+1 for components 64ch.., 2 row steps
#pragma unroll
    for (int ks = 0; ks < 2; ++ks) {
      V8 aq[4];
#pragma unroll
      for (int ct = 0; ct < 4; ++ct)
        aq[ct] = *reinterpret_cast<const V8*>(&sm.qt[64 * ch + 16 * ct + lr][32 * ks + 8 * lg]);
#pragma unroll
      for (int p2 = 0; p2 < 2; ++p2)
#pragma unroll
        for (int ft = 0; ft < NS; ++ft) {
          const V8 bx = *reinterpret_cast<const V8*>(&sm.xct[(2 * pp + p2) * Dp + 16 * ft + lr][32 * ks + 8 * lg]);
#pragma unroll
          for (int ct = 0; ct < 4; ++ct) macc[ct][ft] = mfma16(aq[ct], bx, macc[ct][ft]);
        }
    }
  }

  // ---- posterior mass per component (lanes lg, then the two row halves)
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    float sv = s0l[c];
    sv += __shfl_xor(sv, 16, 64);
    sv += __shfl_xor(sv, 32, 64);
    s0l[c] = sv;
  }
  __syncthreads();
  if (lg == 0) {
#pragma unroll
    for (int c = 0; c < 4; ++c) sm.s0[wr][64 * wc + 16 * c + lr] = s0l[c];
  }
  __syncthreads();

  // ---- finalisation: every wave dumps its (plane pair, component half)
  // moments into LDS as FP32 -- the staging buffers are free now -- and all
  // threads then finish (component, dimension) pairs.
  float* F = reinterpret_cast<float*>(&sm.xr[0][0]);  // [2 planes: S1, S2][128 comps][Dp]; spans xr, xct, qt
#pragma unroll
  for (int ct = 0; ct < 4; ++ct)
#pragma unroll
    for (int ft = 0; ft < NS; ++ft)
#pragma unroll
      for (int r = 0; r < 4; ++r)
        F[((size_t)pp * MF_COMPS + 64 * ch + 16 * ct + 4 * lg + r) * Dp + 16 * ft + lr] = macc[ct][ft][r];
  __syncthreads();
  const double invq = 1.0 / (double)sc.sq;
  const bool sqrtFlag = (flags & 1) != 0;
  T* encU = enc + (size_t)img * 2 * K * D;
  T* encV = encU + (size_t)K * D;
  double nrm = 0.0;
#pragma nounroll
  for (int idx = tid; idx < MF_COMPS * Dp; idx += 256) {
    const int cl = idx / Dp, d = idx - cl * Dp;
    const int k = k0 + cl;
    if (k >= K || d >= D) continue;
    const double S0 = ((double)sm.s0[0][cl] + (double)sm.s0[1][cl]) * invq;
    const double S1 = (double)F[(size_t)cl * Dp + d] * invq / (double)sc.s1;
    const double S2 = (double)F[((size_t)MF_COMPS + cl) * Dp + d] * invq / (double)sc.s2;
    const double prior = (double)priors[k];
    T u = T(0), v = T(0);
    if (prior >= 1e-6) {
      const double mu = (double)mup[(size_t)k * Dp + d];
      double uu = (double)isg[(size_t)k * Dp + d] * (S1 - mu * S0);
      double vv = (double)sinv[(size_t)k * Dp + d] * (S2 - 2.0 * mu * S1 + mu * mu * S0) - S0;
      if (nImg > 0) {
        uu *= 1.0 / ((double)nImg * sqrt(prior));
        vv *= 1.0 / ((double)nImg * sqrt(2.0 * prior));
      }
      u = (T)uu;
      v = (T)vv;
    }
    if (sqrtFlag) { u = signed_sqrt(u); v = signed_sqrt(v); }
    encU[(size_t)k * D + d] = u;
    encV[(size_t)k * D + d] = v;
    nrm += (double)u * (double)u + (double)v * (double)v;
  }
  const double bn = block_sum_256(nrm, red);
  if (tid == 0) normpart[(size_t)img * numKB + kb] = bn;
  if (numTerms) {
    const double bt = block_sum_256((double)terms, red);
    if (tid == 0) atomicAdd(&numTerms[img], (unsigned long long)bt);
  }
}

// ---------------------------------------------------------------------------

template <typename H, int NS>
static void run_mfma(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                     const float* x, const uint64_t* offsets, size_t numImages, int flags,
                     unsigned long long* numTerms, uint64_t total) {
  constexpr int Dp = NS * 16;
  const int Kp = (int)((K + MF_COMPS - 1) / MF_COMPS) * MF_COMPS;
  const bool isHalf = sizeof(H) == 2 && std::is_same<H, _Float16>::value;
  // tables: [scales][g Dp][bias Kp][mup, isg, sinv Kp*Dp][Whi, Wlo Kp*2Dp]
  auto al = [](size_t b) { return (b + 255) & ~size_t(255); };
  const size_t bS = al(sizeof(MfmaScales)), bG = al(sizeof(float) * Dp), bB = al(sizeof(float) * Kp),
               bT = al(sizeof(float) * (size_t)Kp * Dp), bW = al(sizeof(H) * (size_t)Kp * 2 * Dp);
  unsigned char* base = (unsigned char*)scratch(SLOT_PREP, bS + bG + bB + 3 * bT + 2 * bW);
  MfmaScales* scp = (MfmaScales*)base;
  float* g = (float*)(base + bS);
  float* bias = (float*)(base + bS + bG);
  float* mup = (float*)(base + bS + bG + bB);
  float* isg = mup + bT / sizeof(float);
  float* sinv = isg + bT / sizeof(float);
  H* Whi = (H*)(base + bS + bG + bB + 3 * bT);
  H* Wlo = (H*)(base + bS + bG + bB + 3 * bT + bW);
  VLHIP_CHECK(hipMemsetAsync(g, 0, bG, ctx().stream));
  VLHIP_LAUNCH("mfma_center", (mfma_center_kernel<float, H>), dim3(1), dim3(256), 0, means, covs, priors, (int)D,
               (int)K, g, scp, isHalf ? 1 : 0);
  VLHIP_LAUNCH("mfma_tables", (mfma_tables_kernel<float, H>), dim3((Kp + 127) / 128), dim3(128), 0, means, covs,
               priors, (int)D, Dp, (int)K, Kp, g, scp, Whi, Wlo, bias, mup, isg, sinv);

  unsigned char* rsb = (unsigned char*)scratch(SLOT_ROWSTAT, (size_t)(total + 1) * 12 + 512);
  float* rm = (float*)rsb;
  float* rs = rm + total + 1;
  int* ra = (int*)(rs + total + 1);
  if (total > 0) {
    const unsigned blocks = (unsigned)((total + ML_ROWS - 1) / ML_ROWS);
    VLHIP_LAUNCH("mfma_lse", (mfma_lse_kernel<H, NS>), dim3(blocks), dim3(256), 0, x, (long long)total, (int)D,
                 Kp, Whi, Wlo, bias, scp, g, rm, rs, ra);
  }
  const int numKB = Kp / MF_COMPS;
  double* normpart = (double*)scratch(SLOT_PARTIAL, numImages * (size_t)numKB * sizeof(double));
  if (numTerms) VLHIP_CHECK(hipMemsetAsync(numTerms, 0, numImages * sizeof(unsigned long long), ctx().stream));
  const size_t shmem = sizeof(MfLds<H, NS>);
  auto kern = mfma_fisher_kernel<float, H, NS>;
  VLHIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  const int padImages = numImages >= 8 ? (int)((numImages + 7) / 8) * 8 : (int)numImages;
  VLHIP_LAUNCH("mfma_fisher", kern, dim3((unsigned)((size_t)padImages * numKB)), dim3(256), shmem, x, offsets,
               (int)numImages, (int)D, (int)K, Whi, Wlo, bias, scp, g, mup, isg, sinv, priors, rm, rs, ra, flags, enc,
               normpart, numTerms, numKB);
  if (flags & 2) {
    launch_fisher_normalize(enc, normpart, numKB, numImages, (size_t)2 * K * D);
  }
}

template <typename H>
static void mfma_dispatch(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                          const float* x, const uint64_t* offsets, size_t numImages, int flags,
                          unsigned long long* numTerms) {
  uint64_t total = 0;
  download(&total, offsets + numImages, sizeof(uint64_t));
  const size_t Dp = (D + 15) / 16 * 16;
  switch (Dp) {
    case 16: run_mfma<H, 1>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, total); break;
    case 32: run_mfma<H, 2>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, total); break;
    case 48: run_mfma<H, 3>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, total); break;
    case 64: run_mfma<H, 4>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, total); break;
    case 80: run_mfma<H, 5>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, total); break;
    case 96: run_mfma<H, 6>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms, total); break;
    default: fatal("vlhip_fisher_encode_batch", __FILE__, __LINE__,
                   "the matrix-core Fisher encoder supports dimension <= 96");
  }
}

void fisher_split16(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                    const float* x, const uint64_t* offsets, size_t numImages, int flags,
                    unsigned long long* numTerms, bool fp16) {
  VLHIP_REQUIRE(K >= 1 && D >= 1, "numClusters and dimension must be positive");
  if (numImages == 0) return;
  if (fp16) mfma_dispatch<_Float16>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms);
  else mfma_dispatch<__bf16>(enc, means, covs, priors, D, K, x, offsets, numImages, flags, numTerms);
}

}  // namespace vlhip
