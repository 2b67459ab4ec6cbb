// kmeans_mfma.hip -- exact L2 quantizer (vl/kmeans.c:655-702) screened on the
// matrix cores, for float data with D <= 128.
//
// The assignment must be bitwise the reference's: the argmin, with the first
// index among equal minima, of the squared distances as the host's SIMD kernel
// rounds them (lane-order partial sums, separate multiply and add;
// mathop_avx.c:43-111, mathop_sse2.c:43-100, mathop.c:296-307).  Evaluating
// that exact distance for every (descriptor, centre) pair is VALU work of
// 3DK per descriptor.  Here the matrix cores first compute, for every pair,
//   a_ik = ||c_k||^2 - 2 x_i.c_k    (the distance less the row's ||x_i||^2)
// with the dot product in split BF16 (x = x_hi + x_lo, three products
// x_hi c_hi + x_lo c_hi + x_hi c_lo, FP32 accumulation), whose distance to the
// exact float distance e_ik - ||x_i||^2 is bounded by eps_i = 2^-13 (||x_i|| + max_k ||c_k||)^2
// (a margin of about 4x over the sum of: the split's representation and
// dropped-product errors 3 * 2^-18 |x_d c_d| per term, the FP32 accumulation
// of the 3 * ceil(D/32) matrix instructions, the roundings of a_ik, and the
// reference's own float rounding of e_ik).  Every centre whose exact distance
// is the minimum then has a_ik <= min_k a_ik + 2 eps_i, so the exact distance
// is evaluated only for those candidates (usually one or two) and the
// smallest, lowest index first, is the reference's assignment.  The screen
// takes one sweep over the centres: each of the 16 lanes that share a row
// keeps its smallest screened value with its centre and its second smallest
// value, and the row's list is assembled from its lanes at the end: a lane's
// centre if its smallest value is within 2 eps of the row's minimum, or every
// centre of that lane if its second smallest is too.  A row with no candidate
// (non-finite data) or more than QM_CAP list entries is scanned exactly over
// all centres.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

#include <algorithm>

namespace vlhip {

namespace {

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

__device__ __forceinline__ f32x4 mfma_bf16(const bf16x8& a, const bf16x8& b, const f32x4& c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(a, b, c, 0, 0, 0);
}

constexpr int QM_WAVES = 4;  // waves per workgroup, 32 descriptors each
constexpr int QM_CAP = 16;   // candidate slots per descriptor
constexpr int QM_BLOCK = 64; // centres per sweep step (4 tiles of 16)

// Squared L2 distance of one pair in the host kernel's order (see the header
// of kmeans_kernels.hip): LANES partial sums, fixed horizontal tree, remainder.
template <int LANES>
__device__ __forceinline__ float exact_l2(const float* __restrict__ a, const float* __restrict__ b, int D) {
#pragma clang fp contract(off)
  const int Dv = (LANES > 1) ? (D / LANES) * LANES : 0;
  float acc[LANES];
#pragma unroll
  for (int l = 0; l < LANES; ++l) acc[l] = 0.f;
  if (LANES >= 4 && ((reinterpret_cast<uintptr_t>(a) | reinterpret_cast<uintptr_t>(b)) & 15) == 0) {
    // the same operations, fed by 16-byte loads
    const f32x4* a4 = reinterpret_cast<const f32x4*>(a);
    const f32x4* b4 = reinterpret_cast<const f32x4*>(b);
#pragma unroll 4
    for (int d = 0; d < Dv; d += LANES) {
#pragma unroll
      for (int h = 0; h < LANES / 4; ++h) {
        const f32x4 av = a4[d / 4 + h], bv = b4[d / 4 + h];
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          const float delta = av[e] - bv[e];
          const float d2 = delta * delta;
          acc[4 * h + e] = acc[4 * h + e] + d2;
        }
      }
    }
  } else {
    for (int d = 0; d < Dv; d += LANES) {
#pragma unroll
      for (int l = 0; l < LANES; ++l) {
        const float delta = a[d + l] - b[d + l];
        const float d2 = delta * delta;
        acc[l] = acc[l] + d2;
      }
    }
  }
  float s;
  if constexpr (LANES == 8) s = ((acc[0] + acc[1]) + (acc[4] + acc[5])) + ((acc[2] + acc[3]) + (acc[6] + acc[7]));
  else if constexpr (LANES == 4) s = (acc[0] + acc[2]) + (acc[1] + acc[3]);
  else if constexpr (LANES == 2) s = acc[0] + acc[1];
  else s = 0.f;
  for (int d = Dv; d < D; ++d) {
    const float delta = a[d] - b[d];
    const float d2 = delta * delta;
    s = s + d2;
  }
  return s;
}

__device__ __forceinline__ bool better(float v, int k, float bv, int bk) { return v < bv || (v == bv && k < bk); }

}  // namespace

__global__ void qm_zero_kernel(unsigned* __restrict__ p) {
  if (threadIdx.x == 0) *p = 0u;
}

// Centre table: split BF16 rows [Kp][Dq] (Dq = D rounded up to 32, zero
// padding), ||c_k||^2 (FP64 sum rounded to float; +inf for padding centres)
// and the bits of max_k ||c_k|| (atomic max of non-negative floats).
__global__ __launch_bounds__(64) void qm_table_kernel(const float* __restrict__ centers, int D, int Dq, int K, int Kp,
                                                      __bf16* __restrict__ chi, __bf16* __restrict__ clo,
                                                      float* __restrict__ cc, unsigned* __restrict__ maxcBits) {
  const int k = blockIdx.x, lane = threadIdx.x;
  double s = 0.0;
  for (int d = lane; d < Dq; d += 64) {
    const float v = (k < K && d < D) ? centers[(size_t)k * D + d] : 0.f;
    const __bf16 h = (__bf16)v;
    chi[(size_t)k * Dq + d] = h;
    clo[(size_t)k * Dq + d] = (__bf16)(v - (float)h);
    s += (double)v * (double)v;
  }
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) s += __shfl_xor(s, o, 64);
  if (lane == 0) {
    cc[k] = k < K ? (float)s : INFINITY;
    if (k < K) {
      const float n = (float)sqrt(s);
      if (n > 0.f) atomicMax(maxcBits, __float_as_uint(n));
    }
  }
}

// Screen statistics, counted while kernel timing is on: rows, exact distances
// of listed entries (a centre, or all centres of a lane), rows scanned over
// all centres.
__device__ unsigned long long qm_screen_stats[3];

template <int NQ, int LANES>
__global__ __launch_bounds__(QM_WAVES * 64, 2) void qm_quantize_kernel(
    const float* __restrict__ x, const float* __restrict__ centers, long long N, int D, int K, int Kp,
    const __bf16* __restrict__ chi, const __bf16* __restrict__ clo, const float* __restrict__ cc,
    const unsigned* __restrict__ maxcBits, uint32_t* __restrict__ assign, float* __restrict__ dist, int stats) {
  constexpr int Dq = NQ * 32;
  constexpr int WS = Dq + 8;                       // LDS row stride (halves) of a staged centre
  constexpr int NT = QM_WAVES * 64;
  constexpr int PARTW = QM_BLOCK * Dq / 8;         // 16-byte words of one part (hi or lo) of a block
  constexpr int STG = 2 * PARTW / NT;              // words per thread per block
  static_assert(STG * NT == 2 * PARTW, "staging must divide evenly");
  __shared__ __align__(16) __bf16 wsh[2][2][QM_BLOCK][WS];  // [buffer][hi, lo][centre][feature]
  __shared__ __align__(16) float ccs[2][QM_BLOCK];  // ||c||^2 of the staged block (+inf for padding)
  __shared__ float epss[QM_WAVES][32];
  __shared__ int cnt[QM_WAVES][32];
  __shared__ int cand[QM_WAVES][32][QM_CAP];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const long long row0 = ((long long)blockIdx.x * QM_WAVES + w) * 32;

  // centre blocks double-buffered in LDS, the next one fetched into
  // registers while the current one is multiplied (one barrier per block);
  // the first block's loads are issued before the descriptors'
  const int nb = Kp / QM_BLOCK;
  int srcOff[STG], dstOff[STG];
#pragma unroll
  for (int q = 0; q < STG; ++q) {
    const int i = tid + NT * q;
    const int part = i / PARTW, rem = i - part * PARTW;
    const int c = rem / (Dq / 8), q8 = rem - c * (Dq / 8);
    srcOff[q] = part * Kp * Dq + c * Dq + 8 * q8;  // + block * 64 * Dq, halves into [chi | clo]
    dstOff[q] = (part * QM_BLOCK + c) * WS + 8 * q8;
  }
  typedef unsigned u32x4 __attribute__((ext_vector_type(4)));
  u32x4 pre[STG];
  f32x4 prec;  // threads 0-15: four ||c||^2 of the block
#define QM_STAGE_LOAD(BLK)                                                              \
  {                                                                                     \
    const int blk_ = (BLK);                                                             \
    _Pragma("unroll") for (int q = 0; q < STG; ++q) {                                   \
      const int o = srcOff[q] + blk_ * QM_BLOCK * Dq;                                   \
      const __bf16* src = o < Kp * Dq ? chi + o : clo + (o - Kp * Dq);                  \
      pre[q] = *reinterpret_cast<const u32x4*>(src);                                    \
    }                                                                                   \
    if (tid < QM_BLOCK / 4) prec = *reinterpret_cast<const f32x4*>(cc + blk_ * QM_BLOCK + 4 * tid); \
  }
#define QM_STAGE_STORE(BUF)                                                             \
  {                                                                                     \
    __bf16* dst_ = &wsh[(BUF)][0][0][0];                                                \
    _Pragma("unroll") for (int q = 0; q < STG; ++q)                                     \
      *reinterpret_cast<u32x4*>(dst_ + dstOff[q]) = pre[q];                             \
    if (tid < QM_BLOCK / 4) *reinterpret_cast<f32x4*>(&ccs[(BUF)][4 * tid]) = prec;     \
  }
  QM_STAGE_LOAD(0)

  // A fragments: row 16t + lr, features 32j + 8lg .. +7, split hi / lo; the
  // squared norm of the row in FP64 from the same coordinates.  Whole runs of
  // 8 features in 16-byte loads when rows are 16-byte aligned.
  const bool vec = (D & 3) == 0 && (reinterpret_cast<uintptr_t>(x) & 15) == 0;
  // every load is issued (out-of-range ones from the first row, their values
  // replaced by zeros) before any is used, so that none waits for another
  float xv[2][NQ][8];
  if (vec) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const long long row = row0 + 16 * t + lr;
      const float* xr = x + (row < N ? row : 0) * D;
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        const int d0 = 32 * j + 8 * lg;
        const bool in0 = row < N && d0 + 4 <= D, in1 = row < N && d0 + 8 <= D;
        const f32x4 u = *reinterpret_cast<const f32x4*>(in0 ? xr + d0 : x);
        const f32x4 v = *reinterpret_cast<const f32x4*>(in1 ? xr + d0 + 4 : x);
#pragma unroll
        for (int e = 0; e < 4; ++e) {
          xv[t][j][e] = in0 ? u[e] : 0.f;
          xv[t][j][4 + e] = in1 ? v[e] : 0.f;
        }
      }
    }
  } else {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const long long row = row0 + 16 * t + lr;
      const float* xr = x + (row < N ? row : 0) * D;
#pragma unroll
      for (int j = 0; j < NQ; ++j)
#pragma unroll
        for (int e = 0; e < 8; ++e) {
          const int d = 32 * j + 8 * lg + e;
          const bool in = row < N && d < D;
          const float u = *(in ? xr + d : x);
          xv[t][j][e] = in ? u : 0.f;
        }
    }
  }
  bf16x8 ahi[2][NQ], alo[2][NQ];
  double nrm[2] = {0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        const __bf16 h = (__bf16)xv[t][j][e];
        ahi[t][j][e] = h;
        alo[t][j][e] = (__bf16)(xv[t][j][e] - (float)h);
        nrm[t] += (double)xv[t][j][e] * (double)xv[t][j][e];
      }
    }
    nrm[t] += __shfl_xor(nrm[t], 16, 64);
    nrm[t] += __shfl_xor(nrm[t], 32, 64);
  }
  const double maxc = (double)__uint_as_float(*maxcBits);
  if (lg == 0) {
#pragma unroll
    for (int t = 0; t < 2; ++t) {
      const double m = sqrt(nrm[t]) + maxc;
      epss[w][16 * t + lr] = (float)(ldexp(m * m, -13) * (1.0 + 1e-6));
    }
  }
  if (lane < 32) cnt[w][lane] = 0;
  QM_STAGE_STORE(0)
  if (nb > 1) QM_STAGE_LOAD(1)
  __syncthreads();
  // accumulator layout: rows 16t + 4lg + r, centre 16c + lr of the block
  float w2[2][4];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) w2[t][r] = 2.f * epss[w][16 * t + 4 * lg + r];

  // One sweep over the centre blocks.  Each lane sees, for each of its 8
  // rows, the centres 16c + lr of every block (K/16 of them), and keeps the
  // smallest screened value m1 with its centre k1 and the second smallest
  // value m2, without branches: per value one multiply-add, a min of a max,
  // a compare and a select.
  float m1[2][4], m2[2][4];
  int k1[2][4];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      m1[t][r] = INFINITY;
      m2[t][r] = INFINITY;
      k1[t][r] = -1;
    }

  for (int st = 0; st < nb; ++st) {
    const int k0 = st * QM_BLOCK;
    if (st + 1 < nb) {
      QM_STAGE_STORE((st + 1) & 1)  // that buffer was last read in step st - 1, before the barrier
      if (st + 2 < nb) QM_STAGE_LOAD(st + 2)
    }
    const __bf16* wb = &wsh[st & 1][0][0][0];
    {
      f32x4 acc[4][2];
#pragma unroll
      for (int c = 0; c < 4; ++c) acc[c][0] = acc[c][1] = f32x4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
      for (int j = 0; j < NQ; ++j) {
        bf16x8 bh[4], bl[4];
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          bh[c] = *reinterpret_cast<const bf16x8*>(wb + (16 * c + lr) * WS + 32 * j + 8 * lg);
          bl[c] = *reinterpret_cast<const bf16x8*>(wb + (QM_BLOCK + 16 * c + lr) * WS + 32 * j + 8 * lg);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c)
#pragma unroll
          for (int t = 0; t < 2; ++t) {
            acc[c][t] = mfma_bf16(ahi[t][j], bh[c], acc[c][t]);
            acc[c][t] = mfma_bf16(alo[t][j], bh[c], acc[c][t]);
            acc[c][t] = mfma_bf16(ahi[t][j], bl[c], acc[c][t]);
          }
      }
      // the screen without the row's constant ||x||^2; padding centres
      // (last block only) are +inf and change nothing
      float ck[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) ck[c] = ccs[st & 1][16 * c + lr];
#pragma unroll
      for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float a = ck[c] - 2.f * acc[c][t][r];
            m2[t][r] = fminf(m2[t][r], fmaxf(m1[t][r], a));
            const bool lt = a < m1[t][r];
            k1[t][r] = lt ? k0 + 16 * c + lr : k1[t][r];
            m1[t][r] = lt ? a : m1[t][r];
          }
    }
    __syncthreads();
  }
#undef QM_STAGE_LOAD
#undef QM_STAGE_STORE

  // The row's minimum M over its 16 lanes and the limit M + 2 eps.  A lane
  // with m1 inside the limit lists k1; if m2 is inside too, the lane may hold
  // more centres inside it and lists itself instead (entry -1 - lane: every
  // centre of that lane is evaluated exactly).  A lane with m2 outside has no
  // centre inside the limit but k1, since all its other values are >= m2.
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      const float lim = row16_min(m1[t][r]) + w2[t][r];
      const int rw = 16 * t + 4 * lg + r;
      if (k1[t][r] >= 0 && m1[t][r] <= lim) {
        const int slot = atomicAdd(&cnt[w][rw], 1);
        if (slot < QM_CAP) cand[w][rw][slot] = m2[t][r] <= lim ? -1 - lr : k1[t][r];
      }
    }
  __syncthreads();

  // Exact distances of the single listed centres: lanes l and l + 32 serve
  // row l & 31.  Rows with a lane entry, with no entry (non-finite data) or
  // with more entries than the slots hold are marked wide.
  const int rw = lane & 31, half = lane >> 5;
  const long long row = row0 + rw;
  float bv = INFINITY;
  int bk = 0x7fffffff;
  int work = 0, fulls = 0;  // exact distances of listed entries; rows scanned over all centres
  bool wide = false;
  if (row < N) {
    const float* xr = x + row * D;
    const int n = cnt[w][rw];
    if (n >= 1 && n <= QM_CAP) {
      for (int s = half; s < n; s += 2) {
        const int k = cand[w][rw][s];
        if (k < 0) {
          wide = true;
          continue;
        }
        const float e = exact_l2<LANES>(xr, centers + (size_t)k * D, D);
        if (better(e, k, bv, bk)) { bv = e; bk = k; }
        ++work;
      }
    } else {
      wide = true;
    }
  }
  {
    const float ov = __shfl_xor(bv, 32, 64);
    const int ok = __shfl_xor(bk, 32, 64);
    if (better(ov, ok, bv, bk)) { bv = ov; bk = ok; }
  }
  // Wide rows one at a time, their centres spread over the whole wave: the
  // centres of each listed lane, or all centres.  A lane's partial best is
  // never NaN (NaN is never better than the initial +inf), so the (value,
  // index) minimum over the wave does not depend on the reduction order.
  const unsigned long long wb = __ballot(wide);
  unsigned wideRows = (unsigned)(wb | (wb >> 32));
  while (wideRows) {
    const int rr = __builtin_ctz(wideRows);
    wideRows &= wideRows - 1;
    const float* xr = x + (row0 + rr) * D;
    const int n = cnt[w][rr];
    float v = INFINITY;
    int kv = 0x7fffffff;
    if (n >= 1 && n <= QM_CAP) {
      for (int s = 0; s < n; ++s) {
        const int e0 = cand[w][rr][s];
        if (e0 >= 0) continue;
        for (int k = -1 - e0 + 16 * lane; k < K; k += 16 * 64) {
          const float e = exact_l2<LANES>(xr, centers + (size_t)k * D, D);
          if (better(e, k, v, kv)) { v = e; kv = k; }
          ++work;
        }
      }
    } else {
      for (int k = lane; k < K; k += 64) {
        const float e = exact_l2<LANES>(xr, centers + (size_t)k * D, D);
        if (better(e, k, v, kv)) { v = e; kv = k; }
      }
      if (lane == 0) ++fulls;
    }
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      const float ov = __shfl_xor(v, o, 64);
      const int ok = __shfl_xor(kv, o, 64);
      if (better(ov, ok, v, kv)) { v = ov; kv = ok; }
    }
    if (rw == rr && better(v, kv, bv, bk)) { bv = v; bk = kv; }
  }
  if (half == 0 && row < N) {
    // no finite distance at all: the first centre, as the exhaustive kernel
    assign[row] = (uint32_t)(bk == 0x7fffffff ? 0 : bk);
    if (dist) dist[row] = bv;
  }
  if (stats) {
    // summed over the wave, one atomic per counter and wave
    int v0 = half == 0 && row < N ? 1 : 0, v1 = work, v2 = fulls;
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
      v0 += __shfl_xor(v0, o, 64);
      v1 += __shfl_xor(v1, o, 64);
      v2 += __shfl_xor(v2, o, 64);
    }
    if (lane == 0) {
      atomicAdd(&qm_screen_stats[0], (unsigned long long)v0);
      atomicAdd(&qm_screen_stats[1], (unsigned long long)v1);
      atomicAdd(&qm_screen_stats[2], (unsigned long long)v2);
    }
  }
}

// Bytes of the centre table [chi | clo | cc | max bits] of a screened shape
// (0 when the shape is not screened).
static constexpr size_t qm_table_bytes(size_t D, size_t K) {
  return 2 * (((K + QM_BLOCK - 1) / QM_BLOCK * QM_BLOCK * ((D + 31) / 32 * 32) * 2 + 255) & ~size_t(255)) +
         (((K + QM_BLOCK - 1) / QM_BLOCK * QM_BLOCK * sizeof(float) + 255) & ~size_t(255)) + 256;
}
size_t quantize_screened_bytes(size_t D, size_t K) {
  if (D > 128 || K < 64) return 0;
  return qm_table_bytes(D, K);
}

bool quantize_screened(uint32_t* assign, float* dist, const float* x, const float* centers, size_t N, size_t D,
                       size_t K, int lanes, void* ws) {
  if (D > 128 || K < 64 || N == 0) return false;
  const int NQ = (int)((D + 31) / 32);
  const int Dq = NQ * 32;
  const int Kp = (int)((K + QM_BLOCK - 1) / QM_BLOCK) * QM_BLOCK;
  const size_t bW = ((size_t)Kp * Dq * sizeof(__bf16) + 255) & ~size_t(255);
  const size_t bC = ((size_t)Kp * sizeof(float) + 255) & ~size_t(255);
  // the caller's workspace (encode plans, the no-allocation entry point) or the
  // library's grow-only table
  unsigned char* base = ws ? (unsigned char*)ws : (unsigned char*)scratch(SLOT_QTAB, 2 * bW + bC + 256);
  __bf16* chi = (__bf16*)base;
  __bf16* clo = (__bf16*)(base + bW);
  float* cc = (float*)(base + 2 * bW);
  unsigned* maxc = (unsigned*)(base + 2 * bW + bC);
  VLHIP_LAUNCH("qm_zero", (qm_zero_kernel), dim3(1), dim3(64), 0, maxc);
  VLHIP_LAUNCH("qm_table", qm_table_kernel, dim3((unsigned)Kp), dim3(64), 0, centers, (int)D, Dq, (int)K, Kp, chi, clo,
               cc, maxc);
  const unsigned blocks = (unsigned)((N + 32 * QM_WAVES - 1) / (32 * QM_WAVES));
  const int stats = ctx().timing && !ctx().capturing;
#define QM_LAUNCH(NQV, LV)                                                                                        \
  VLHIP_LAUNCH("kmeans_quantize_mfma", (qm_quantize_kernel<NQV, LV>), dim3(blocks), dim3(QM_WAVES * 64), 0, x,      \
               centers, (long long)N, (int)D, (int)K, Kp, chi, clo, cc, maxc, assign, dist, stats)
#define QM_LANES(NQV)                               \
  switch (lanes) {                                  \
    case 8: QM_LAUNCH(NQV, 8); break;               \
    case 4: QM_LAUNCH(NQV, 4); break;               \
    case 2: QM_LAUNCH(NQV, 2); break;               \
    default: QM_LAUNCH(NQV, 1); break;              \
  }
  switch (NQ) {
    case 1: QM_LANES(1) break;
    case 2: QM_LANES(2) break;
    case 3: QM_LANES(3) break;
    default: QM_LANES(4) break;
  }
#undef QM_LANES
#undef QM_LAUNCH
  return true;
}

// ---------------------------------------------------------------------------
// No-allocation form of the quantizer, for callers that cannot allocate device
// memory (code under graph capture, serving loops).  The screened kernel's
// centre table lives in a static device array, one copy per GPU, held until
// the process ends; its device address is looked up once per device.  Every
// call on a GPU shares that GPU's table, so calls that overlap on one GPU race.
constexpr size_t kNoAllocMaxCenters = 4096;
constexpr size_t kNoAllocTableBytes = qm_table_bytes(128, kNoAllocMaxCenters);  // 2.0 MiB

}  // namespace vlhip

static __device__ __align__(256) unsigned char noalloc_scratch[vlhip::kNoAllocTableBytes];

namespace vlhip {

constexpr int kNoAllocMaxDevices = 64;
static void* g_noalloc_table[kNoAllocMaxDevices] = {};

// Device address of this GPU's table (looked up on the first call on the
// device, with the device current).
static void* noalloc_table() {
  int dev = 0;
  VLHIP_CHECK(hipGetDevice(&dev));
  VLHIP_REQUIRE(dev >= 0 && dev < kNoAllocMaxDevices, "device index out of range");
  if (!g_noalloc_table[dev]) {
    void* p = nullptr;
    VLHIP_CHECK(hipGetSymbolAddress(&p, HIP_SYMBOL(noalloc_scratch)));
    g_noalloc_table[dev] = p;
  }
  return g_noalloc_table[dev];
}

}  // namespace vlhip

using namespace vlhip;

extern "C" void vlhip_kmeans_quantize_noalloc(int dtype, uint32_t* assign, void* dist, void const* data,
                                              void const* centers, size_t N, size_t D, size_t K, int lanes) {
  // argument checks come first: they need no device
  VLHIP_REQUIRE(dtype == 1 || dtype == 2, "unknown data type");
  VLHIP_REQUIRE(lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8, "lanes must be 1, 2, 4 or 8");
  VLHIP_REQUIRE(K >= 1 && D >= 1, "numCenters and dimension must be positive");
  const bool screened = dtype == 1 && quantize_screened_bytes(D, K) != 0;
  VLHIP_REQUIRE(!screened || K <= kNoAllocMaxCenters,
                "vlhip_kmeans_quantize_noalloc: float data with dimension <= 128 takes at most 4096 centers "
                "(the size of its static table)");
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  // no timing events: the launches below create nothing
  const bool savedTiming = ctx().timing;
  ctx().timing = false;
  if (dtype == 1) {
    if (screened && N > 0) {
      quantize_screened(assign, (float*)dist, (const float*)data, (const float*)centers, N, D, K, lanes,
                        noalloc_table());
    } else {
      // the exhaustive kernel, which uses no work buffer
      quantize_t<float>(assign, (float*)dist, (const float*)data, (const float*)centers, N, D, K, lanes);
    }
  } else {
    quantize_t<double>(assign, (double*)dist, (const double*)data, (const double*)centers, N, D, K, lanes);
  }
  ctx().timing = savedTiming;
}

extern "C" void vlhip_kmeans_screen_stats(uint64_t* counts, int reset) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  unsigned long long h[3] = {0, 0, 0};
  VLHIP_CHECK(hipStreamSynchronize(ctx().stream));
  VLHIP_CHECK(hipMemcpyFromSymbol(h, HIP_SYMBOL(qm_screen_stats), sizeof(h), 0, hipMemcpyDeviceToHost));
  if (counts)
    for (int i = 0; i < 3; ++i) counts[i] = (uint64_t)h[i];
  if (reset) {
    const unsigned long long z[3] = {0, 0, 0};
    VLHIP_CHECK(hipMemcpyToSymbol(HIP_SYMBOL(qm_screen_stats), z, sizeof(z), 0, hipMemcpyHostToDevice));
  }
}
