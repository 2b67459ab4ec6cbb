// kmeans_mfma.hip -- exact L2 quantizer (vl/kmeans.c:655-702) screened on the
// matrix cores, for float data with D <= 128.
//
// The assignment must be bitwise the reference's: the argmin, with the first
// index among equal minima, of the squared distances as the host's SIMD kernel
// rounds them (lane-order partial sums, separate multiply and add;
// mathop_avx.c:43-111, mathop_sse2.c:43-100, mathop.c:296-307).  Evaluating
// that exact distance for every (descriptor, centre) pair is VALU work of
// 3DK per descriptor.  Here the matrix cores first compute, for every pair,
//   a_ik = ||x_i||^2 - 2 x_i.c_k + ||c_k||^2
// with the dot product in split BF16 (x = x_hi + x_lo, three products
// x_hi c_hi + x_lo c_hi + x_hi c_lo, FP32 accumulation), whose distance to the
// exact float distance e_ik is bounded by eps_i = 2^-13 (||x_i|| + max_k ||c_k||)^2
// (a margin of about 4x over the sum of: the split's representation and
// dropped-product errors 3 * 2^-18 |x_d c_d| per term, the FP32 accumulation
// of the 3 * ceil(D/32) matrix instructions, the roundings of a_ik, and the
// reference's own float rounding of e_ik).  Every centre whose exact distance
// is the minimum then has a_ik <= min_k a_ik + 2 eps_i, so the exact distance
// is evaluated only for those candidates (usually one or two) and the
// smallest, lowest index first, is the reference's assignment.  A row with no
// candidate (non-finite data) or more than QM_CAP of them is scanned exactly
// over all centres.
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

template <int NQ, int LANES>
__global__ __launch_bounds__(QM_WAVES * 64) void qm_quantize_kernel(
    const float* __restrict__ x, const float* __restrict__ centers, long long N, int D, int K, int Kp,
    const __bf16* __restrict__ chi, const __bf16* __restrict__ clo, const float* __restrict__ cc,
    const unsigned* __restrict__ maxcBits, uint32_t* __restrict__ assign, float* __restrict__ dist) {
  constexpr int Dq = NQ * 32;
  constexpr int WS = Dq + 8;                       // LDS row stride (halves) of a staged centre
  constexpr int NT = QM_WAVES * 64;
  constexpr int PARTW = QM_BLOCK * Dq / 8;         // 16-byte words of one part (hi or lo) of a block
  constexpr int STG = 2 * PARTW / NT;              // words per thread per block
  static_assert(STG * NT == 2 * PARTW, "staging must divide evenly");
  __shared__ __align__(16) __bf16 wsh[2][2][QM_BLOCK][WS];  // [buffer][hi, lo][centre][feature]
  __shared__ float xxs[QM_WAVES][32];
  __shared__ float epss[QM_WAVES][32];
  __shared__ int cnt[QM_WAVES][32];
  __shared__ int cand[QM_WAVES][32][QM_CAP];

  const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
  const int lr = lane & 15, lg = lane >> 4;
  const long long row0 = ((long long)blockIdx.x * QM_WAVES + w) * 32;

  // A fragments: row 16t + lr, features 32j + 8lg .. +7, split hi / lo; the
  // squared norm of the row in FP64 from the same coordinates
  bf16x8 ahi[2][NQ], alo[2][NQ];
  double nrm[2] = {0.0, 0.0};
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const long long row = row0 + 16 * t + lr;
#pragma unroll
    for (int j = 0; j < NQ; ++j) {
#pragma unroll
      for (int v = 0; v < 8; ++v) {
        const int d = 32 * j + 8 * lg + v;
        const float xv = (row < N && d < D) ? x[row * D + d] : 0.f;
        const __bf16 h = (__bf16)xv;
        ahi[t][j][v] = h;
        alo[t][j][v] = (__bf16)(xv - (float)h);
        nrm[t] += (double)xv * (double)xv;
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
      xxs[w][16 * t + lr] = (float)nrm[t];
      epss[w][16 * t + lr] = (float)(ldexp(m * m, -13) * (1.0 + 1e-6));
    }
  }
  if (lane < 32) cnt[w][lane] = 0;
  __syncthreads();
  // accumulator layout: rows 16t + 4lg + r, centre 16c + lr of the block
  float xx[2][4], ep[2][4];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) {
      xx[t][r] = xxs[w][16 * t + 4 * lg + r];
      ep[t][r] = epss[w][16 * t + 4 * lg + r];
    }

  float thr[2][4];
#pragma unroll
  for (int t = 0; t < 2; ++t)
#pragma unroll
    for (int r = 0; r < 4; ++r) thr[t][r] = INFINITY;

  // centre blocks double-buffered in LDS, the next one fetched into
  // registers while the current one is multiplied (one barrier per block);
  // steps 0 .. nb-1 are sweep 0 (smallest screened distance per row), steps
  // nb .. 2nb-1 sweep 1 (the candidates)
  const int nb = Kp / QM_BLOCK, steps = 2 * nb;
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
#define QM_STAGE_LOAD(BLK)                                                              \
  {                                                                                     \
    const int blk_ = (BLK);                                                             \
    _Pragma("unroll") for (int q = 0; q < STG; ++q) {                                   \
      const int o = srcOff[q] + blk_ * QM_BLOCK * Dq;                                   \
      const __bf16* src = o < Kp * Dq ? chi + o : clo + (o - Kp * Dq);                  \
      pre[q] = *reinterpret_cast<const u32x4*>(src);                                    \
    }                                                                                   \
  }
#define QM_STAGE_STORE(BUF)                                                             \
  {                                                                                     \
    __bf16* dst_ = &wsh[(BUF)][0][0][0];                                                \
    _Pragma("unroll") for (int q = 0; q < STG; ++q)                                     \
      *reinterpret_cast<u32x4*>(dst_ + dstOff[q]) = pre[q];                             \
  }
  QM_STAGE_LOAD(0)
  QM_STAGE_STORE(0)
  if (steps > 1) QM_STAGE_LOAD(1 % nb)
  __syncthreads();
  for (int st = 0; st < steps; ++st) {
    const int sweep = st >= nb ? 1 : 0;
    const int k0 = (st - sweep * nb) * QM_BLOCK;
    if (st + 1 < steps) {
      QM_STAGE_STORE((st + 1) & 1)  // that buffer was last read in step st - 1, before the barrier
      if (st + 2 < steps) QM_STAGE_LOAD((st + 2) % nb)
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
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int k = k0 + 16 * c + lr;
        const float ck = cc[k];  // +inf for padding centres
#pragma unroll
        for (int t = 0; t < 2; ++t)
#pragma unroll
          for (int r = 0; r < 4; ++r) {
            const float a = (xx[t][r] - 2.f * acc[c][t][r]) + ck;
            if (sweep == 0) {
              thr[t][r] = fminf(thr[t][r], a);
            } else if (k < K && a <= thr[t][r]) {
              const int rw = 16 * t + 4 * lg + r;
              const int slot = atomicAdd(&cnt[w][rw], 1);
              if (slot < QM_CAP) cand[w][rw][slot] = k;
            }
          }
      }
    }
    if (st == nb - 1) {
      // row minimum over the 16 lanes holding the row's centres, plus 2 eps
#pragma unroll
      for (int t = 0; t < 2; ++t)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          float m = thr[t][r];
#pragma unroll
          for (int o = 1; o < 16; o <<= 1) m = fminf(m, __shfl_xor(m, o, 64));
          thr[t][r] = m + 2.f * ep[t][r];
        }
    }
    __syncthreads();
  }
#undef QM_STAGE_LOAD
#undef QM_STAGE_STORE

  // exact distances: lanes l and l + 32 serve row l & 31
  const int rw = lane & 31, half = lane >> 5;
  const long long row = row0 + rw;
  float bv = INFINITY;
  int bk = 0x7fffffff;
  if (row < N) {
    const float* xr = x + row * D;
    const int n = cnt[w][rw];
    if (n >= 1 && n <= QM_CAP) {
      for (int s = half; s < n; s += 2) {
        const int k = cand[w][rw][s];
        const float e = exact_l2<LANES>(xr, centers + (size_t)k * D, D);
        if (better(e, k, bv, bk)) { bv = e; bk = k; }
      }
    } else {
      for (int k = half; k < K; k += 2) {
        const float e = exact_l2<LANES>(xr, centers + (size_t)k * D, D);
        if (better(e, k, bv, bk)) { bv = e; bk = k; }
      }
    }
  }
  const float ov = __shfl_xor(bv, 32, 64);
  const int ok = __shfl_xor(bk, 32, 64);
  if (better(ov, ok, bv, bk)) { bv = ov; bk = ok; }
  if (half == 0 && row < N) {
    // no finite distance at all: the first centre, as the exhaustive kernel
    assign[row] = (uint32_t)(bk == 0x7fffffff ? 0 : bk);
    if (dist) dist[row] = bv;
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
#define QM_LAUNCH(NQV, LV)                                                                                        \
  VLHIP_LAUNCH("kmeans_quantize_mfma", (qm_quantize_kernel<NQV, LV>), dim3(blocks), dim3(QM_WAVES * 64), 0, x,      \
               centers, (long long)N, (int)D, (int)K, Kp, chi, clo, cc, maxc, assign, dist)
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
