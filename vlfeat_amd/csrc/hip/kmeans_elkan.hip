// kmeans_elkan.hip -- Elkan's k-means refinement (vl/kmeans.c:1162-1745) on
// the GPU, bit for bit.
//
// The reference keeps, per datum x, an upper bound u(x) on its distance to its
// center (with a flag telling whether the bound is the exact distance) and a
// lower bound l(x, c) for every center c; all distances are squared L2 in the
// host's SIMD lane order.  Every per-datum step of an iteration depends only on
// that datum's bounds and on per-center tables the host prepares (center-center
// distances, each center's distance to its nearest other center, each center's
// move), so the steps run as one thread per datum with the reference's
// operations in the reference's order:
//   init      assign to center 0, then scan c = 1 .. K-1, skipping c when
//             4 u <= d(c_a, c) (kmeans.c:1224-1256);
//   bounds    u <- float(double(u + m) + 2 sqrt(u m)) with m the assigned
//             center's move, flag cleared; l(x, c) <- 0 if l < m_c, else
//             float(double(l + m_c) - 2 sqrt(l m_c)) (kmeans.c:1400-1450;
//             the float sum rounds first, then the double expression);
//   assign    skip the datum when 4 u <= its center's nearest-center
//             distance; otherwise scan c in order, skipping c when 4 u <=
//             d(c_a, c) or u <= l(x, c), refreshing u once with the exact
//             distance when it is not, and moving to c on a strictly smaller
//             distance (kmeans.c:1487-1575).
// The square root is sqrtf / sqrt correctly rounded (the reference's libm);
// no contraction.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip.h"

#include <type_traits>

namespace vlhip {

#pragma clang fp contract(off)

// Squared L2 distance in the host's lane order: LANES partial sums over d = l
// (mod LANES), the host kernel's horizontal-add tree, then the remaining
// coordinates one by one.
template <typename T, int LANES>
__device__ __forceinline__ T elkan_l2(const T* __restrict__ a, const T* __restrict__ b, int D) {
  const int Dv = LANES > 1 ? (D / LANES) * LANES : 0;
  T acc[LANES];
#pragma unroll
  for (int l = 0; l < LANES; ++l) acc[l] = T(0);
  for (int d = 0; d < Dv; d += LANES) {
#pragma unroll
    for (int l = 0; l < LANES; ++l) {
      const T delta = a[d + l] - b[d + l];
      const T d2 = delta * delta;
      acc[l] = acc[l] + d2;
    }
  }
  T s;
  if constexpr (LANES == 8) s = ((acc[0] + acc[1]) + (acc[4] + acc[5])) + ((acc[2] + acc[3]) + (acc[6] + acc[7]));
  else if constexpr (LANES == 4) s = (acc[0] + acc[2]) + (acc[1] + acc[3]);
  else if constexpr (LANES == 2) s = acc[0] + acc[1];
  else s = T(0);
  for (int d = Dv; d < D; ++d) {
    const T delta = a[d] - b[d];
    const T d2 = delta * delta;
    s = s + d2;
  }
  return s;
}

__device__ __forceinline__ float elkan_sqrt(float v) { return __fsqrt_rn(v); }
__device__ __forceinline__ double elkan_sqrt(double v) { return __dsqrt_rn(v); }

template <typename T, int LANES>
__global__ __launch_bounds__(256) void elkan_init_kernel(const T* __restrict__ x, const T* __restrict__ centers,
                                                         const T* __restrict__ cd, long long N, int D, int K,
                                                         uint32_t* __restrict__ assign, T* __restrict__ ub,
                                                         unsigned char* __restrict__ strict, T* __restrict__ lb) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= N) return;
  const T* xi = x + i * D;
  T* L = lb + i * K;
  for (int c = 0; c < K; ++c) L[c] = T(0);
  uint32_t a = 0;
  T u = elkan_l2<T, LANES>(xi, centers, D);
  L[0] = u;
  for (int c = 1; c < K; ++c) {
    if (4.0 * (double)u <= (double)cd[(size_t)c + (size_t)a * K]) continue;
    const T dist = elkan_l2<T, LANES>(xi, centers + (size_t)c * D, D);
    L[c] = dist;
    if (dist < u) {
      u = dist;
      a = (uint32_t)c;
    }
  }
  assign[i] = a;
  ub[i] = u;
  strict[i] = 1;
}

// One thread per (datum, center) for the lower bounds; the threads of center
// 0 also move the datum's upper bound.
template <typename T>
__global__ __launch_bounds__(256) void elkan_bounds_kernel(long long N, int K, const uint32_t* __restrict__ assign,
                                                           const T* __restrict__ move, T* __restrict__ ub,
                                                           unsigned char* __restrict__ strict, T* __restrict__ lb) {
  const long long t = blockIdx.x * 256LL + threadIdx.x;
  if (t >= N * (long long)K) return;
  const long long i = t / K;
  const int c = (int)(t - i * K);
  {
    const T a = lb[t], b = move[c];
    if (a < b) {
      lb[t] = T(0);
    } else {
      const T s = elkan_sqrt(a * b);
      lb[t] = (T)((double)(a + b) - 2.0 * (double)s);
    }
  }
  if (c == 0) {
    const T a = ub[i], b = move[assign[i]];
    const T s = elkan_sqrt(a * b);
    ub[i] = (T)((double)(a + b) + 2.0 * (double)s);
    strict[i] = 0;
  }
}

template <typename T, int LANES>
__global__ __launch_bounds__(256) void elkan_assign_kernel(const T* __restrict__ x, const T* __restrict__ centers,
                                                           const T* __restrict__ cd, const T* __restrict__ nearest,
                                                           long long N, int D, int K, uint32_t* __restrict__ assign,
                                                           T* __restrict__ ub, unsigned char* __restrict__ strict,
                                                           T* __restrict__ lb, unsigned* __restrict__ moved) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= N) return;
  uint32_t cx = assign[i];
  T u = ub[i];
  if (4.0 * (double)u <= (double)nearest[cx]) return;
  bool st = strict[i] != 0, any = false;
  const T* xi = x + i * D;
  T* L = lb + i * K;
  for (int c = 0; c < K; ++c) {
    if ((uint32_t)c == cx) continue;
    if (4.0 * (double)u <= (double)cd[(size_t)c + (size_t)cx * K]) continue;
    if (u <= L[c]) continue;
    if (!st) {
      const T dist = elkan_l2<T, LANES>(xi, centers + (size_t)cx * D, D);
      u = dist;
      st = true;
      L[cx] = dist;
      if (4.0 * (double)u <= (double)cd[(size_t)c + (size_t)cx * K]) continue;
      if (u <= L[c]) continue;
    }
    const T dist = elkan_l2<T, LANES>(xi, centers + (size_t)c * D, D);
    L[c] = dist;
    if (dist < u) {
      cx = (uint32_t)c;
      u = dist;
      any = true;
    }
  }
  assign[i] = cx;
  ub[i] = u;
  strict[i] = st ? 1 : 0;
  if (any) atomicOr(moved, 1u);
}

template <typename T, int LANES>
__global__ __launch_bounds__(256) void elkan_energy_kernel(const T* __restrict__ x, const T* __restrict__ centers,
                                                           long long N, int D, const uint32_t* __restrict__ assign,
                                                           T* __restrict__ dist) {
  const long long i = blockIdx.x * 256LL + threadIdx.x;
  if (i >= N) return;
  dist[i] = elkan_l2<T, LANES>(x + i * D, centers + (size_t)assign[i] * D, D);
}

#pragma clang fp contract(on)

template <typename T>
static void elkan_step(int step, const T* x, const T* centers, const T* cd, const T* aux, size_t N, size_t D,
                       size_t K, int lanes, uint32_t* assign, T* ub, unsigned char* strict, T* lb, unsigned* moved,
                       T* dist) {
  const unsigned blocks = (unsigned)((N + 255) / 256);
#define ELKAN_LANES(L)                                                                                              \
  case L:                                                                                                           \
    if (step == 0)                                                                                                  \
      VLHIP_LAUNCH("elkan_init", (elkan_init_kernel<T, L>), dim3(blocks), dim3(256), 0, x, centers, cd,             \
                   (long long)N, (int)D, (int)K, assign, ub, strict, lb);                                           \
    else if (step == 2)                                                                                             \
      VLHIP_LAUNCH("elkan_assign", (elkan_assign_kernel<T, L>), dim3(blocks), dim3(256), 0, x, centers, cd, aux,    \
                   (long long)N, (int)D, (int)K, assign, ub, strict, lb, moved);                                    \
    else                                                                                                            \
      VLHIP_LAUNCH("elkan_energy", (elkan_energy_kernel<T, L>), dim3(blocks), dim3(256), 0, x, centers,             \
                   (long long)N, (int)D, assign, dist);                                                             \
    break;
  switch (lanes) {
    ELKAN_LANES(8) ELKAN_LANES(4) ELKAN_LANES(2) ELKAN_LANES(1)
    default: fatal("vlhip_kmeans_elkan", __FILE__, __LINE__, "lanes must be 1, 2, 4 or 8");
  }
#undef ELKAN_LANES
}

template <typename T>
static void elkan_bounds(const T* move, size_t N, size_t K, const uint32_t* assign, T* ub, unsigned char* strict,
                         T* lb) {
  const size_t total = N * K;
  VLHIP_REQUIRE((total + 255) / 256 < ((size_t)1 << 31), "Elkan bounds: too many data x centers");
  VLHIP_LAUNCH("elkan_bounds", elkan_bounds_kernel<T>, dim3((unsigned)((total + 255) / 256)), dim3(256), 0,
               (long long)N, (int)K, assign, move, ub, strict, lb);
}

}  // namespace vlhip

using namespace vlhip;

extern "C" void vlhip_kmeans_elkan(int dtype, int step, void const* data, void const* centers, void const* table,
                                   void const* aux, size_t numData, size_t dimension, size_t numCenters, int lanes,
                                   uint32_t* assignments, void* upperBounds, unsigned char* upperBoundIsExact,
                                   void* lowerBounds, unsigned* moved, void* distances) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  VLHIP_REQUIRE(dtype == 1 || dtype == 2, "vlhip_kmeans_elkan: unknown data type");
  VLHIP_REQUIRE(step >= 0 && step <= 3, "vlhip_kmeans_elkan: unknown step");
  VLHIP_REQUIRE(numCenters >= 1 && numCenters < ((size_t)1 << 31) && dimension >= 1 && dimension < ((size_t)1 << 31),
                "vlhip_kmeans_elkan: sizes out of range");
  if (numData == 0) return;
  VLHIP_REQUIRE((numData + 255) / 256 < ((size_t)1 << 31), "vlhip_kmeans_elkan: too many data");
  if (dtype == 1) {
    if (step == 1)
      elkan_bounds<float>((const float*)aux, numData, numCenters, assignments, (float*)upperBounds,
                          upperBoundIsExact, (float*)lowerBounds);
    else
      elkan_step<float>(step, (const float*)data, (const float*)centers, (const float*)table, (const float*)aux,
                        numData, dimension, numCenters, lanes, assignments, (float*)upperBounds, upperBoundIsExact,
                        (float*)lowerBounds, moved, (float*)distances);
  } else {
    if (step == 1)
      elkan_bounds<double>((const double*)aux, numData, numCenters, assignments, (double*)upperBounds,
                           upperBoundIsExact, (double*)lowerBounds);
    else
      elkan_step<double>(step, (const double*)data, (const double*)centers, (const double*)table, (const double*)aux,
                         numData, dimension, numCenters, lanes, assignments, (double*)upperBounds, upperBoundIsExact,
                         (double*)lowerBounds, moved, (double*)distances);
  }
}
