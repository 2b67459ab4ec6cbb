// kmeans_kernels.hip -- exact L2 quantizer (vl/kmeans.c:655-702), the
// point-to-data distances used by k-means++ seeding (kmeans.c:610-616), and
// the center update and ANN keep-previous rule of the refinement loops.
//
// Bit-exact argmin parity: the reference evaluates each squared L2 distance
// with the host's SIMD kernel -- LANES partial sums over d = l (mod LANES) for
// the first floor(D/LANES)*LANES coordinates, a fixed horizontal-add tree, then
// the remaining coordinates one by one (mathop_avx.c:43-111,
// mathop_sse2.c:43-100, mathop.c:296-307) -- with separate multiply and add
// (no fused multiply-add).  The kernels below reproduce that order with FP
// contraction disabled, so every distance is bitwise the reference's and the
// strict '<' / first-index rule (kmeans.c:690-695) yields the same assignment.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

#include <algorithm>
#include <type_traits>

namespace vlhip {

// Horizontal sum of the lane accumulators in the order of the host kernel.
//  8 lanes (AVX float):   ((a0+a1)+(a4+a5)) + ((a2+a3)+(a6+a7))
//  4 lanes (AVX double, SSE2 float): (a0+a2) + (a1+a3)
//  2 lanes (SSE2 double): a0 + a1
template <typename T, int LANES>
__device__ __forceinline__ T lane_hsum(const T (&a)[LANES]) {
#pragma clang fp contract(off)
  if constexpr (LANES == 8) {
    return ((a[0] + a[1]) + (a[4] + a[5])) + ((a[2] + a[3]) + (a[6] + a[7]));
  } else if constexpr (LANES == 4) {
    return (a[0] + a[2]) + (a[1] + a[3]);
  } else if constexpr (LANES == 2) {
    return a[0] + a[1];
  } else {
    return a[0];
  }
}

constexpr int QZ_ROWS = 32;
constexpr int QZ_COLS = 32;
constexpr int QZ_DC = 16;  // multiple of every lane count

// Thread (ty, tx) of a 16x16 grid owns rows ty*2.., centers tx*2...
template <typename T, int LANES>
__global__ __launch_bounds__(256) void quantize_kernel(const T* __restrict__ x, const T* __restrict__ centers,
                                                       long long N, int D, int K, uint32_t* __restrict__ assign,
                                                       T* __restrict__ dist) {
#pragma clang fp contract(off)
  __shared__ T xs[QZ_DC][QZ_ROWS + 1];
  __shared__ T cs[QZ_DC][QZ_COLS + 1];
  __shared__ T red_v[QZ_ROWS][16];
  __shared__ int red_i[QZ_ROWS][16];

  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const long long row0 = (long long)blockIdx.x * QZ_ROWS;
  const int Dv = (LANES > 1) ? (D / LANES) * LANES : 0;  // vectorised prefix

  T best[2];
  int bidx[2];
  best[0] = best[1] = (T)INFINITY;
  bidx[0] = bidx[1] = 0;

  for (int k0 = 0; k0 < K; k0 += QZ_COLS) {
    T acc[2][2][LANES];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j)
#pragma unroll
        for (int l = 0; l < LANES; ++l) acc[i][j][l] = T(0);
    for (int d0 = 0; d0 < Dv; d0 += QZ_DC) {
      const int dc = min(QZ_DC, Dv - d0);
      __syncthreads();
      for (int idx = tid; idx < QZ_ROWS * QZ_DC; idx += 256) {
        const int r = idx / QZ_DC, dd = idx % QZ_DC;
        const long long row = row0 + r;
        xs[dd][r] = (dd < dc && row < N) ? x[row * D + d0 + dd] : T(0);
      }
      for (int idx = tid; idx < QZ_COLS * QZ_DC; idx += 256) {
        const int c = idx / QZ_DC, dd = idx % QZ_DC;
        const int k = k0 + c;
        cs[dd][c] = (dd < dc && k < K) ? centers[(size_t)k * D + d0 + dd] : T(0);
      }
      __syncthreads();
      for (int dd = 0; dd < dc; dd += LANES) {
#pragma unroll
        for (int l = 0; l < LANES; ++l) {
#pragma unroll
          for (int i = 0; i < 2; ++i)
#pragma unroll
            for (int j = 0; j < 2; ++j) {
              const T delta = xs[dd + l][ty * 2 + i] - cs[dd + l][tx * 2 + j];
              const T d2 = delta * delta;
              acc[i][j][l] = acc[i][j][l] + d2;
            }
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 2; ++i) {
      const long long row = row0 + ty * 2 + i;
#pragma unroll
      for (int j = 0; j < 2; ++j) {
        const int k = k0 + tx * 2 + j;
        T s = (LANES > 1) ? lane_hsum<T, LANES>(acc[i][j]) : T(0);
        if (row < N && k < K) {
          for (int d = Dv; d < D; ++d) {
            const T delta = x[row * D + d] - centers[(size_t)k * D + d];
            const T d2 = delta * delta;
            s = s + d2;
          }
          if (s < best[i]) { best[i] = s; bidx[i] = k; }
        }
      }
    }
  }

#pragma unroll
  for (int i = 0; i < 2; ++i) {
    red_v[ty * 2 + i][tx] = best[i];
    red_i[ty * 2 + i][tx] = bidx[i];
  }
  __syncthreads();
  if (tid < QZ_ROWS) {
    const long long row = row0 + tid;
    if (row < N) {
      // thread tx visited centers in increasing order within itself; merge
      // with the smallest distance, ties to the smallest index.
      T bv = red_v[tid][0];
      int bi = red_i[tid][0];
      for (int t = 1; t < 16; ++t) {
        const T v = red_v[tid][t];
        const int a = red_i[tid][t];
        if (v < bv || (v == bv && a < bi)) { bv = v; bi = a; }
      }
      assign[row] = (uint32_t)bi;
      if (dist) dist[row] = bv;
    }
  }
}

// Squared L2 distance of each datum to `point`, host lane order.
template <typename T, int LANES>
__global__ __launch_bounds__(256) void l2_to_point_kernel(const T* __restrict__ x, const T* __restrict__ point,
                                                          long long N, int D, T* __restrict__ out) {
#pragma clang fp contract(off)
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= N) return;
  const T* xi = x + i * D;
  const int Dv = (LANES > 1) ? (D / LANES) * LANES : 0;
  T acc[LANES];
#pragma unroll
  for (int l = 0; l < LANES; ++l) acc[l] = T(0);
  for (int d = 0; d < Dv; d += LANES) {
#pragma unroll
    for (int l = 0; l < LANES; ++l) {
      const T delta = point[d + l] - xi[d + l];
      const T d2 = delta * delta;
      acc[l] = acc[l] + d2;
    }
  }
  T s = (LANES > 1) ? lane_hsum<T, LANES>(acc) : T(0);
  for (int d = Dv; d < D; ++d) {
    const T delta = point[d] - xi[d];
    const T d2 = delta * delta;
    s = s + d2;
  }
  out[i] = s;
}

// Center update of Lloyd / ANN refinement (kmeans.c:899-931 and 1075-1107):
// the reference zeroes the centers and adds every datum to its center in data
// order, then divides by the cluster mass, or copies a random datum into an
// empty center.  Here members[starts[c] .. starts[c+1]) lists the data of
// cluster c in increasing data order, so each (c, d) sum below is the same
// sequence of float (or double) additions as the reference's and the result
// is bitwise equal; restartRow[c] is the datum drawn for an empty cluster.
// Block (c, dimension slice of 64); the member indices are uniform per block,
// eight data loads are issued before their adds.
template <typename T>
__global__ __launch_bounds__(64) void center_update_kernel(const T* __restrict__ x,
                                                           const uint32_t* __restrict__ members,
                                                           const uint64_t* __restrict__ starts,
                                                           const uint64_t* __restrict__ restartRow, int D,
                                                           T* __restrict__ centers) {
#pragma clang fp contract(off)
  const int c = blockIdx.x;
  const int d = blockIdx.y * 64 + threadIdx.x;
  if (d >= D) return;
  const uint64_t b = starts[c], e = starts[c + 1];
  T out;
  if (e > b) {
    T s = T(0);
    uint64_t j = b;
    for (; j + 8 <= e; j += 8) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) v[u] = x[(size_t)members[j + u] * D + d];
#pragma unroll
      for (int u = 0; u < 8; ++u) s = s + v[u];
    }
    for (; j < e; ++j) s = s + x[(size_t)members[j] * D + d];
    out = s / (T)(e - b);
  } else {
    out = x[(size_t)restartRow[c] * D + d];
  }
  centers[(size_t)c * D + d] = out;
}

// ANN assignment update (kmeans.c:750-763 with update): a datum keeps its
// previous center, at the distance to that center, unless the new nearest
// center is strictly closer.  Distances in the host lane order.
template <typename T, int LANES>
__global__ __launch_bounds__(256) void ann_merge_kernel(const T* __restrict__ x, const T* __restrict__ centers,
                                                        long long N, int D, const uint32_t* __restrict__ prev,
                                                        uint32_t* __restrict__ assign, T* __restrict__ dist) {
#pragma clang fp contract(off)
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= N) return;
  const uint32_t a = prev[i];
  const T* xi = x + i * D;
  const T* ci = centers + (size_t)a * D;
  const int Dv = (LANES > 1) ? (D / LANES) * LANES : 0;
  T acc[LANES];
#pragma unroll
  for (int l = 0; l < LANES; ++l) acc[l] = T(0);
  for (int d = 0; d < Dv; d += LANES) {
#pragma unroll
    for (int l = 0; l < LANES; ++l) {
      const T delta = xi[d + l] - ci[d + l];
      const T d2 = delta * delta;
      acc[l] = acc[l] + d2;
    }
  }
  T s = (LANES > 1) ? lane_hsum<T, LANES>(acc) : T(0);
  for (int d = Dv; d < D; ++d) {
    const T delta = xi[d] - ci[d];
    const T d2 = delta * delta;
    s = s + d2;
  }
  if (!(s > dist[i])) {
    assign[i] = a;
    dist[i] = s;
  }
}

template <typename T>
void quantize_t(uint32_t* assign, T* dist, const T* x, const T* centers, size_t N, size_t D, size_t K, int lanes,
                void* ws) {
  VLHIP_REQUIRE(K >= 1 && D >= 1, "numCenters and dimension must be positive");
  if (N == 0) return;
  if constexpr (std::is_same<T, float>::value) {
    if ((lanes == 1 || lanes == 2 || lanes == 4 || lanes == 8) &&
        quantize_screened(assign, dist, x, centers, N, D, K, lanes, ws))
      return;
  }
  const unsigned blocks = (unsigned)((N + QZ_ROWS - 1) / QZ_ROWS);
  switch (lanes) {
    case 8: VLHIP_LAUNCH("kmeans_quantize", (quantize_kernel<T, 8>), dim3(blocks), dim3(256), 0, x, centers,
                         (long long)N, (int)D, (int)K, assign, dist); break;
    case 4: VLHIP_LAUNCH("kmeans_quantize", (quantize_kernel<T, 4>), dim3(blocks), dim3(256), 0, x, centers,
                         (long long)N, (int)D, (int)K, assign, dist); break;
    case 2: VLHIP_LAUNCH("kmeans_quantize", (quantize_kernel<T, 2>), dim3(blocks), dim3(256), 0, x, centers,
                         (long long)N, (int)D, (int)K, assign, dist); break;
    case 1: VLHIP_LAUNCH("kmeans_quantize", (quantize_kernel<T, 1>), dim3(blocks), dim3(256), 0, x, centers,
                         (long long)N, (int)D, (int)K, assign, dist); break;
    default: fatal("vlhip_kmeans_quantize", __FILE__, __LINE__, "lanes must be 1, 2, 4 or 8");
  }
}

template <typename T>
void l2_to_point_t(T* out, const T* x, const T* point, size_t N, size_t D, int lanes) {
  if (N == 0) return;
  const unsigned blocks = (unsigned)((N + 255) / 256);
  switch (lanes) {
    case 8: VLHIP_LAUNCH("l2_to_point", (l2_to_point_kernel<T, 8>), dim3(blocks), dim3(256), 0, x, point,
                         (long long)N, (int)D, out); break;
    case 4: VLHIP_LAUNCH("l2_to_point", (l2_to_point_kernel<T, 4>), dim3(blocks), dim3(256), 0, x, point,
                         (long long)N, (int)D, out); break;
    case 2: VLHIP_LAUNCH("l2_to_point", (l2_to_point_kernel<T, 2>), dim3(blocks), dim3(256), 0, x, point,
                         (long long)N, (int)D, out); break;
    case 1: VLHIP_LAUNCH("l2_to_point", (l2_to_point_kernel<T, 1>), dim3(blocks), dim3(256), 0, x, point,
                         (long long)N, (int)D, out); break;
    default: fatal("vlhip_l2_to_point", __FILE__, __LINE__, "lanes must be 1, 2, 4 or 8");
  }
}

template void quantize_t<float>(uint32_t*, float*, const float*, const float*, size_t, size_t, size_t, int, void*);
template void quantize_t<double>(uint32_t*, double*, const double*, const double*, size_t, size_t, size_t, int,
                                 void*);

template <typename T>
void ann_merge_t(uint32_t* assign, T* dist, const uint32_t* prev, const T* x, const T* centers, size_t N,
                        size_t D, int lanes) {
  if (N == 0) return;
  const unsigned blocks = (unsigned)((N + 255) / 256);
  switch (lanes) {
    case 8: VLHIP_LAUNCH("kmeans_ann_merge", (ann_merge_kernel<T, 8>), dim3(blocks), dim3(256), 0, x, centers,
                         (long long)N, (int)D, prev, assign, dist); break;
    case 4: VLHIP_LAUNCH("kmeans_ann_merge", (ann_merge_kernel<T, 4>), dim3(blocks), dim3(256), 0, x, centers,
                         (long long)N, (int)D, prev, assign, dist); break;
    case 2: VLHIP_LAUNCH("kmeans_ann_merge", (ann_merge_kernel<T, 2>), dim3(blocks), dim3(256), 0, x, centers,
                         (long long)N, (int)D, prev, assign, dist); break;
    case 1: VLHIP_LAUNCH("kmeans_ann_merge", (ann_merge_kernel<T, 1>), dim3(blocks), dim3(256), 0, x, centers,
                         (long long)N, (int)D, prev, assign, dist); break;
    default: fatal("vlhip_kmeans_ann_merge", __FILE__, __LINE__, "lanes must be 1, 2, 4 or 8");
  }
}

}  // namespace vlhip

using namespace vlhip;

extern "C" {

void vlhip_kmeans_quantize(int dtype, uint32_t* assign, void* dist, void const* data, void const* centers,
                           size_t N, size_t D, size_t K, int lanes) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (dtype == 1)
    quantize_t<float>(assign, (float*)dist, (const float*)data, (const float*)centers, N, D, K, lanes);
  else if (dtype == 2)
    quantize_t<double>(assign, (double*)dist, (const double*)data, (const double*)centers, N, D, K, lanes);
  else
    fatal("vlhip_kmeans_quantize", __FILE__, __LINE__, "unknown data type");
}

void vlhip_l2_to_point(int dtype, void* out, void const* data, void const* point, size_t N, size_t D, int lanes) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (dtype == 1)
    l2_to_point_t<float>((float*)out, (const float*)data, (const float*)point, N, D, lanes);
  else if (dtype == 2)
    l2_to_point_t<double>((double*)out, (const double*)data, (const double*)point, N, D, lanes);
  else
    fatal("vlhip_l2_to_point", __FILE__, __LINE__, "unknown data type");
}

void vlhip_kmeans_update_centers(int dtype, void* centers, void const* data, uint32_t const* members,
                                 uint64_t const* starts, uint64_t const* restartRow, size_t numCenters,
                                 size_t dimension) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  VLHIP_REQUIRE(dimension >= 1 && dimension < (1u << 30), "dimension out of range");
  VLHIP_REQUIRE(numCenters < (1u << 31), "numCenters out of range");
  if (numCenters == 0) return;
  const dim3 grid((unsigned)numCenters, (unsigned)((dimension + 63) / 64));
  if (dtype == 1)
    VLHIP_LAUNCH("kmeans_update", (center_update_kernel<float>), grid, dim3(64), 0, (const float*)data, members,
                 starts, restartRow, (int)dimension, (float*)centers);
  else if (dtype == 2)
    VLHIP_LAUNCH("kmeans_update", (center_update_kernel<double>), grid, dim3(64), 0, (const double*)data, members,
                 starts, restartRow, (int)dimension, (double*)centers);
  else
    fatal("vlhip_kmeans_update_centers", __FILE__, __LINE__, "unknown data type");
}

void vlhip_kmeans_ann_merge(int dtype, uint32_t* assignments, void* distances, uint32_t const* previous,
                            void const* data, void const* centers, size_t numData, size_t dimension, int lanes) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (dtype == 1)
    ann_merge_t<float>(assignments, (float*)distances, previous, (const float*)data, (const float*)centers, numData,
                       dimension, lanes);
  else if (dtype == 2)
    ann_merge_t<double>(assignments, (double*)distances, previous, (const double*)data, (const double*)centers,
                        numData, dimension, lanes);
  else
    fatal("vlhip_kmeans_ann_merge", __FILE__, __LINE__, "unknown data type");
}

}  // extern "C"
