// vlad_kernels.hip -- VLAD encoder (vl/vlad.c:163-242).
//
// Workgroup (image, block of VL_COMPS clusters) accumulates, for each of its
// clusters k and every dimension d, enc[k,d] = sum_i q_ik x_id over the image's
// descriptors in index order, with the reference's arithmetic: the product and
// sum are formed in double and the running sum is rounded back to the data
// type after every term (vlad.c:181-188, where q is double and enc is TYPE),
// and only q > 0 contributes.  The per-cluster corrections, signed square root
// and per-cluster L2 normalisation follow vlad.c:191-227 in the same order and
// precision; the global normalisation (vlad.c:230-241) sums the squared
// entries sequentially in index order, so for float data the result matches
// the reference's float path element for element.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

#include <algorithm>

namespace vlhip {

constexpr int VL_COMPS = 16;
constexpr int VL_ROWS = 64;

template <typename T, int DPT, bool HARD>
__global__ __launch_bounds__(256) void vlad_accum_kernel(
    const T* __restrict__ x, const uint64_t* __restrict__ offsets, int numImages, int D, int K,
    const T* __restrict__ means, const T* __restrict__ assign, const uint32_t* __restrict__ hard,
    int flags, T* __restrict__ enc, int numKB) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* xs = reinterpret_cast<T*>(smem_raw);           // [VL_ROWS][D + 1]
  T* qs = xs + (size_t)VL_ROWS * (D + 1);            // [VL_ROWS][VL_COMPS]
  T* zs = qs + VL_ROWS * VL_COMPS;                   // [VL_COMPS][D] final values
  __shared__ double massSh[VL_COMPS];

  const int b = blockIdx.x;
  int img, kb;
  if (numImages >= 8) {
    const int xcd = b & 7, slot = b >> 3;
    img = (slot / numKB) * 8 + xcd;
    kb = slot % numKB;
  } else {
    img = b / numKB;
    kb = b % numKB;
  }
  if (img >= numImages) return;

  const int tid = threadIdx.x;
  const int c = tid >> 4, dg = tid & 15;
  const int k = kb * VL_COMPS + c;
  const long long rBegin = (long long)offsets[img], rEnd = (long long)offsets[img + 1];

  T acc[DPT];
#pragma unroll
  for (int j = 0; j < DPT; ++j) acc[j] = T(0);
  double mass = 0.0;

  for (long long r0 = rBegin; r0 < rEnd; r0 += VL_ROWS) {
    const int nr = (int)min((long long)VL_ROWS, rEnd - r0);
    __syncthreads();
    for (int idx = tid; idx < nr * D; idx += 256) {
      const int r = idx / D, d = idx - r * D;
      xs[r * (D + 1) + d] = x[(r0 + r) * D + d];
    }
    for (int idx = tid; idx < VL_ROWS * VL_COMPS; idx += 256) {
      const int r = idx / VL_COMPS, cc = idx - r * VL_COMPS;
      const int kk = kb * VL_COMPS + cc;
      T qv = T(0);
      if (r < nr && kk < K) {
        if (HARD) qv = (hard[r0 + r] == (uint32_t)kk) ? T(1) : T(0);
        else qv = assign[(r0 + r) * K + kk];
      }
      qs[idx] = qv;
    }
    __syncthreads();
    if (k < K) {
      for (int r = 0; r < nr; ++r) {
        const T qv = qs[r * VL_COMPS + c];
        if (qv > T(0)) {
          const double qd = (double)qv;
          mass += qd;
#pragma unroll
          for (int j = 0; j < DPT; ++j) {
            const int d = dg + 16 * j;
            if (d < D) acc[j] = (T)((double)acc[j] + qd * (double)xs[r * (D + 1) + d]);
          }
        }
      }
    }
  }

  // vlad.c:191-213: mean subtraction and signed square root
  if (k < K) {
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const int d = dg + 16 * j;
      if (d >= D) continue;
      T e = acc[j];
      const T mu = means[(size_t)k * D + d];
      if (mass > 0) {
        if (flags & 8) {
          e = (T)((double)e / mass);
          e = e - mu;
        } else {
          e = (T)((double)e - mass * (double)mu);
        }
      }
      if (flags & 2) e = signed_sqrt(e);
      zs[c * D + d] = e;
    }
  }
  __syncthreads();
  // vlad.c:215-227: per-cluster L2 normalisation, sequential sum over d
  if ((flags & 1) && dg == 0 && k < K) {
    T n = T(0);
    for (int d = 0; d < D; ++d) {
      const T z = zs[c * D + d];
      n = n + z * z;
    }
    n = dev_sqrt<T>(n);
    massSh[c] = ((double)n > 1e-12) ? (double)n : 1e-12;
  }
  __syncthreads();
  if (k < K) {
    T* e = enc + (size_t)img * K * D + (size_t)k * D;
    const T n = (flags & 1) ? (T)massSh[c] : T(1);
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const int d = dg + 16 * j;
      if (d >= D) continue;
      const T z = zs[c * D + d];
      e[d] = (flags & 1) ? z / n : z;
    }
  }
}

// vlad.c:230-241: enc /= max(||enc||, 1e-12), the squared norm summed in
// index order by one lane (bitwise the reference's sequential loop).  The
// workgroup stages the squares z*z (each rounded as the reference rounds it)
// through LDS in chunks, so that the lane's only work is the chain of
// dependent additions, fed by 16-byte LDS reads.
constexpr int VL_NCHUNK = 4096;
template <typename T>
__global__ __launch_bounds__(256) void vlad_normalize_kernel(T* __restrict__ enc, size_t len) {
#pragma clang fp contract(off)
  __shared__ T nsh;
  __shared__ __align__(16) T buf[VL_NCHUNK];
  T* e = enc + (size_t)blockIdx.x * len;
  T n = T(0);
  for (size_t c0 = 0; c0 < len; c0 += VL_NCHUNK) {
    const int m = (int)min((size_t)VL_NCHUNK, len - c0);
    __syncthreads();
    for (int i = threadIdx.x; i < m; i += blockDim.x) {
      const T z = e[c0 + i];
      buf[i] = z * z;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
      constexpr int V = 16 / sizeof(T);
      typedef T vec __attribute__((ext_vector_type(V)));
      const int mv = m / V;
      for (int i = 0; i < mv; ++i) {
        const vec q = reinterpret_cast<const vec*>(buf)[i];
#pragma unroll
        for (int t = 0; t < V; ++t) n = n + q[t];
      }
      for (int i = mv * V; i < m; ++i) n = n + buf[i];
    }
  }
  if (threadIdx.x == 0) {
    n = dev_sqrt<T>(n);
    nsh = (T)(((double)n > 1e-12) ? (double)n : 1e-12);
  }
  __syncthreads();
  const T nn = nsh;
  for (size_t i = threadIdx.x; i < len; i += blockDim.x) e[i] = e[i] / nn;
}

// Hard assignments (one cluster per descriptor): workgroup (image, block of
// CB clusters), thread d owns dimension d of every cluster of the block and
// walks the image's descriptors in index order, adding x_id to the LDS
// accumulator of the descriptor's cluster when it is in the block -- the same
// sequence of roundings per (cluster, dimension) as vlad_accum_kernel with a
// one-hot q (vlad.c:181-188), without scanning every cluster for every row.
// The rows of the block are found from a ballot over 64 cluster indices and
// their coordinates loaded VH_RS at a time, so that a few memory round trips
// serve 64 rows.
// Finalisation as vlad_accum_kernel (vlad.c:191-227).  D <= 128.  Blocks of
// 32 clusters: every workgroup walks all of its image's cluster indices, so
// small blocks mean many short-lived workgroups (16 per image at K = 512)
// that keep the device busy whatever the number of images.
#ifndef VH_RS
#define VH_RS 16
#endif
constexpr int VH_COMPS = 32;

template <typename T, int CB>
__global__ __launch_bounds__(128) void vlad_hard_kernel(const T* __restrict__ x, const uint64_t* __restrict__ offsets,
                                                        int numImages, int D, int K, const T* __restrict__ means,
                                                        const uint32_t* __restrict__ hard, int flags,
                                                        T* __restrict__ enc, int numKB) {
#pragma clang fp contract(off)
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* acc = reinterpret_cast<T*>(smem_raw);                             // [CB][D]
  double* mass = reinterpret_cast<double*>(smem_raw + (((size_t)CB * D * sizeof(T) + 15) & ~size_t(15)));
  T* nrm = reinterpret_cast<T*>(mass + CB);                           // [CB]
  const int b = blockIdx.x;
  int img, kb;
  if (numImages >= 8) {
    const int xcd = b & 7, slot = b >> 3;
    img = (slot / numKB) * 8 + xcd;
    kb = slot % numKB;
  } else {
    img = b / numKB;
    kb = b % numKB;
  }
  if (img >= numImages) return;
  const int tid = threadIdx.x;
  const int k0 = kb * CB;
  for (int i = tid; i < CB * D; i += blockDim.x) acc[i] = T(0);
  for (int i = tid; i < CB; i += blockDim.x) mass[i] = 0.0;
  __syncthreads();
  const long long rBegin = (long long)offsets[img], rEnd = (long long)offsets[img + 1];
  const int d = tid, lane = tid & 63;
  // 64 rows at a time: each wave reads their cluster indices (one per lane),
  // the rows of this block are picked from the ballot in index order, their
  // coordinates fetched VH_RS at a time, then added in order.  The two waves
  // own disjoint dimensions and walk the rows independently.
  for (long long r0 = rBegin; r0 < rEnd; r0 += 64) {
    const long long rr = r0 + lane;
    const uint32_t kk = rr < rEnd ? hard[rr] : 0xffffffffu;
    const int kl = (int)kk - k0;
    unsigned long long mask = __ballot(kk != 0xffffffffu && kl >= 0 && kl < CB);
    while (mask) {
      int us[VH_RS], ks[VH_RS];
#pragma unroll
      for (int q = 0; q < VH_RS; ++q) {
        if (mask) {
          const int u = __builtin_ctzll(mask);
          mask &= mask - 1;
          us[q] = u;
          ks[q] = __shfl(kl, u, 64);
        } else {
          us[q] = -1;
          ks[q] = -1;
        }
      }
      T xv[VH_RS];
#pragma unroll
      for (int q = 0; q < VH_RS; ++q) xv[q] = (us[q] >= 0 && d < D) ? x[(r0 + us[q]) * D + d] : T(0);
#pragma unroll
      for (int q = 0; q < VH_RS; ++q) {
        if (ks[q] >= 0) {
          if (d < D) acc[ks[q] * D + d] = (T)((double)acc[ks[q] * D + d] + (double)xv[q]);
          if (tid == 0) mass[ks[q]] += 1.0;
        }
      }
    }
  }
  __syncthreads();
  // vlad.c:191-213 per (cluster, dimension), then the per-cluster norms
  for (int i = tid; i < CB * D; i += blockDim.x) {
    const int c = i / D, dd = i - c * D;
    const int k = k0 + c;
    if (k >= K) continue;
    T e = acc[i];
    const T mu = means[(size_t)k * D + dd];
    const double ms = mass[c];
    if (ms > 0) {
      if (flags & 8) {
        e = (T)((double)e / ms);
        e = e - mu;
      } else {
        e = (T)((double)e - ms * (double)mu);
      }
    }
    if (flags & 2) e = signed_sqrt(e);
    acc[i] = e;
  }
  __syncthreads();
  if (flags & 1) {
    for (int c = tid; c < CB; c += blockDim.x) {
      T n = T(0);
      for (int dd = 0; dd < D; ++dd) {
        const T z = acc[c * D + dd];
        n = n + z * z;
      }
      n = dev_sqrt<T>(n);
      nrm[c] = (T)(((double)n > 1e-12) ? (double)n : 1e-12);
    }
    __syncthreads();
  }
  for (int i = tid; i < CB * D; i += blockDim.x) {
    const int c = i / D;
    const int k = k0 + c;
    if (k >= K) continue;
    const T z = acc[i];
    enc[(size_t)img * K * D + (size_t)k * D + (i - c * D)] = (flags & 1) ? z / nrm[c] : z;
  }
}

template <typename T, int CB>
static void launch_vlad_hard(T* enc, const T* means, int D, int K, const T* x, const uint64_t* offsets, int numImages,
                             const uint32_t* hard, int flags) {
  const int nkb = (K + CB - 1) / CB;
  const int padImages = numImages >= 8 ? ((numImages + 7) / 8) * 8 : numImages;
  const size_t sh = (((size_t)CB * D * sizeof(T) + 15) & ~size_t(15)) + CB * sizeof(double) + CB * sizeof(T);
  auto kern = vlad_hard_kernel<T, CB>;
  VLHIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  VLHIP_LAUNCH("vlad_hard", kern, dim3((unsigned)((size_t)padImages * nkb)), dim3(128), sh, x, offsets, numImages, D,
               K, means, hard, flags, enc, nkb);
}

template <typename T, int DPT>
static void launch_vlad(T* enc, const T* means, int D, int K, const T* x, const uint64_t* offsets, int numImages,
                        const T* assign, const uint32_t* hard, int flags) {
  const int numKB = (K + VL_COMPS - 1) / VL_COMPS;
  const size_t shmem = ((size_t)VL_ROWS * (D + 1) + VL_ROWS * VL_COMPS + (size_t)VL_COMPS * D) * sizeof(T);
  const int padImages = numImages >= 8 ? ((numImages + 7) / 8) * 8 : numImages;
  const unsigned blocks = (unsigned)((size_t)padImages * numKB);
  if (hard && D <= 128) {
    launch_vlad_hard<T, VH_COMPS>(enc, means, D, K, x, offsets, numImages, hard, flags);
  } else if (hard) {
    auto kern = vlad_accum_kernel<T, DPT, true>;
    VLHIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    VLHIP_LAUNCH("vlad_accum", kern, dim3(blocks), dim3(256), shmem, x, offsets, numImages, D, K, means, assign,
                 hard, flags, enc, numKB);
  } else {
    auto kern = vlad_accum_kernel<T, DPT, false>;
    VLHIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
    VLHIP_LAUNCH("vlad_accum", kern, dim3(blocks), dim3(256), shmem, x, offsets, numImages, D, K, means, assign,
                 hard, flags, enc, numKB);
  }
  if (!(flags & 4)) {
    VLHIP_LAUNCH("vlad_normalize", (vlad_normalize_kernel<T>), dim3((unsigned)numImages), dim3(256), 0, enc,
                 (size_t)K * D);
  }
}

template <typename T>
void vlad_t(T* enc, const T* means, size_t D, size_t K, const T* x, const uint64_t* offsets, size_t numImages,
            const T* assign, const uint32_t* hard, int flags) {
  VLHIP_REQUIRE(K >= 1 && D >= 1, "numClusters and dimension must be positive");
  VLHIP_REQUIRE((assign != nullptr) != (hard != nullptr), "exactly one of assignments / hardAssignments");
  VLHIP_REQUIRE(D <= 512, "GPU VLAD encoder supports dimension <= 512");
  if (numImages == 0) return;
  const int d = (int)D, k = (int)K, n = (int)numImages;
  if (D <= 16) launch_vlad<T, 1>(enc, means, d, k, x, offsets, n, assign, hard, flags);
  else if (D <= 32) launch_vlad<T, 2>(enc, means, d, k, x, offsets, n, assign, hard, flags);
  else if (D <= 64) launch_vlad<T, 4>(enc, means, d, k, x, offsets, n, assign, hard, flags);
  else if (D <= 80) launch_vlad<T, 5>(enc, means, d, k, x, offsets, n, assign, hard, flags);
  else if (D <= 128) launch_vlad<T, 8>(enc, means, d, k, x, offsets, n, assign, hard, flags);
  else if (D <= 256) launch_vlad<T, 16>(enc, means, d, k, x, offsets, n, assign, hard, flags);
  else launch_vlad<T, 32>(enc, means, d, k, x, offsets, n, assign, hard, flags);
}

template void vlad_t<float>(float*, const float*, size_t, size_t, const float*, const uint64_t*, size_t,
                            const float*, const uint32_t*, int);
template void vlad_t<double>(double*, const double*, size_t, size_t, const double*, const uint64_t*, size_t,
                             const double*, const uint32_t*, int);

}  // namespace vlhip

using namespace vlhip;

extern "C" void vlhip_vlad_encode_batch(int dtype, void* enc, void const* means, size_t D, size_t K,
                                        void const* data, uint64_t const* offsets, size_t numImages,
                                        void const* assignments, uint32_t const* hard, int flags) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (dtype == 1)
    vlad_t<float>((float*)enc, (const float*)means, D, K, (const float*)data, offsets, numImages,
                  (const float*)assignments, hard, flags);
  else if (dtype == 2)
    vlad_t<double>((double*)enc, (const double*)means, D, K, (const double*)data, offsets, numImages,
                   (const double*)assignments, hard, flags);
  else
    fatal("vlhip_vlad_encode_batch", __FILE__, __LINE__, "unknown data type");
}
