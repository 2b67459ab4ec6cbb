// fisher_kernels.hip -- Fisher vector encoder (vl/fisher.c), FP32/FP64
// direct-form path.
//
// Per group of images whose N x K posterior matrix fits a 4 GiB budget:
//   1. the E-step kernel (gmm_kernels.hip) writes the normalised posteriors
//      of the group's descriptors (gmm.c:758-791) and their argmax;
//   2. fisher_accum_kernel: workgroup (image, component block) thresholds the
//      posteriors as fisher.c:436-441 does, stages them in LDS and accumulates
//        u[k,d] = sum_i q_ik t_ikd,  v[k,d] = sum_i q_ik t_ikd^2 - sum_i q_ik,
//        t_ikd = (x_id - mu_kd) * sigma_kd^-1            (fisher.c:442-447)
//      in registers, then applies the prefixes, the signed square root and
//      writes its slice of the encoding plus its share of the squared norm;
// then fisher_normalize_kernel divides each image's encoding by its L2 norm
// (fisher.c:474-485).
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

#include <algorithm>
#include <vector>

namespace vlhip {

constexpr int FA_ROWS = 64;   // descriptors per chunk
constexpr int FA_COMPS = 64;  // components per workgroup
constexpr int FA_PAD = 4;

// XCD-aware block mapping: the hardware deals workgroups to the 8 XCDs round
// robin, so all component blocks of one image are placed on the same XCD and
// share its L2 copy of the image's descriptors and posteriors.
__device__ __forceinline__ void fisher_block_coords(int numKB, int numImages, int& img, int& kb) {
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

// Finalisation of one workgroup's slice (fisher.c:449-458, 466-471): prefixes
// 1/(n sqrt(pi)) and 1/(n sqrt(2 pi)), v -= S0, the signed square root, and the
// slice's share of the squared norm.  Thread (kg, dg) owns components
// k0 + kg*4 + c and dimensions dg + 16 j; s0 is indexed by kg*4 + c.
template <typename T, int DPT>
__device__ __forceinline__ double fisher_finalize_slice(const T (&u)[4][DPT], const T (&v)[4][DPT], const T* s0,
                                                        const T* __restrict__ priors, int k0, int kg, int dg,
                                                        int D, int K, long long nImg, int flags, T* encU, T* encV) {
  const bool sqrtFlag = (flags & 1) != 0;
  double nrm = 0.0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int cl = kg * 4 + c;
    const int k = k0 + cl;
    if (k >= K) continue;
    const double prior = (double)priors[k];
    const bool skip = prior < 1e-6;
    T up = T(1), vp = T(1);
    if (nImg > 0 && !skip) {
      up = (T)(1.0 / ((double)nImg * sqrt(prior)));
      vp = (T)(1.0 / ((double)nImg * sqrt(2.0 * prior)));
    }
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const int d = dg + 16 * j;
      if (d >= D) continue;
      T uu = T(0), vv = T(0);
      if (!skip) {
        uu = u[c][j];
        vv = v[c][j] - s0[cl];
        if (nImg > 0) { uu = uu * up; vv = vv * vp; }
      }
      if (sqrtFlag) { uu = signed_sqrt(uu); vv = signed_sqrt(vv); }
      encU[(size_t)k * D + d] = uu;
      encV[(size_t)k * D + d] = vv;
      nrm += (double)uu * (double)uu + (double)vv * (double)vv;
    }
  }
  return nrm;
}

// Posterior-weighted accumulation, fisher.c:419-458.  The normalised
// posteriors q (rows of this image group x K) come from the E-step kernel.
// Workgroup (image, 64 components): per 64-row chunk the descriptors
// (transposed) and the thresholded posteriors are staged in LDS; thread
// (kg, dg) owns components kg*4.. and dimensions dg + 16 j and accumulates
//   u[k,d] = sum_i q_ik t_ikd,  v[k,d] = sum_i q_ik t_ikd^2 - sum_i q_ik,
//   t_ikd = (x_id - mu_kd) * sigma_kd^-1
// in registers, following the reference's per-component loop over data in
// index order.  Posteriors below 1e-6 and components with prior below 1e-6
// are skipped and counted as in fisher.c:436-441; with FAST the posterior is
// the one-hot argmax of the E-step (fisher.c:399-417).
//
// The rows of an image are summed in segments of 4 chunks: each segment's
// sums start from zero and are added, segment after segment, to the image's
// totals (kept in the workgroup's own slice of the encoding until the
// finalisation).  When there are too few (image, component block) pairs to
// fill the device, the segments run in parallel instead (gridDim.y = the
// largest segment count): segment s of an image writes its sums to
// part[(image within the launch, s)] = [u K x D | v K x D | S0 K] and
// fisher_split_combine_kernel adds them in segment order.  Both orders of
// operations are the same, so an image's encoding does not depend on which
// way it was computed (nor, therefore, on the other images of a batch).
template <typename T, int DPT>
__global__ __launch_bounds__(256) void fisher_accum_kernel(
    const T* __restrict__ x, const uint64_t* __restrict__ offsets, int imgBegin, int numImages,
    long long qRowBase, int D, int K, const T* __restrict__ means, const T* __restrict__ isig,
    const T* __restrict__ priors, const T* __restrict__ q, const int* __restrict__ row_arg, int flags,
    T* __restrict__ enc, double* __restrict__ normpart, unsigned long long* __restrict__ numTerms, int numKB,
    int chunk, T* __restrict__ part) {
  __shared__ __align__(16) T qs[FA_ROWS][FA_COMPS + FA_PAD];
  __shared__ double red[256];
  __shared__ T s0sh[4][FA_COMPS];
  __shared__ T s0tot[FA_COMPS];
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* xs = reinterpret_cast<T*>(smem_raw);  // [D][chunk + FA_PAD]
  const int XS = chunk + FA_PAD;

  int li, kb;
  fisher_block_coords(numKB, numImages, li, kb);
  if (li >= numImages) return;
  const int img = imgBegin + li;

  const int tid = threadIdx.x;
  const int sc = tid & 63, sr = tid >> 6;  // staging: component sc, rows sr + 4 j
  const int kg = tid >> 4, dg = tid & 15;  // accumulation: comps kg*4.., dims dg + 16 j
  const int k0 = kb * FA_COMPS;
  const long long iBegin = (long long)offsets[img];
  const long long iEnd = (long long)offsets[img + 1];
  const long long nImg = iEnd - iBegin;
  const long long segRows = 4LL * chunk;
  const long long nseg = (nImg + segRows - 1) / segRows;
  const bool splitMode = gridDim.y > 1;
  const long long sgBegin = splitMode ? (long long)blockIdx.y : 0;
  const long long sgEnd = splitMode ? min(nseg, sgBegin + 1) : nseg;
  if (splitMode && sgBegin >= nseg) return;
  const bool fast = (flags & 4) != 0;
  const int ks = k0 + sc;
  const bool ksok = ks < K && (double)priors[ks] >= 1e-6;
  T* encU = enc + (size_t)img * 2 * K * D;
  T* encV = encU + (size_t)K * D;

  T mu[4][DPT], sg[4][DPT], u[4][DPT], v[4][DPT];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const int k = k0 + kg * 4 + c, d = dg + 16 * j;
      const bool ok = (k < K && d < D);
      mu[c][j] = ok ? means[(size_t)k * D + d] : T(0);
      sg[c][j] = ok ? isig[(size_t)k * D + d] : T(0);
    }
  unsigned terms = 0;

  for (long long seg = sgBegin; seg < sgEnd; ++seg) {
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int j = 0; j < DPT; ++j) { u[c][j] = T(0); v[c][j] = T(0); }
    T s0acc = T(0);
    const long long sBegin = iBegin + seg * segRows, sEnd = min(iEnd, sBegin + segRows);
    for (long long r0 = sBegin; r0 < sEnd; r0 += chunk) {
      const int nr = (int)min((long long)chunk, sEnd - r0);
      __syncthreads();
      // staging: each thread's loads are issued together (16 at a time), then stored
      for (int base = 0; base < chunk * D; base += 256 * 16) {
        T xv[16];
#pragma unroll
        for (int u2 = 0; u2 < 16; ++u2) {
          const int idx = base + tid + 256 * u2;
          const int r = idx / D;
          xv[u2] = (idx < chunk * D && r < nr) ? x[(r0 + r) * D + (idx - r * D)] : T(0);
        }
#pragma unroll
        for (int u2 = 0; u2 < 16; ++u2) {
          const int idx = base + tid + 256 * u2;
          if (idx < chunk * D) {
            const int r = idx / D;
            xs[(idx - r * D) * XS + r] = xv[u2];
          }
        }
      }
      {
        T pv[FA_ROWS / 4];
#pragma unroll
        for (int j = 0; j < FA_ROWS / 4; ++j) {
          const int r = sr + 4 * j;
          T p = T(0);
          if (j < chunk / 4 && r < nr && ksok) {
            const long long row = r0 + r;
            if (fast) p = (ks == row_arg[row]) ? T(1) : T(0);
            else p = q[(row - qRowBase) * K + ks];
          }
          pv[j] = p;
        }
#pragma unroll
        for (int j = 0; j < FA_ROWS / 4; ++j) {
          if (j >= chunk / 4) break;
          const int r = sr + 4 * j;
          T p = pv[j];
          if (r < nr && ksok) {
            if ((double)p < 1e-6) p = T(0); else terms += 1;
          }
          qs[r][sc] = p;
          s0acc += p;
        }
      }
      __syncthreads();
      for (int r = 0; r < nr; ++r) {
        T qv[4];
        lds4(&qs[r][kg * 4], qv);
        if (qv[0] == T(0) && qv[1] == T(0) && qv[2] == T(0) && qv[3] == T(0)) continue;
        T xv[DPT];
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
          const int d = dg + 16 * j;
          xv[j] = (d < D) ? xs[d * XS + r] : T(0);
        }
#pragma unroll
        for (int c = 0; c < 4; ++c) {
          const T qc = qv[c];
#pragma unroll
          for (int j = 0; j < DPT; ++j) {
            const T t = (xv[j] - mu[c][j]) * sg[c][j];
            const T qt = qc * t;
            u[c][j] += qt;
            v[c][j] = fma(qt, t, v[c][j]);
          }
        }
      }
    }
    // segment mass per component: the 4 staging threads of each component
    s0sh[sr][sc] = s0acc;
    __syncthreads();
    if (splitMode) {
      T* P = part + ((size_t)li * gridDim.y + seg) * (size_t)K * (2 * D + 1);
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int k = k0 + kg * 4 + c;
        if (k >= K) continue;
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
          const int d = dg + 16 * j;
          if (d >= D) continue;
          P[(size_t)k * D + d] = u[c][j];
          P[(size_t)K * D + (size_t)k * D + d] = v[c][j];
        }
      }
      if (tid < FA_COMPS && k0 + tid < K)
        P[2 * (size_t)K * D + k0 + tid] = ((s0sh[0][tid] + s0sh[1][tid]) + s0sh[2][tid]) + s0sh[3][tid];
    } else {
      // running totals in this workgroup's own slice of the encoding
#pragma unroll
      for (int c = 0; c < 4; ++c) {
        const int k = k0 + kg * 4 + c;
        if (k >= K) continue;
#pragma unroll
        for (int j = 0; j < DPT; ++j) {
          const int d = dg + 16 * j;
          if (d >= D) continue;
          const size_t o = (size_t)k * D + d;
          encU[o] = seg == 0 ? u[c][j] : encU[o] + u[c][j];
          encV[o] = seg == 0 ? v[c][j] : encV[o] + v[c][j];
        }
      }
      if (tid < FA_COMPS) {
        const T s = ((s0sh[0][tid] + s0sh[1][tid]) + s0sh[2][tid]) + s0sh[3][tid];
        s0tot[tid] = seg == 0 ? s : s0tot[tid] + s;
      }
    }
  }
  if (numTerms) {
    const double bt = block_sum_256((double)terms, red);
    if (tid == 0 && bt > 0) atomicAdd(&numTerms[img], (unsigned long long)bt);
  }
  if (splitMode) return;

  // totals of the image (zero without rows), then the finalisation
  __syncthreads();
  if (nImg == 0 && tid < FA_COMPS) s0tot[tid] = T(0);
  __syncthreads();
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int k = k0 + kg * 4 + c;
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const int d = dg + 16 * j;
      const bool ok = nImg > 0 && k < K && d < D;
      u[c][j] = ok ? encU[(size_t)k * D + d] : T(0);
      v[c][j] = ok ? encV[(size_t)k * D + d] : T(0);
    }
  }
  const double nrm = fisher_finalize_slice<T, DPT>(u, v, s0tot, priors, k0, kg, dg, D, K, nImg, flags, encU, encV);
  const double bn = block_sum_256(nrm, red);
  if (tid == 0) normpart[(size_t)img * numKB + kb] = bn;
}

// Sum of the segments of fisher_accum_kernel's split mode (in segment order,
// in T, from zero -- the order the sequential mode adds them in) and the
// finalisation; same workgroup / thread ownership as the accumulation.
template <typename T, int DPT>
__global__ __launch_bounds__(256) void fisher_split_combine_kernel(
    const T* __restrict__ part, int splits, long long segRows, const uint64_t* __restrict__ offsets, int imgBegin,
    int numImages, int D, int K, const T* __restrict__ priors, int flags, T* __restrict__ enc,
    double* __restrict__ normpart, int numKB) {
  __shared__ double red[256];
  __shared__ T s0c[FA_COMPS];
  int li, kb;
  fisher_block_coords(numKB, numImages, li, kb);
  if (li >= numImages) return;
  const int img = imgBegin + li;
  const int tid = threadIdx.x, kg = tid >> 4, dg = tid & 15;
  const int k0 = kb * FA_COMPS;
  const long long nImg = (long long)offsets[img + 1] - (long long)offsets[img];
  const int nseg = (int)((nImg + segRows - 1) / segRows);
  const size_t PS = (size_t)K * (2 * D + 1);
  const T* P0 = part + (size_t)li * splits * PS;
  T u[4][DPT], v[4][DPT];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int j = 0; j < DPT; ++j) { u[c][j] = T(0); v[c][j] = T(0); }
  // segment by segment (every load of a segment issued together), each sum
  // from zero in segment order
  for (int sp = 0; sp < nseg; ++sp) {
    T pu[4][DPT], pv[4][DPT];
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int j = 0; j < DPT; ++j) {
        const int k = k0 + kg * 4 + c, d = dg + 16 * j;
        const bool ok = k < K && d < D;
        pu[c][j] = ok ? P0[sp * PS + (size_t)k * D + d] : T(0);
        pv[c][j] = ok ? P0[sp * PS + (size_t)K * D + (size_t)k * D + d] : T(0);
      }
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
      for (int j = 0; j < DPT; ++j) {
        u[c][j] += pu[c][j];
        v[c][j] += pv[c][j];
      }
  }
  if (tid < FA_COMPS) {
    T s0 = T(0);
    if (k0 + tid < K)
      for (int sp = 0; sp < nseg; ++sp) s0 += P0[sp * PS + 2 * (size_t)K * D + k0 + tid];
    s0c[tid] = s0;
  }
  __syncthreads();
  T* encU = enc + (size_t)img * 2 * K * D;
  T* encV = encU + (size_t)K * D;
  const double nrm = fisher_finalize_slice<T, DPT>(u, v, s0c, priors, k0, kg, dg, D, K, nImg, flags, encU, encV);
  const double bn = block_sum_256(nrm, red);
  if (tid == 0) normpart[(size_t)img * numKB + kb] = bn;
}

// fisher.c:474-485: enc /= max(||enc||_2, 1e-12).  Workgroup (image = x,
// slice of FN_SLICE entries = y): each workgroup of an image sums the same norm partials in
// the same order, so they all divide by the same value.  The slice's loads are
// issued before the norm is formed, so that they overlap its latency.
constexpr int FN_SLICE = 8192;
template <typename T>
__global__ __launch_bounds__(256) void fisher_normalize_kernel(T* __restrict__ enc, const double* __restrict__ normpart,
                                                              int numKB, size_t len) {
  __shared__ T nsh;
  const int img = blockIdx.x;
  T* e = enc + (size_t)img * len;
  const size_t s0 = (size_t)blockIdx.y * FN_SLICE;
  const size_t end = min(len, s0 + FN_SLICE);
  constexpr int V = 16 / sizeof(T);
  typedef T vec __attribute__((ext_vector_type(V)));
  auto norm_of_image = [&]() {
    if (threadIdx.x == 0) {
      double s = 0.0;
      for (int b = 0; b < numKB; ++b) s += normpart[(size_t)img * numKB + b];
      const T n = (T)sqrt(s);
      nsh = (T)(((double)n > 1e-12) ? (double)n : 1e-12);
    }
    __syncthreads();
    return nsh;
  };
  if (len % V == 0 && ((uintptr_t)e & 15) == 0) {
    // 16-byte pieces, every load of the thread issued before the first store
    constexpr int U = FN_SLICE / V / 256;
    vec* ev = reinterpret_cast<vec*>(e);
    vec v[U];
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = s0 / V + threadIdx.x + 256 * u;
      if (i * V < end) v[u] = ev[i];
    }
    const T n = norm_of_image();
#pragma unroll
    for (int u = 0; u < U; ++u) {
      const size_t i = s0 / V + threadIdx.x + 256 * u;
      if (i * V < end) {
#pragma unroll
        for (int t = 0; t < V; ++t) v[u][t] = v[u][t] / n;
        ev[i] = v[u];
      }
    }
  } else {
    const T n = norm_of_image();
    for (size_t i = s0 + threadIdx.x; i < end; i += blockDim.x) e[i] = e[i] / n;
  }
}

template <typename T>
static void launch_normalize_t(T* enc, const double* normpart, int numKB, size_t numImages, size_t len) {
  VLHIP_LAUNCH("fisher_normalize", (fisher_normalize_kernel<T>),
               dim3((unsigned)numImages, (unsigned)((len + FN_SLICE - 1) / FN_SLICE)), dim3(256), 0, enc, normpart,
               numKB, len);
}

void launch_fisher_normalize(float* enc, const double* normpart, int numKB, size_t numImages, size_t len) {
  launch_normalize_t<float>(enc, normpart, numKB, numImages, len);
}

constexpr size_t kAccumLds = 96 * 1024;

template <typename T, int DPT>
static void launch_fisher_accum(const T* x, const uint64_t* offsets, int imgBegin, int numImages, long long qRowBase,
                                int D, int K, const T* means, const GmmPrep<T>& prep, const T* priors, const T* q,
                                const int* ra, int flags, T* enc, double* normpart, unsigned long long* numTerms,
                                int numKB, long long maxRows) {
  // rows per chunk: 64, fewer when the transposed descriptor tile would exceed 96 KiB of LDS
  int chunk = FA_ROWS;
  while (chunk > 4 && (size_t)D * (chunk + FA_PAD) * sizeof(T) > kAccumLds) chunk -= 4;
  const size_t shmem = (size_t)D * (chunk + FA_PAD) * sizeof(T);
  auto kern = fisher_accum_kernel<T, DPT>;
  VLHIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  const int padImages = numImages >= 8 ? ((numImages + 7) / 8) * 8 : numImages;
  const unsigned blocks = (unsigned)((size_t)padImages * numKB);
  // segments in parallel when (image, component block) pairs alone would
  // leave most of the device idle (at most 64k workgroups, 1 GiB of sums)
  const int cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
  const long long segRows = 4LL * chunk;
  const long long maxSeg = (maxRows + segRows - 1) / segRows;
  const size_t partBytes = (size_t)numImages * (size_t)maxSeg * K * (2 * D + 1) * sizeof(T);
  const bool split = (long long)numImages * numKB < 2LL * cus && maxSeg > 1 &&
                     (long long)blocks * maxSeg <= 65536 && partBytes <= (size_t(1) << 30);
  T* part = split ? (T*)scratch(SLOT_SEGMENTS, partBytes) : nullptr;
  VLHIP_LAUNCH("fisher_accum", kern, dim3(blocks, split ? (unsigned)maxSeg : 1u), dim3(256), shmem, x, offsets,
               imgBegin, numImages, qRowBase, D, K, means, prep.isig, priors, q, ra, flags, enc, normpart, numTerms,
               numKB, chunk, part);
  if (split) {
    VLHIP_LAUNCH("fisher_split_combine", (fisher_split_combine_kernel<T, DPT>), dim3(blocks), dim3(256), 0, part,
                 (int)maxSeg, segRows, offsets, imgBegin, numImages, D, K, priors, flags, enc, normpart, numKB);
  }
}

// Rows of posteriors materialised at once (bounds the q buffer to 4 GiB of FP32).
constexpr size_t kPosteriorBudget = size_t(4) << 30;

template <typename T>
void fisher_fp_t(T* enc, const T* means, const T* covs, const T* priors, size_t D, size_t K, const T* x,
                 const uint64_t* offsets, size_t numImages, int flags, unsigned long long* numTerms) {
  VLHIP_REQUIRE(K >= 1 && D >= 1, "numClusters and dimension must be positive");
  VLHIP_REQUIRE(D <= 256, "GPU Fisher encoder supports dimension <= 256");
  if (numImages == 0) return;
  std::vector<uint64_t> off(numImages + 1);
  download(off.data(), offsets, (numImages + 1) * sizeof(uint64_t));
  const uint64_t total = off[numImages];
  GmmPrep<T> prep = gmm_prepare<T>(priors, covs, D, K);
  int* ra = (int*)scratch(SLOT_ROWSTAT, (size_t)(total + 1) * sizeof(int) + 512);
  const int numKB = (int)((K + FA_COMPS - 1) / FA_COMPS);
  double* normpart = (double*)scratch(SLOT_PARTIAL, numImages * (size_t)numKB * sizeof(double));
  if (numTerms) VLHIP_CHECK(hipMemsetAsync(numTerms, 0, numImages * sizeof(unsigned long long), ctx().stream));
  const int d = (int)D, k = (int)K;
  // image groups whose posterior matrix fits the budget (at least one image)
  size_t g0 = 0;
  while (g0 < numImages) {
    size_t g1 = g0 + 1;
    while (g1 < numImages && (off[g1 + 1] - off[g0]) * K * sizeof(T) <= kPosteriorBudget) ++g1;
    const uint64_t rows = off[g1] - off[g0];
    T* q = (T*)scratch(SLOT_FISHER_WORK, (size_t)(rows ? rows : 1) * K * sizeof(T));
    run_estep<T>(x + off[g0] * D, rows, D, K, means, prep, q, nullptr, nullptr, ra + off[g0], "fisher_estep", false);
    const int n = (int)(g1 - g0), b = (int)g0;
    const long long qb = (long long)off[g0];
    long long mr = 0;
    for (size_t i = g0; i < g1; ++i) mr = std::max<long long>(mr, (long long)(off[i + 1] - off[i]));
    if (D <= 16)       launch_fisher_accum<T, 1>(x, offsets, b, n, qb, d, k, means, prep, priors, q, ra, flags, enc, normpart, numTerms, numKB, mr);
    else if (D <= 32)  launch_fisher_accum<T, 2>(x, offsets, b, n, qb, d, k, means, prep, priors, q, ra, flags, enc, normpart, numTerms, numKB, mr);
    else if (D <= 64)  launch_fisher_accum<T, 4>(x, offsets, b, n, qb, d, k, means, prep, priors, q, ra, flags, enc, normpart, numTerms, numKB, mr);
    else if (D <= 80)  launch_fisher_accum<T, 5>(x, offsets, b, n, qb, d, k, means, prep, priors, q, ra, flags, enc, normpart, numTerms, numKB, mr);
    else if (D <= 128) launch_fisher_accum<T, 8>(x, offsets, b, n, qb, d, k, means, prep, priors, q, ra, flags, enc, normpart, numTerms, numKB, mr);
    else               launch_fisher_accum<T, 16>(x, offsets, b, n, qb, d, k, means, prep, priors, q, ra, flags, enc, normpart, numTerms, numKB, mr);
    g0 = g1;
  }
  if (flags & 2) {
    launch_normalize_t<T>(enc, normpart, numKB, numImages, (size_t)2 * K * D);
  }
}

}  // namespace vlhip

using namespace vlhip;

extern "C" void vlhip_fisher_set_gate(void* gate, int candidates, float need) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  VLHIP_REQUIRE(gate == nullptr || (candidates >= 1 && candidates <= 16), "gate candidates must be in [1, 16]");
  ctx().gate = gate;
  ctx().gate_m = candidates;
  ctx().gate_need = need;
}

extern "C" void vlhip_fisher_gate_stats(unsigned long long* groups, unsigned long long* served, int reset) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (groups) *groups = ctx().gate_groups;
  if (served) *served = ctx().gate_served;
  if (reset) ctx().gate_groups = ctx().gate_served = 0;
}

extern "C" void vlhip_fisher_encode_batch(int dtype, void* enc, void const* means, void const* covs,
                                          void const* priors, size_t D, size_t K, void const* data,
                                          uint64_t const* offsets, size_t numImages, int flags, int mode,
                                          unsigned long long* numTerms) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (dtype == 2) {
    VLHIP_REQUIRE(mode == VLHIP_MODE_FP32, "double data supports only the direct-form mode");
    fisher_fp_t<double>((double*)enc, (const double*)means, (const double*)covs, (const double*)priors, D, K,
                        (const double*)data, offsets, numImages, flags, numTerms);
    return;
  }
  VLHIP_REQUIRE(dtype == 1, "unknown data type");
  if (mode == VLHIP_MODE_FP32) {
    fisher_fp_t<float>((float*)enc, (const float*)means, (const float*)covs, (const float*)priors, D, K,
                       (const float*)data, offsets, numImages, flags, numTerms);
  } else if (mode == VLHIP_MODE_BF16X3 || mode == VLHIP_MODE_F16X3) {
    fisher_split16((float*)enc, (const float*)means, (const float*)covs, (const float*)priors, D, K,
                   (const float*)data, offsets, numImages, flags, numTerms, mode == VLHIP_MODE_F16X3, nullptr);
  } else if (mode == VLHIP_MODE_FP32_MFMA) {
    fisher_f32mfma((float*)enc, (const float*)means, (const float*)covs, (const float*)priors, D, K,
                   (const float*)data, offsets, numImages, flags, numTerms);
  } else if (mode == VLHIP_MODE_GATED) {
    VLHIP_REQUIRE(ctx().gate != nullptr, "VLHIP_MODE_GATED needs a gate (vlhip_fisher_set_gate)");
    const GateUse gu{ctx().gate, ctx().gate_m, ctx().gate_need};
    fisher_split16((float*)enc, (const float*)means, (const float*)covs, (const float*)priors, D, K,
                   (const float*)data, offsets, numImages, flags, numTerms, true, &gu);
  } else {
    fatal("vlhip_fisher_encode_batch", __FILE__, __LINE__, "unknown mode");
  }
}
