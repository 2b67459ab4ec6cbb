// fisher_kernels.hip -- Fisher vector encoder (vl/fisher.c), FP32/FP64
// direct-form path.
//
// Two passes per batch of images:
//   1. the E-step kernel (gmm_kernels.hip) sweeps all K components for every
//      descriptor and keeps only the row statistics (max logit, sum of
//      exponentials, argmax) -- no N x K posterior matrix is written;
//   2. fisher_accum_kernel: workgroup (image, component block) recomputes the
//      logits of its FA_COMPS components with the same instructions as pass 1
//      (so the posteriors are bit-identical to the E-step's), thresholds them
//      as fisher.c:436-441 does, stages them in LDS and accumulates
//        u[k,d] = sum_i q_ik t_ikd,  v[k,d] = sum_i q_ik t_ikd^2 - sum_i q_ik,
//        t_ikd = (x_id - mu_kd) * sigma_kd^-1            (fisher.c:442-447)
//      in registers, then applies the prefixes, the signed square root and
//      writes its slice of the encoding plus its share of the squared norm;
//   3. fisher_normalize_kernel divides each image's encoding by its L2 norm
//      (fisher.c:474-485).
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

// LDS layout of fisher_accum_kernel (in elements of T):
//   xs [D][FA_ROWS + FA_PAD]      descriptors of the chunk, transposed
//   ms [D][FA_COMPS + FA_PAD]     means of the block, transposed   (RESIDENT)
//   ss [D][FA_COMPS + FA_PAD]     1/sigma^2 of the block, transposed (RESIDENT)
//   qs [FA_ROWS][FA_COMPS + FA_PAD] thresholded posteriors
//   rd [16][FA_COMPS]             per-thread-row partial posterior mass
inline size_t fisher_accum_lds_elems(int D, bool resident) {
  size_t e = (size_t)D * (FA_ROWS + FA_PAD) + (size_t)FA_ROWS * (FA_COMPS + FA_PAD) + 16 * FA_COMPS;
  if (resident) e += 2 * (size_t)D * (FA_COMPS + FA_PAD);
  return e;
}

// XCD-aware block mapping: the hardware deals workgroups to the 8 XCDs round
// robin, so all component blocks of one image are placed on the same XCD and
// share its L2 copy of the image's descriptors.
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

template <typename T, int DPT, bool RESIDENT>
__global__ __launch_bounds__(256) void fisher_accum_kernel(
    const T* __restrict__ x, const uint64_t* __restrict__ offsets, int numImages, int D, int K,
    const T* __restrict__ means, const T* __restrict__ inv, const T* __restrict__ isig,
    const double* __restrict__ cterm, const T* __restrict__ priors,
    const T* __restrict__ row_m, const T* __restrict__ row_s, const int* __restrict__ row_arg,
    int flags, T* __restrict__ enc, double* __restrict__ normpart,
    unsigned long long* __restrict__ numTerms, int numKB) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  int img, kb;
  fisher_block_coords(numKB, numImages, img, kb);
  if (img >= numImages) return;

  const int XS = FA_ROWS + FA_PAD, CS = FA_COMPS + FA_PAD;
  T* xs = reinterpret_cast<T*>(smem_raw);
  T* qs = xs + (size_t)D * XS;
  T* rd = qs + (size_t)FA_ROWS * CS;
  T* ms = rd + 16 * FA_COMPS;
  T* ss = ms + (size_t)D * CS;
  __shared__ double red[256];

  const int tid = threadIdx.x;
  const int tx = tid & 15, ty = tid >> 4;  // phase A: rows ty*4.., comps tx*4..
  const int kg = tid >> 4, dg = tid & 15;  // phase B: comps kg*4.., dims dg + 16 j
  const int k0 = kb * FA_COMPS;
  const long long rBegin = (long long)offsets[img];
  const long long rEnd = (long long)offsets[img + 1];
  const long long nImg = rEnd - rBegin;
  const bool fast = (flags & 4) != 0;

  if (RESIDENT) {
    for (int idx = tid; idx < FA_COMPS * D; idx += 256) {
      const int c = idx / D, d = idx - c * D;
      const int k = k0 + c;
      ms[d * CS + c] = (k < K) ? means[(size_t)k * D + d] : T(0);
      ss[d * CS + c] = (k < K) ? inv[(size_t)k * D + d] : T(0);
    }
  }

  // phase-A constants of this thread's components
  double ct[4];
  bool kok[4];
#pragma unroll
  for (int j = 0; j < 4; ++j) {
    const int k = k0 + tx * 4 + j;
    kok[j] = (k < K) && ((double)priors[k] >= 1e-6);
    ct[j] = (k < K) ? cterm[k] : 0.0;
  }
  // phase-B constants of this thread's (component, dimension) pairs
  T mu[4][DPT], sg[4][DPT], u[4][DPT], v[4][DPT];
#pragma unroll
  for (int c = 0; c < 4; ++c)
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const int k = k0 + kg * 4 + c, d = dg + 16 * j;
      const bool ok = (k < K && d < D);
      mu[c][j] = ok ? means[(size_t)k * D + d] : T(0);
      sg[c][j] = ok ? isig[(size_t)k * D + d] : T(0);
      u[c][j] = T(0);
      v[c][j] = T(0);
    }
  T s0acc[4] = {T(0), T(0), T(0), T(0)};
  unsigned long long terms = 0;

  for (long long r0 = rBegin; r0 < rEnd; r0 += FA_ROWS) {
    const int nr = (int)min((long long)FA_ROWS, rEnd - r0);
    __syncthreads();
    for (int idx = tid; idx < FA_ROWS * D; idx += 256) {
      const int r = idx / D, d = idx - r * D;
      xs[d * XS + r] = (r < nr) ? x[(r0 + r) * D + d] : T(0);
    }
    __syncthreads();

    // ---- phase A: logits of FA_ROWS x FA_COMPS, posteriors into LDS
    T acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = T(0);
    for (int d = 0; d < D; ++d) {
      T xv[4], mv[4], sv[4];
      lds4(&xs[d * XS + ty * 4], xv);
      if (RESIDENT) {
        lds4(&ms[d * CS + tx * 4], mv);
        lds4(&ss[d * CS + tx * 4], sv);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = min(k0 + tx * 4 + j, K - 1);
          mv[j] = means[(size_t)k * D + d];
          sv[j] = inv[(size_t)k * D + d];
        }
      }
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const T t = xv[i] - mv[j];
          acc[i][j] = fma(t * t, sv[j], acc[i][j]);
        }
    }
#pragma unroll
    for (int i = 0; i < 4; ++i) {
      const int r = ty * 4 + i;
      const bool rok = r < nr;
      const long long row = r0 + r;
      const T M = rok ? row_m[row] : T(0);
      const T S = rok ? row_s[row] : T(1);
      const int arg = rok ? row_arg[row] : -1;
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = k0 + tx * 4 + j;
        T p = T(0);
        if (rok && kok[j]) {
          if (fast) {
            p = (k == arg) ? T(1) : T(0);
          } else {
            const T l = (T)(ct[j] - 0.5 * (double)acc[i][j]);
            p = dev_exp<T>(l - M) / S;
          }
          if ((double)p < 1e-6) p = T(0); else terms += 1;
        }
        qs[r * CS + tx * 4 + j] = p;
        s0acc[j] += p;
      }
    }
    __syncthreads();

    // ---- phase B: posterior-weighted first and second order residuals
    for (int r = 0; r < nr; ++r) {
      T qv[4];
      lds4(&qs[r * CS + kg * 4], qv);
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

  // ---- per-component posterior mass (sum over the 16 phase-A row groups)
  __syncthreads();
#pragma unroll
  for (int j = 0; j < 4; ++j) rd[ty * FA_COMPS + tx * 4 + j] = s0acc[j];
  __syncthreads();

  // ---- finalize, fisher.c:450-472
  const bool sqrtFlag = (flags & 1) != 0;
  double nrm = 0.0;
  T* encU = enc + (size_t)img * 2 * K * D;
  T* encV = encU + (size_t)K * D;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    const int k = k0 + kg * 4 + c;
    if (k >= K) continue;
    T s0 = T(0);
    for (int t = 0; t < 16; ++t) s0 += rd[t * FA_COMPS + kg * 4 + c];
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
        vv = v[c][j] - s0;
        if (nImg > 0) { uu = uu * up; vv = vv * vp; }
      }
      if (sqrtFlag) { uu = signed_sqrt(uu); vv = signed_sqrt(vv); }
      encU[(size_t)k * D + d] = uu;
      encV[(size_t)k * D + d] = vv;
      nrm += (double)uu * (double)uu + (double)vv * (double)vv;
    }
  }
  const double bn = block_sum_256(nrm, red);
  if (tid == 0) normpart[(size_t)img * numKB + kb] = bn;
  if (numTerms) {
    const double bt = block_sum_256((double)terms, red);
    if (tid == 0) atomicAdd(&numTerms[img], (unsigned long long)bt);
  }
}

// fisher.c:474-485: enc /= max(||enc||_2, 1e-12).
template <typename T>
__global__ __launch_bounds__(256) void fisher_normalize_kernel(T* __restrict__ enc, const double* __restrict__ normpart,
                                                              int numKB, size_t len) {
  __shared__ T nsh;
  const int img = blockIdx.x;
  if (threadIdx.x == 0) {
    double s = 0.0;
    for (int b = 0; b < numKB; ++b) s += normpart[(size_t)img * numKB + b];
    const T n = (T)sqrt(s);
    nsh = (T)(((double)n > 1e-12) ? (double)n : 1e-12);
  }
  __syncthreads();
  const T n = nsh;
  T* e = enc + (size_t)img * len;
  for (size_t i = threadIdx.x; i < len; i += blockDim.x) e[i] = e[i] / n;
}

void launch_fisher_normalize(float* enc, const double* normpart, int numKB, size_t numImages, size_t len) {
  VLHIP_LAUNCH("fisher_normalize", (fisher_normalize_kernel<float>), dim3((unsigned)numImages), dim3(256), 0, enc,
               normpart, numKB, len);
}

template <typename T, int DPT, bool RESIDENT>
static void launch_fisher_accum(const T* x, const uint64_t* offsets, int numImages, int D, int K, const T* means,
                                const GmmPrep<T>& prep, const T* priors, const T* rm, const T* rs, const int* ra,
                                int flags, T* enc, double* normpart, unsigned long long* numTerms, int numKB) {
  const size_t shmem = fisher_accum_lds_elems(D, RESIDENT) * sizeof(T);
  auto kern = fisher_accum_kernel<T, DPT, RESIDENT>;
  VLHIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)shmem));
  const int padImages = numImages >= 8 ? ((numImages + 7) / 8) * 8 : numImages;
  const unsigned blocks = (unsigned)((size_t)padImages * numKB);
  VLHIP_LAUNCH("fisher_accum", kern, dim3(blocks), dim3(256), shmem, x, offsets, numImages, D, K, means,
               prep.inv, prep.isig, prep.cterm, priors, rm, rs, ra, flags, enc, normpart, numTerms, numKB);
}

template <typename T, int DPT>
static void dispatch_resident(bool resident, const T* x, const uint64_t* offsets, int numImages, int D, int K,
                              const T* means, const GmmPrep<T>& prep, const T* priors, const T* rm, const T* rs,
                              const int* ra, int flags, T* enc, double* normpart, unsigned long long* numTerms,
                              int numKB) {
  if (resident)
    launch_fisher_accum<T, DPT, true>(x, offsets, numImages, D, K, means, prep, priors, rm, rs, ra, flags, enc,
                                      normpart, numTerms, numKB);
  else
    launch_fisher_accum<T, DPT, false>(x, offsets, numImages, D, K, means, prep, priors, rm, rs, ra, flags, enc,
                                       normpart, numTerms, numKB);
}

constexpr size_t kMaxLds = 160 * 1024;

template <typename T>
void fisher_fp_t(T* enc, const T* means, const T* covs, const T* priors, size_t D, size_t K, const T* x,
                 const uint64_t* offsets, size_t numImages, int flags, unsigned long long* numTerms) {
  VLHIP_REQUIRE(K >= 1 && D >= 1, "numClusters and dimension must be positive");
  if (numImages == 0) return;
  uint64_t total = 0;
  download(&total, offsets + numImages, sizeof(uint64_t));
  GmmPrep<T> prep = gmm_prepare<T>(priors, covs, D, K);
  unsigned char* rsb = (unsigned char*)scratch(SLOT_ROWSTAT, (size_t)(total + 1) * (2 * sizeof(T) + sizeof(int)) + 512);
  T* rm = (T*)rsb;
  T* rs = rm + total + 1;
  int* ra = (int*)(rs + total + 1);
  run_estep<T>(x, total, D, K, means, prep, nullptr, rm, rs, ra, "fisher_rowstats", false);

  const int numKB = (int)((K + FA_COMPS - 1) / FA_COMPS);
  double* normpart = (double*)scratch(SLOT_PARTIAL, numImages * (size_t)numKB * sizeof(double));
  if (numTerms) VLHIP_CHECK(hipMemsetAsync(numTerms, 0, numImages * sizeof(unsigned long long), ctx().stream));
  const bool resident = fisher_accum_lds_elems((int)D, true) * sizeof(T) + 2048 <= kMaxLds;
  VLHIP_REQUIRE(fisher_accum_lds_elems((int)D, false) * sizeof(T) + 2048 <= kMaxLds,
                "dimension too large for the GPU Fisher encoder's LDS tiles");
  const int d = (int)D, k = (int)K, n = (int)numImages;
  if (D <= 16)       dispatch_resident<T, 1>(resident, x, offsets, n, d, k, means, prep, priors, rm, rs, ra, flags, enc, normpart, numTerms, numKB);
  else if (D <= 32)  dispatch_resident<T, 2>(resident, x, offsets, n, d, k, means, prep, priors, rm, rs, ra, flags, enc, normpart, numTerms, numKB);
  else if (D <= 64)  dispatch_resident<T, 4>(resident, x, offsets, n, d, k, means, prep, priors, rm, rs, ra, flags, enc, normpart, numTerms, numKB);
  else if (D <= 80)  dispatch_resident<T, 5>(resident, x, offsets, n, d, k, means, prep, priors, rm, rs, ra, flags, enc, normpart, numTerms, numKB);
  else if (D <= 128) dispatch_resident<T, 8>(resident, x, offsets, n, d, k, means, prep, priors, rm, rs, ra, flags, enc, normpart, numTerms, numKB);
  else if (D <= 256) dispatch_resident<T, 16>(resident, x, offsets, n, d, k, means, prep, priors, rm, rs, ra, flags, enc, normpart, numTerms, numKB);
  else VLHIP_REQUIRE(false, "GPU Fisher encoder supports dimension <= 256");

  if (flags & 2) {
    VLHIP_LAUNCH("fisher_normalize", (fisher_normalize_kernel<T>), dim3((unsigned)numImages), dim3(256), 0, enc,
                 normpart, numKB, (size_t)2 * K * D);
  }
}

}  // namespace vlhip

using namespace vlhip;

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
                   (const float*)data, offsets, numImages, flags, numTerms, mode == VLHIP_MODE_F16X3);
  } else {
    fatal("vlhip_fisher_encode_batch", __FILE__, __LINE__, "unknown mode");
  }
}
