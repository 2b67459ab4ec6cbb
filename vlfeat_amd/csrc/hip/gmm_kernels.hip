// gmm_kernels.hip -- E-step, M-step moments, empty-mode split and one-hot
// posteriors of the diagonal GMM (vl/gmm.c).
//
// Data layout (as the reference, gmm.c:775, 1093-1149): descriptors x[i*D+d],
// means / covariances [k*D+d], priors [k], posteriors q[i*K+k], all row-major.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

namespace vlhip {

// ---------------------------------------------------------------------------
// Per-component terms, gmm.c:741-756 and fisher.c:386-390.  The log-prior and
// the log-determinant are accumulated in FP64:
//   cterm[k] = log pi_k - D/2 log(2 pi) - 1/2 sum_d log sigma2_kd
// (-inf when pi_k < 1e-6, gmm.c:746-747), inv[k,d] = 1/sigma2 in T
// (gmm.c:753), isig[k,d] = sqrt(1/sigma2) rounded from FP64 (fisher.c:388).
// One workgroup per component: the threads take the dimensions (log, 1/sigma2
// and sqrt(1/sigma2) of each), the logarithms pass through LDS in chunks and
// thread 0 adds them in dimension order -- the same FP64 sum as a serial loop.
constexpr int PREP_CHUNK = 1024;
template <typename T>
__global__ __launch_bounds__(64) void gmm_prep_kernel(const T* __restrict__ priors, const T* __restrict__ covs,
                                                      int D, int K, T* __restrict__ inv, T* __restrict__ isig,
                                                      double* __restrict__ cterm) {
  __shared__ double lg[PREP_CHUNK];
  const int k = blockIdx.x;
  if (k >= K) return;
  double logdet = 0.0;
  for (int d0 = 0; d0 < D; d0 += PREP_CHUNK) {
    const int n = min(PREP_CHUNK, D - d0);
    for (int i = threadIdx.x; i < n; i += blockDim.x) {
      const size_t at = (size_t)k * D + d0 + i;
      const T c = covs[at];
      lg[i] = log((double)c);
      inv[at] = T(1) / c;
      isig[at] = (T)sqrt(1.0 / (double)c);
    }
    __syncthreads();
    if (threadIdx.x == 0)
      for (int i = 0; i < n; ++i) logdet += lg[i];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    const double p = (double)priors[k];
    const double logw = (p < 1e-6) ? -(double)INFINITY : log(p);
    cterm[k] = logw - 0.5 * (double)D * kLog2Pi - 0.5 * logdet;
  }
}

// ---------------------------------------------------------------------------
// E-step, gmm.c:758-791.  One workgroup owns 16R descriptors (R = 4: EM_ROWS)
// and sweeps all K components in tiles of EM_COLS; every thread keeps an R x 4
// register tile of squared Mahalanobis distances, accumulated over D in
// LDS-staged chunks.  A row's arithmetic does not depend on R (same chunks,
// tiles and merge order), so R = 1 -- four times the workgroups -- serves
// small launches with the same posteriors bit for bit.  The
// row log-sum-exp is kept online per thread and merged across the 16 threads
// of a row at the end.  With WRITE_Q the logits are written to q during the
// sweep and normalised in place afterwards by the same thread that wrote them.
//
// Outputs per row (when the pointers are non-null): row_m = max logit,
// row_s = sum_k exp(logit - max), row_arg = first index of the max logit
// (the FAST rule of fisher.c:399-417).  ll_part[block] = sum over the block's
// rows of log(row_s) + row_m, in FP64 (gmm.c:786).
template <typename T, bool WRITE_Q, int R = 4>
__global__ __launch_bounds__(256) void estep_kernel(
    const T* __restrict__ x, long long N, int D, int K,
    const T* __restrict__ means, const T* __restrict__ inv, const double* __restrict__ cterm,
    T* __restrict__ q, T* __restrict__ row_m, T* __restrict__ row_s, int* __restrict__ row_arg,
    double* __restrict__ ll_part) {
  constexpr int ROWS = 16 * R;  // rows of the workgroup, R per thread
  __shared__ __align__(16) T xs[EM_DC][ROWS + 4];
  __shared__ __align__(16) T ms[EM_DC][EM_COLS + 4];
  __shared__ __align__(16) T ss[EM_DC][EM_COLS + 4];
  __shared__ T red_m[ROWS][17];
  __shared__ T red_s[ROWS][17];
  __shared__ T red_b[ROWS][16];
  __shared__ int red_a[ROWS][16];
  __shared__ double red_ll[256];

  const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
  const long long row0 = (long long)blockIdx.x * ROWS;

  T m_run[R], s_run[R], b_val[R];
  int b_idx[R];
#pragma unroll
  for (int i = 0; i < R; ++i) {
    m_run[i] = neg_inf<T>(); s_run[i] = T(0); b_val[i] = neg_inf<T>(); b_idx[i] = 0;
  }

  for (int k0 = 0; k0 < K; k0 += EM_COLS) {
    T acc[R][4];
#pragma unroll
    for (int i = 0; i < R; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = T(0);

    for (int d0 = 0; d0 < D; d0 += EM_DC) {
      const int dc = min(EM_DC, D - d0);
      __syncthreads();
      {
        // every load of the step issued together, then stored
        constexpr int XN = (ROWS * EM_DC + 255) / 256, MN = (EM_COLS * EM_DC + 255) / 256;
        T xv[XN], mv[MN], iv[MN];
#pragma unroll
        for (int u = 0; u < XN; ++u) {
          const int idx = tid + 256 * u;
          const int r = idx / EM_DC, dd = idx % EM_DC;
          const long long row = row0 + r;
          xv[u] = (idx < ROWS * EM_DC && dd < dc && row < N) ? x[row * D + d0 + dd] : T(0);
        }
#pragma unroll
        for (int u = 0; u < MN; ++u) {
          const int idx = tid + 256 * u;
          const int c = idx / EM_DC, dd = idx % EM_DC;
          const int k = k0 + c;
          const bool ok = (idx < EM_COLS * EM_DC && dd < dc && k < K);
          mv[u] = ok ? means[(size_t)k * D + d0 + dd] : T(0);
          iv[u] = ok ? inv[(size_t)k * D + d0 + dd] : T(0);
        }
#pragma unroll
        for (int u = 0; u < XN; ++u) {
          const int idx = tid + 256 * u;
          if (idx < ROWS * EM_DC) xs[idx % EM_DC][idx / EM_DC] = xv[u];
        }
#pragma unroll
        for (int u = 0; u < MN; ++u) {
          const int idx = tid + 256 * u;
          if (idx < EM_COLS * EM_DC) {
            ms[idx % EM_DC][idx / EM_DC] = mv[u];
            ss[idx % EM_DC][idx / EM_DC] = iv[u];
          }
        }
      }
      __syncthreads();
      for (int dd = 0; dd < dc; ++dd) {
        T xv[4], mv[4], sv[4];
        if constexpr (R == 4) lds4(&xs[dd][ty * R], xv); else xv[0] = xs[dd][ty];
        lds4(&ms[dd][tx * 4], mv);
        lds4(&ss[dd][tx * 4], sv);
#pragma unroll
        for (int i = 0; i < R; ++i)
#pragma unroll
          for (int j = 0; j < 4; ++j) {
            const T t = xv[i] - mv[j];
            acc[i][j] = fma(t * t, sv[j], acc[i][j]);
          }
      }
    }

    // logits and online log-sum-exp
#pragma unroll
    for (int i = 0; i < R; ++i) {
      const long long row = row0 + ty * R + i;
      T l[4];
      T lmax = neg_inf<T>();
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const int k = k0 + tx * 4 + j;
        l[j] = (k < K) ? (T)(cterm[k] - 0.5 * (double)acc[i][j]) : neg_inf<T>();
        lmax = l[j] > lmax ? l[j] : lmax;
        if (l[j] > b_val[i]) { b_val[i] = l[j]; b_idx[i] = k; }
      }
      const T mnew = lmax > m_run[i] ? lmax : m_run[i];
      if (mnew != neg_inf<T>()) {
        T s = s_run[i] * dev_exp<T>(m_run[i] - mnew);
#pragma unroll
        for (int j = 0; j < 4; ++j) s += dev_exp<T>(l[j] - mnew);
        s_run[i] = s;
        m_run[i] = mnew;
      }
      if (WRITE_Q && row < N) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = k0 + tx * 4 + j;
          if (k < K) q[row * K + k] = l[j];
        }
      }
    }
  }

  // merge the 16 per-thread partial statistics of each row
#pragma unroll
  for (int i = 0; i < R; ++i) {
    red_m[ty * R + i][tx] = m_run[i];
    red_s[ty * R + i][tx] = s_run[i];
    red_b[ty * R + i][tx] = b_val[i];
    red_a[ty * R + i][tx] = b_idx[i];
  }
  __syncthreads();
  double ll = 0.0;
  if (tid < ROWS) {
    const int r = tid;
    T M = neg_inf<T>();
    for (int t = 0; t < 16; ++t) M = red_m[r][t] > M ? red_m[r][t] : M;
    T S = T(0);
    if (M != neg_inf<T>()) {
      for (int t = 0; t < 16; ++t)
        if (red_m[r][t] != neg_inf<T>()) S += red_s[r][t] * dev_exp<T>(red_m[r][t] - M);
    }
    T bv = red_b[r][0];
    int bi = red_a[r][0];
    for (int t = 1; t < 16; ++t) {
      const T v = red_b[r][t];
      const int a = red_a[r][t];
      if (v > bv || (v == bv && a < bi)) { bv = v; bi = a; }
    }
    red_m[r][16] = M;
    red_s[r][16] = S;
    const long long row = row0 + r;
    if (row < N) {
      if (row_m) row_m[row] = M;
      if (row_s) row_s[row] = S;
      if (row_arg) row_arg[row] = bi;
      ll = log((double)S) + (double)M;
    }
  }
  const double bl = block_sum_256(ll, red_ll);
  if (tid == 0 && ll_part) ll_part[blockIdx.x] = bl;

  if (WRITE_Q) {
    // each thread rescales exactly the entries it wrote (no cross-thread reads)
    T M[R], S[R];
#pragma unroll
    for (int i = 0; i < R; ++i) { M[i] = red_m[ty * R + i][16]; S[i] = red_s[ty * R + i][16]; }
    for (int k0 = 0; k0 < K; k0 += EM_COLS) {
#pragma unroll
      for (int i = 0; i < R; ++i) {
        const long long row = row0 + ty * R + i;
        if (row >= N) continue;
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const int k = k0 + tx * 4 + j;
          if (k < K) {
            const long long idx = row * K + k;
            q[idx] = dev_exp<T>(q[idx] - M[i]) / S[i];
          }
        }
      }
    }
  }
}

// ---------------------------------------------------------------------------
// M-step moments, gmm.c:1093-1149.  Workgroup (kb, split) owns MS_COMPS
// components and the rows of one split; thread t owns component t/8 and the
// dimension pairs d = 2(t%8) + 16j, 2(t%8) + 16j + 1 (one 8-byte LDS read each).  Posteriors below the threshold (0.01/K,
// gmm.c:1099) contribute nothing, exactly as the reference's `continue`.
// Partial sums per split are written in T and reduced in FP64 afterwards.
template <typename T, int DPT>
__global__ __launch_bounds__(256) void moments_kernel(
    const T* __restrict__ q, const T* __restrict__ x, const T* __restrict__ oldMeans,
    long long N, int D, int K, long long rowsPerSplit, double thr,
    T* __restrict__ S0p, T* __restrict__ S1p, T* __restrict__ S2p, T* __restrict__ massp) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* qs = reinterpret_cast<T*>(smem_raw);                    // [MS_ROWS][MS_COMPS + 1]
  const int XR = D + (D & 1) + 2;                             // row stride, even
  T* xs = qs + MS_ROWS * (MS_COMPS + 1);                      // [MS_ROWS][XR]
  T* red = xs + (size_t)MS_ROWS * XR;                         // [2][256]

  const int tid = threadIdx.x;
  const int c = tid >> 3, g = tid & 7;             // compute ownership
  const int cs = tid & (MS_COMPS - 1), rs = tid >> 5;  // staging ownership
  const int k0 = blockIdx.x * MS_COMPS;
  const int split = blockIdx.y;
  const long long rBegin = (long long)split * rowsPerSplit;
  const long long rEnd = min(N, rBegin + rowsPerSplit);
  const int kc = k0 + c;

  static_assert(DPT % 2 == 0, "dimensions are owned in pairs");
  // element j of the thread: dimension 2g + 16 (j / 2) + (j % 2)
  auto dim_of = [&](int j) __attribute__((always_inline)) { return 2 * g + 16 * (j >> 1) + (j & 1); };
  T mu[DPT], s1[DPT], s2[DPT];
#pragma unroll
  for (int j = 0; j < DPT; ++j) {
    const int d = dim_of(j);
    mu[j] = (kc < K && d < D) ? oldMeans[(size_t)kc * D + d] : T(0);
    s1[j] = T(0);
    s2[j] = T(0);
  }
  T s0acc = T(0), massacc = T(0);

  for (long long r0 = rBegin; r0 < rEnd; r0 += MS_ROWS) {
    const int nr = (int)min((long long)MS_ROWS, rEnd - r0);
    __syncthreads();
    // stage posteriors (thresholded) and descriptors
#pragma unroll
    for (int j = 0; j < MS_ROWS / 8; ++j) {
      const int r = rs + 8 * j;
      const int k = k0 + cs;
      T p = T(0);
      if (r < nr && k < K) {
        p = q[(r0 + r) * K + k];
        massacc += p;
        if ((double)p < thr) p = T(0);
        s0acc += p;
      }
      qs[r * (MS_COMPS + 1) + cs] = p;
    }
    // descriptors: 8 loads per thread issued together, then stored
    for (int base = 0; base < nr * D; base += 256 * 8) {
      T v[8];
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = base + tid + 256 * u;
        v[u] = idx < nr * D ? x[r0 * D + idx] : T(0);
      }
#pragma unroll
      for (int u = 0; u < 8; ++u) {
        const int idx = base + tid + 256 * u;
        if (idx < nr * D) {
          const int r = idx / D;
          xs[r * XR + (idx - r * D)] = v[u];
        }
      }
    }
    __syncthreads();
    typedef T pair2 __attribute__((ext_vector_type(2)));
    for (int r = 0; r < nr; ++r) {
      const T p = qs[r * (MS_COMPS + 1) + c];
      if (p != T(0)) {
#pragma unroll
        for (int jj = 0; jj < DPT / 2; ++jj) {
          const int d0 = 2 * g + 16 * jj;
          if (d0 < D) {
            // xs rows hold at least D + 1 values, so the pair's second (padding) read stays in the row
            const pair2 xp = *reinterpret_cast<const pair2*>(&xs[r * XR + d0]);
#pragma unroll
            for (int h = 0; h < 2; ++h) {
              const int j = 2 * jj + h;
              if (d0 + h < D) {
                const T xv = xp[h];
                const T diff = xv - mu[j];
                s1[j] = fma(p, xv, s1[j]);
                s2[j] = fma(p * diff, diff, s2[j]);
              }
            }
          }
        }
      }
    }
  }

  if (kc < K) {
#pragma unroll
    for (int j = 0; j < DPT; ++j) {
      const int d = dim_of(j);
      if (d < D) {
        const size_t o = ((size_t)split * K + kc) * D + d;
        S1p[o] = s1[j];
        S2p[o] = s2[j];
      }
    }
  }
  // reduce the 8 staging partials of each component
  __syncthreads();
  red[tid] = s0acc;
  red[256 + tid] = massacc;
  __syncthreads();
  if (tid < MS_COMPS) {
    T a = T(0), b = T(0);
    for (int j = 0; j < 8; ++j) { a += red[tid + 32 * j]; b += red[256 + tid + 32 * j]; }
    const int k = k0 + tid;
    if (k < K) {
      S0p[(size_t)split * K + k] = a;
      massp[(size_t)split * K + k] = b;
    }
  }
}

// Fixed-order FP64 reduction of the per-split partials.
template <typename T>
__global__ void moments_reduce_kernel(const T* __restrict__ S0p, const T* __restrict__ S1p,
                                      const T* __restrict__ S2p, const T* __restrict__ massp,
                                      int splits, int D, int K, double* __restrict__ out) {
  // out layout: [S0 K][mass K][S1 K*D][S2 K*D]
  const size_t KD = (size_t)K * D;
  const size_t total = 2 * (size_t)K + 2 * KD;
  for (size_t i = blockIdx.x * (size_t)blockDim.x + threadIdx.x; i < total;
       i += (size_t)gridDim.x * blockDim.x) {
    double s = 0.0;
    if (i < (size_t)K) {
      for (int p = 0; p < splits; ++p) s += (double)S0p[(size_t)p * K + i];
    } else if (i < 2 * (size_t)K) {
      const size_t k = i - K;
      for (int p = 0; p < splits; ++p) s += (double)massp[(size_t)p * K + k];
    } else if (i < 2 * (size_t)K + KD) {
      const size_t e = i - 2 * (size_t)K;
      for (int p = 0; p < splits; ++p) s += (double)S1p[(size_t)p * KD + e];
    } else {
      const size_t e = i - 2 * (size_t)K - KD;
      for (int p = 0; p < splits; ++p) s += (double)S2p[(size_t)p * KD + e];
    }
    out[i] = s;
  }
}

// Empty-mode split, gmm.c:968-990.
template <typename T>
__global__ void split_kernel(T* __restrict__ q, const T* __restrict__ x, long long N, int D, int K,
                             int icl, int jcl, int dim, T mu) {
  const long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x;
  if (i >= N) return;
  const T p = q[i * K + jcl];
  const T qq = q[i * K + icl];
  if (x[i * D + dim] < mu) {
    q[i * K + icl] = p + qq;
    q[i * K + jcl] = T(0);
  } else {
    q[i * K + icl] = T(0);
    q[i * K + jcl] = p + qq;
  }
}

template <typename T>
__global__ void onehot_kernel(T* __restrict__ q, const uint32_t* __restrict__ a, long long N, int K) {
  const long long total = N * (long long)K;
  for (long long i = blockIdx.x * (long long)blockDim.x + threadIdx.x; i < total;
       i += (long long)gridDim.x * blockDim.x) {
    const long long row = i / K;
    const int k = (int)(i - row * K);
    q[i] = (a[row] == (uint32_t)k) ? T(1) : T(0);
  }
}

// ---------------------------------------------------------------------------
// Host-side launchers shared with the Fisher path.

template <typename T>
GmmPrep<T> gmm_prepare(const T* priors, const T* covs, size_t D, size_t K) {
  const size_t KD = K * D;
  // [inv KD][isig KD][cterm K doubles], each block 256-byte aligned
  const size_t bInv = (KD * sizeof(T) + 255) & ~size_t(255);
  const size_t bSig = bInv;
  const size_t bCt = (K * sizeof(double) + 255) & ~size_t(255);
  unsigned char* base = (unsigned char*)scratch(SLOT_PREP, bInv + bSig + bCt);
  GmmPrep<T> p;
  p.inv = (T*)base;
  p.isig = (T*)(base + bInv);
  p.cterm = (double*)(base + bInv + bSig);
  VLHIP_LAUNCH("gmm_prep", (gmm_prep_kernel<T>), dim3((unsigned)K), dim3(64), 0, priors, covs, (int)D, (int)K, p.inv,
               p.isig, p.cterm);
  return p;
}

template <typename T>
double run_estep(const T* x, size_t N, size_t D, size_t K, const T* means, const GmmPrep<T>& prep,
                 T* q, T* row_m, T* row_s, int* row_arg, const char* name, bool want_ll) {
  if (N == 0) return 0.0;
  const unsigned blocks = (unsigned)((N + EM_ROWS - 1) / EM_ROWS);
  const int cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
  if (!want_ll && blocks < (unsigned)(2 * cus)) {
    // few rows (one image of a drop-in call): 16-row workgroups; the
    // log-likelihood partials (grouped by workgroup) are not produced here
    const unsigned b1 = (unsigned)((N + 15) / 16);
    if (q)
      VLHIP_LAUNCH(name, (estep_kernel<T, true, 1>), dim3(b1), dim3(256), 0, x, (long long)N, (int)D, (int)K, means,
                   prep.inv, prep.cterm, q, row_m, row_s, row_arg, (double*)nullptr);
    else
      VLHIP_LAUNCH(name, (estep_kernel<T, false, 1>), dim3(b1), dim3(256), 0, x, (long long)N, (int)D, (int)K, means,
                   prep.inv, prep.cterm, q, row_m, row_s, row_arg, (double*)nullptr);
    return 0.0;
  }
  double* llp = want_ll ? (double*)scratch(SLOT_PARTIAL, (size_t)blocks * sizeof(double)) : nullptr;
  if (q) {
    VLHIP_LAUNCH(name, (estep_kernel<T, true>), dim3(blocks), dim3(256), 0, x, (long long)N, (int)D,
                 (int)K, means, prep.inv, prep.cterm, q, row_m, row_s, row_arg, llp);
  } else {
    VLHIP_LAUNCH(name, (estep_kernel<T, false>), dim3(blocks), dim3(256), 0, x, (long long)N, (int)D,
                 (int)K, means, prep.inv, prep.cterm, q, row_m, row_s, row_arg, llp);
  }
  if (!want_ll) return 0.0;
  std::vector<double> h(blocks);
  download(h.data(), llp, blocks * sizeof(double));
  double ll = 0.0;
  for (unsigned b = 0; b < blocks; ++b) ll += h[b];
  return ll;
}

template <typename T>
double gmm_posteriors_t(T* q, size_t K, size_t N, const T* priors, const T* means, size_t D,
                        const T* covs, const T* x) {
  VLHIP_REQUIRE(K >= 1 && D >= 1, "numClusters and dimension must be positive");
  GmmPrep<T> prep = gmm_prepare<T>(priors, covs, D, K);
  return run_estep<T>(x, N, D, K, means, prep, q, nullptr, nullptr, nullptr, "gmm_estep", true);
}

template <typename T, int DPT>
static void launch_moments(const T* q, const T* x, const T* oldMeans, size_t N, size_t D, size_t K,
                           double thr, double* out_host) {
  const int numKB = (int)((K + MS_COMPS - 1) / MS_COMPS);
  const int cus = ctx().num_cus > 0 ? ctx().num_cus : 256;
  long long maxSplits = (long long)((N + MS_ROWS - 1) / MS_ROWS);
  long long want = (4LL * cus + numKB - 1) / numKB;
  long long G = std::max(1LL, std::min(maxSplits, want));
  long long rowsPerSplit = (long long)((N + G - 1) / G);
  rowsPerSplit = ((rowsPerSplit + MS_ROWS - 1) / MS_ROWS) * MS_ROWS;
  G = (long long)((N + rowsPerSplit - 1) / rowsPerSplit);
  if (G < 1) G = 1;
  const size_t KD = K * D;
  T* part = (T*)scratch(SLOT_PARTIAL, (size_t)G * (2 * K + 2 * KD) * sizeof(T));
  T* S0p = part;
  T* massp = S0p + (size_t)G * K;
  T* S1p = massp + (size_t)G * K;
  T* S2p = S1p + (size_t)G * KD;
  const size_t shmem = (size_t)(MS_ROWS * (MS_COMPS + 1) + MS_ROWS * (D + (D & 1) + 2) + 512) * sizeof(T);
  if (N > 0) {
    VLHIP_LAUNCH("gmm_moments", (moments_kernel<T, DPT>), dim3(numKB, (unsigned)G), dim3(256), shmem,
                 q, x, oldMeans, (long long)N, (int)D, (int)K, rowsPerSplit, thr, S0p, S1p, S2p, massp);
  } else {
    G = 0;
  }
  double* out = (double*)scratch(SLOT_REDUCED, (2 * K + 2 * KD) * sizeof(double));
  const size_t total = 2 * K + 2 * KD;
  VLHIP_LAUNCH("gmm_moments_reduce", (moments_reduce_kernel<T>),
               dim3((unsigned)std::min<size_t>((total + 255) / 256, 4096)), dim3(256), 0,
               S0p, S1p, S2p, massp, (int)G, (int)D, (int)K, out);
  download(out_host, out, total * sizeof(double));
}

template <typename T>
void gmm_moments_t(double* out_host, const T* q, const T* x, const T* oldMeans, size_t N, size_t D,
                   size_t K, double thr) {
  VLHIP_REQUIRE(D <= 256, "GPU M-step supports dimension <= 256");
  if (D <= 64) launch_moments<T, 8>(q, x, oldMeans, N, D, K, thr, out_host);
  else if (D <= 80) launch_moments<T, 10>(q, x, oldMeans, N, D, K, thr, out_host);
  else if (D <= 128) launch_moments<T, 16>(q, x, oldMeans, N, D, K, thr, out_host);
  else launch_moments<T, 32>(q, x, oldMeans, N, D, K, thr, out_host);
}

template GmmPrep<float> gmm_prepare<float>(const float*, const float*, size_t, size_t);
template GmmPrep<double> gmm_prepare<double>(const double*, const double*, size_t, size_t);
template double run_estep<float>(const float*, size_t, size_t, size_t, const float*, const GmmPrep<float>&,
                                 float*, float*, float*, int*, const char*, bool);
template double run_estep<double>(const double*, size_t, size_t, size_t, const double*,
                                  const GmmPrep<double>&, double*, double*, double*, int*, const char*, bool);

}  // namespace vlhip

using namespace vlhip;

extern "C" {

double vlhip_gmm_posteriors(int dtype, void* posteriors, size_t K, size_t N, void const* priors,
                            void const* means, size_t D, void const* covs, void const* data) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (dtype == 1)
    return gmm_posteriors_t<float>((float*)posteriors, K, N, (const float*)priors, (const float*)means, D,
                                   (const float*)covs, (const float*)data);
  if (dtype == 2)
    return gmm_posteriors_t<double>((double*)posteriors, K, N, (const double*)priors, (const double*)means,
                                    D, (const double*)covs, (const double*)data);
  fatal("vlhip_gmm_posteriors", __FILE__, __LINE__, "unknown data type");
}

void vlhip_gmm_moments(int dtype, double* out_host, void const* posteriors, void const* data,
                       void const* oldMeans, size_t N, size_t D, size_t K, double threshold) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (dtype == 1)
    gmm_moments_t<float>(out_host, (const float*)posteriors, (const float*)data, (const float*)oldMeans, N, D,
                         K, threshold);
  else if (dtype == 2)
    gmm_moments_t<double>(out_host, (const double*)posteriors, (const double*)data, (const double*)oldMeans,
                          N, D, K, threshold);
  else
    fatal("vlhip_gmm_moments", __FILE__, __LINE__, "unknown data type");
}

void vlhip_gmm_split(int dtype, void* posteriors, void const* data, size_t N, size_t D, size_t K, size_t iCl,
                     size_t jCl, size_t dim, double mu) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (N == 0) return;
  const unsigned blocks = (unsigned)((N + 255) / 256);
  if (dtype == 1)
    VLHIP_LAUNCH("gmm_split", (split_kernel<float>), dim3(blocks), dim3(256), 0, (float*)posteriors,
                 (const float*)data, (long long)N, (int)D, (int)K, (int)iCl, (int)jCl, (int)dim, (float)mu);
  else if (dtype == 2)
    VLHIP_LAUNCH("gmm_split", (split_kernel<double>), dim3(blocks), dim3(256), 0, (double*)posteriors,
                 (const double*)data, (long long)N, (int)D, (int)K, (int)iCl, (int)jCl, (int)dim, mu);
  else
    fatal("vlhip_gmm_split", __FILE__, __LINE__, "unknown data type");
}

void vlhip_gmm_onehot(int dtype, void* posteriors, uint32_t const* assignments, size_t N, size_t K) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (N == 0) return;
  const size_t total = N * K;
  const unsigned blocks = (unsigned)std::min<size_t>((total + 255) / 256, 65536);
  if (dtype == 1)
    VLHIP_LAUNCH("gmm_onehot", (onehot_kernel<float>), dim3(blocks), dim3(256), 0, (float*)posteriors,
                 assignments, (long long)N, (int)K);
  else if (dtype == 2)
    VLHIP_LAUNCH("gmm_onehot", (onehot_kernel<double>), dim3(blocks), dim3(256), 0, (double*)posteriors,
                 assignments, (long long)N, (int)K);
  else
    fatal("vlhip_gmm_onehot", __FILE__, __LINE__, "unknown data type");
}

}  // extern "C"
