// vlad_kernels.hip -- VLAD encoder (vl/vlad.c:163-242).
//
// For each cluster k and dimension d, enc[k,d] = sum_i q_ik x_id over the
// image's descriptors in index order, with the reference's arithmetic: the
// product and sum are formed in double and the running sum is rounded back to
// the data type after every term (vlad.c:181-188, where q is double and enc is
// TYPE), and only q > 0 contributes.  The per-cluster corrections, signed
// square root and per-cluster L2 normalisation follow vlad.c:191-227 in the
// same order and precision; the global normalisation (vlad.c:230-241) gives
// the reference's sequential sum of the squared entries, so for float data the
// result matches the reference's float path element for element.  Dense
// assignments: vlad_slice_kernel; hard assignments (cluster indices, D <= 128):
// vlad_hard_kernel.
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip_kernels.h"
#include "vlhip.h"

#include <algorithm>

namespace vlhip {

// Dense (soft) assignments, and hard ones with D > 128: workgroup (image,
// block of VS_COMPS clusters, slice of VS_DIMS dimensions); thread (c, l) owns
// one (cluster, dimension) and adds q_ik x_id over the image's descriptors in
// index order (vlad.c:181-188: the product and sum formed in double, the sum
// rounded back to the data type after every term, only q > 0 counted).  One
// chain per thread and the dimensions spread over workgroups, so that one
// image fills K D / 256 workgroups.  Descriptors and assignments are staged
// through LDS VS_ROWS rows at a time, the next chunk's loads in flight while
// the current one is added up.  Mean subtraction and signed square root
// (vlad.c:191-213) follow per element; the per-cluster normalisation
// (vlad.c:215-227) needs every dimension of a cluster and runs afterwards
// (vlad_component_norm_kernel).
constexpr int VS_COMPS = 16;
constexpr int VS_DIMS = 16;
constexpr int VS_ROWS = 256;

template <typename T, bool HARD>
__global__ __launch_bounds__(256) void vlad_slice_kernel(
    const T* __restrict__ x, const uint64_t* __restrict__ offsets, int numImages, int D, int K,
    const T* __restrict__ means, const T* __restrict__ assign, const uint32_t* __restrict__ hard,
    int flags, T* __restrict__ enc, int numKB, int numDS) {
#pragma clang fp contract(off)
  __shared__ T xs[VS_ROWS][VS_DIMS + 1];
  __shared__ T qs[VS_ROWS][VS_COMPS + 1];
  const int per = numKB * numDS;
  const int b = blockIdx.x;
  int img, rem;
  if (numImages >= 8) {
    const int xcd = b & 7, slot = b >> 3;
    img = (slot / per) * 8 + xcd;
    rem = slot % per;
  } else {
    img = b / per;
    rem = b % per;
  }
  if (img >= numImages) return;
  const int kb = rem / numDS, ds = rem % numDS;
  const int tid = threadIdx.x, c = tid >> 4, l = tid & 15;
  const int k = kb * VS_COMPS + c, d = ds * VS_DIMS + l;
  const long long rBegin = (long long)offsets[img], rEnd = (long long)offsets[img + 1];
  constexpr int PF = VS_ROWS * 16 / 256;  // staged values of each kind per thread
  T px[PF], pq[PF];
  auto load_chunk = [&](long long r0, int nr) {
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int idx = tid + 256 * u;
      const int r = idx >> 4, j = idx & 15;
      const int dd = ds * VS_DIMS + j, kk = kb * VS_COMPS + j;
      px[u] = (r < nr && dd < D) ? x[(r0 + r) * D + dd] : T(0);
      T qv = T(0);
      if (r < nr && kk < K) {
        if (HARD) qv = (hard[r0 + r] == (uint32_t)kk) ? T(1) : T(0);
        else qv = assign[(r0 + r) * K + kk];
      }
      pq[u] = qv;
    }
  };
  T acc = T(0);
  double mass = 0.0;
  if (rBegin < rEnd) load_chunk(rBegin, (int)min((long long)VS_ROWS, rEnd - rBegin));
  for (long long r0 = rBegin; r0 < rEnd; r0 += VS_ROWS) {
    const int nr = (int)min((long long)VS_ROWS, rEnd - r0);
    __syncthreads();
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int idx = tid + 256 * u;
      xs[idx >> 4][idx & 15] = px[u];
      qs[idx >> 4][idx & 15] = pq[u];
    }
    __syncthreads();
    if (r0 + VS_ROWS < rEnd) load_chunk(r0 + VS_ROWS, (int)min((long long)VS_ROWS, rEnd - r0 - VS_ROWS));
    if (k < K && d < D) {
      for (int r = 0; r < nr; ++r) {
        const T qv = qs[r][c];
        if (qv > T(0)) {
          const double qd = (double)qv;
          mass += qd;
          acc = (T)((double)acc + qd * (double)xs[r][l]);
        }
      }
    }
  }
  if (k < K && d < D) {
    T e = acc;
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
    enc[(size_t)img * K * D + (size_t)k * D + d] = e;
  }
}

// vlad.c:215-227: each cluster's D values divided by max(||.||, 1e-12), the
// squared norm summed over d in order by one lane.  Workgroup per (image,
// cluster), in encoding order.  Clusters of up to VN_STAGE values are staged
// in LDS; longer ones are summed and divided straight from global memory
// (same additions in the same order).
constexpr int VN_STAGE = 512;
template <typename T>
__global__ __launch_bounds__(64) void vlad_component_norm_kernel(T* __restrict__ enc, int D, int K) {
#pragma clang fp contract(off)
  __shared__ T z[VN_STAGE];
  __shared__ T nsh;
  T* e = enc + (size_t)blockIdx.x * D;
  const bool staged = D <= VN_STAGE;
  if (staged)
    for (int i = threadIdx.x; i < D; i += 64) z[i] = e[i];
  __syncthreads();
  if (threadIdx.x == 0) {
    T n = T(0);
    if (staged) {
      for (int i = 0; i < D; ++i) n = n + z[i] * z[i];
    } else {
      for (int i = 0; i < D; ++i) {
        const T v = e[i];
        n = n + v * v;
      }
    }
    n = dev_sqrt<T>(n);
    nsh = (T)(((double)n > 1e-12) ? (double)n : 1e-12);
  }
  __syncthreads();
  const T n = nsh;
  for (int i = threadIdx.x; i < D; i += 64) e[i] = (staged ? z[i] : e[i]) / n;
}

// vlad.c:230-241: enc /= max(||enc||, 1e-12), the squared norm summed in
// index order (bitwise the reference's sequential loop).  The squares z*z
// (each rounded as the reference rounds it) pass through two LDS buffers.
// Float sums take the exact block form below (vlad_exact_chain4) on all four
// waves, the next buffer's loads in flight meanwhile; double sums run as one
// lane's chain of dependent additions, fed by 16-byte LDS reads issued 8
// ahead, while waves 1-3 fill the other buffer.  The division then reads the
// encoding again with independent loads.
template <typename T>
struct VladNorm {
  static constexpr int CH = 16384 / sizeof(T);  // values per LDS buffer
  static constexpr int V = 16 / sizeof(T);      // values per 16-byte LDS read
};

// Threads t = 0..NT-1 of the workgroup write the squares of e[0..m) into sq:
// CH / NT values each, every load issued before the first square is stored.
template <typename T, int NT>
__device__ __forceinline__ void vlad_norm_stage(const T* __restrict__ e, int m, T* __restrict__ sq, int t) {
#pragma clang fp contract(off)
  constexpr int PER = (VladNorm<T>::CH + NT - 1) / NT;
  T v[PER];
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = t + NT * j;
    v[j] = i < m ? e[i] : T(0);
  }
#pragma unroll
  for (int j = 0; j < PER; ++j) {
    const int i = t + NT * j;
    if (i < VladNorm<T>::CH) sq[i] = v[j] * v[j];
  }
}

// n + sq[0] + sq[1] + ... + sq[m-1], in that order, one lane.
template <typename T>
__device__ __forceinline__ T vlad_norm_chain(T n, const T* __restrict__ sq, int m) {
#pragma clang fp contract(off)
  constexpr int V = VladNorm<T>::V, G = 8;
  typedef T vec __attribute__((ext_vector_type(V)));
  const vec* q = reinterpret_cast<const vec*>(sq);
  const int mv = m / V;
  const int full = (mv / (2 * G)) * (2 * G);
  vec a[G], b[G];
  if (full > 0) {
#pragma unroll
    for (int j = 0; j < G; ++j) a[j] = q[j];
  }
  for (int i = 0; i < full; i += 2 * G) {
#pragma unroll
    for (int j = 0; j < G; ++j) b[j] = q[i + G + j];
#pragma unroll
    for (int j = 0; j < G; ++j)
#pragma unroll
      for (int t = 0; t < V; ++t) n = n + a[j][t];
    if (i + 2 * G < full) {
#pragma unroll
      for (int j = 0; j < G; ++j) a[j] = q[i + 2 * G + j];
    }
#pragma unroll
    for (int j = 0; j < G; ++j)
#pragma unroll
      for (int t = 0; t < V; ++t) n = n + b[j][t];
  }
  for (int i = full * V; i < m; ++i) n = n + sq[i];
  return n;
}

// The same sum for float terms q >= 0 (or inf / NaN), bit for bit, a block
// of terms at a time.  While the running sum S stays
// below 2^(e+1), where 2^e <= S (e = -126 for S below the normal range),
// every float it takes is a multiple of u = 2^(e-23), and adding a term q
// rounds S + q to S + u * k with k = q/u rounded to the nearest integer --
// independent of S, unless q/u is an exact half (a tie, broken towards the
// even result, which depends on S).  So the sequential sum over a block equals
// S + u * (sum of the k) when no term ties, no term is inf or NaN or
// >= 2^(e+1), and the integer total stays below 2^24 (S remains in the
// binade).  Otherwise the prefix before the first term that breaks one of
// these is added the same way, that term is added as the reference adds it,
// and the next block starts after it (the rare cases: a tie, and the few
// times S crosses a power of two).  vlad_exact_chain4 below runs this.
constexpr int VN_J = 8;

// Inclusive prefix sum over the 64 lanes of a wave: within each row of 16
// lanes by shifts of 1, 2, 4 and 8, then row 15's total into rows 1 and 3 and
// lane 31's into rows 2 and 3 (DPP row_shr / row_bcast; lanes without a
// source add 0).
__device__ __forceinline__ unsigned vn_wave_scan(unsigned v) {
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x111, 0xf, 0xf, false);
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x112, 0xf, 0xf, false);
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x114, 0xf, 0xf, false);
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x118, 0xf, 0xf, false);
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x142, 0xa, 0xf, false);
  v += (unsigned)__builtin_amdgcn_update_dpp(0, (int)v, 0x143, 0xc, 0xf, false);
  return v;
}

// The block form above on the four waves of a workgroup: each step takes
// 4 x 512 terms, wave w the w-th 512 (8 per lane).  The waves exchange their
// integer totals through LDS; each then looks for its first term that cannot
// take the block path, counting from the totals of the waves before it, and
// the first such term of the step (if any) ends the step: the prefix before
// it is added as a block, the term as the reference adds it, and the next
// step starts after it.  The first `head` terms are added one by one (at the
// start of the sum S is no larger than single terms and would leave its
// binade every few terms).  Every thread computes the same S.  Slots of the
// exchange alternate between steps, so that two barriers per step suffice.
struct VnExchange {
  unsigned tot[2][4];
  int stop[2][4];
  unsigned before[2][4];
};
__device__ __forceinline__ float vlad_exact_chain4(float S, const float* __restrict__ sq, int m, int head, int tid,
                                                   VnExchange& x, int& step) {
#pragma clang fp contract(off)
  const int lane = tid & 63, w = tid >> 6;
#pragma unroll 8
  for (int i = 0; i < head; ++i) S = S + sq[i];
  int p = head;
  while (p < m) {
    if (!(S < INFINITY)) {
      bool nan = S != S;
      for (int i = p + tid; i < m; i += 256) nan |= sq[i] != sq[i];
      return __syncthreads_or(nan) ? __builtin_nanf("") : S;
    }
    const int par = step & 1;
    ++step;
    int e = S == 0.f ? -126 : __builtin_amdgcn_frexp_expf(S) - 1;
    if (e < -126) e = -126;
    const float lim = ldexpf(1.f, e + 1);
    const unsigned Sm = (unsigned)ldexpf(S, 23 - e);
    unsigned k[VN_J];
    int fj = VN_J;
#pragma unroll
    for (int j = 0; j < VN_J; ++j) {
      const int i = p + 512 * w + lane * VN_J + j;
      const float q = i < m ? sq[i] : 0.f;
      const bool ok = q < lim;
      const float sc = ldexpf(ok ? q : 0.f, 23 - e);
      const float fl = floorf(sc);
      const float fr = sc - fl;
      k[j] = (unsigned)fl + (fr > 0.5f ? 1u : 0u);
      if ((!ok || fr == 0.5f) && fj == VN_J) fj = j;
    }
    unsigned L = 0;
#pragma unroll
    for (int j = 0; j < VN_J; ++j) L += k[j];
    const unsigned Lc = L < (1u << 25) ? L : (1u << 25);
    const unsigned incl = vn_wave_scan(Lc);
    const unsigned excl = incl - Lc;
    const unsigned wt = (unsigned)__builtin_amdgcn_readlane((int)incl, 63);
    if (lane == 0) x.tot[par][w] = wt < (1u << 25) ? wt : (1u << 25);
    __syncthreads();
    unsigned base = 0;
    for (int v = 0; v < w; ++v) base += x.tot[par][v];
    unsigned run = Sm + (base + excl < (1u << 25) ? base + excl : (1u << 25));
    int f = fj;
    unsigned pre = 0;
#pragma unroll
    for (int j = 0; j < VN_J; ++j) {
      if (j < f && run + k[j] >= (1u << 24)) f = j;
      if (j < f) {
        run += k[j];
        pre += k[j];
      }
    }
    const unsigned long long stop = __ballot(f < VN_J);
    const int sl = stop ? __builtin_ctzll(stop) : 0;
    const int fsl = __builtin_amdgcn_readlane(f, sl);
    const unsigned bsl = (unsigned)__builtin_amdgcn_readlane((int)(base + excl + pre), sl);
    if (lane == 0) {
      x.stop[par][w] = stop ? sl * VN_J + fsl : 512;
      x.before[par][w] = bsl;
    }
    __syncthreads();
    int W = 4;
    for (int v = 3; v >= 0; --v)
      if (x.stop[par][v] < 512) W = v;
    if (W == 4) {
      unsigned tot = 0;
      for (int v = 0; v < 4; ++v) tot += x.tot[par][v];
      S = ldexpf((float)(Sm + tot), e - 23);
      p += 4 * 512;
    } else {
      S = ldexpf((float)(Sm + x.before[par][W]), e - 23);
      const int fi = p + 512 * W + x.stop[par][W];
      S = S + sq[fi];
      p = fi + 1;
    }
  }
  return S;
}

// Loads of vlad_norm_stage (issued) and its stores (squares), apart, so that
// the next chunk's loads are in flight while the current one is summed.
template <typename T, int NT>
struct VnPrefetch {
  static constexpr int PER = (VladNorm<T>::CH + NT - 1) / NT;
  T v[PER];
  __device__ __forceinline__ void load(const T* __restrict__ e, int m, int t) {
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int i = t + NT * j;
      v[j] = i < m ? e[i] : T(0);
    }
  }
  __device__ __forceinline__ void store(T* __restrict__ sq, int t) const {
#pragma clang fp contract(off)
#pragma unroll
    for (int j = 0; j < PER; ++j) {
      const int i = t + NT * j;
      if (i < VladNorm<T>::CH) sq[i] = v[j] * v[j];
    }
  }
};

template <typename T>
__global__ __launch_bounds__(256) void vlad_normalize_kernel(T* __restrict__ enc, size_t len) {
#pragma clang fp contract(off)
  constexpr int CH = VladNorm<T>::CH;
  __shared__ __align__(16) T sq[2][CH];
  __shared__ T nsh;
  T* e = enc + (size_t)blockIdx.x * len;
  const int tid = threadIdx.x;
  const size_t nch = (len + CH - 1) / CH;
  vlad_norm_stage<T, 256>(e, (int)min((size_t)CH, len), sq[0], tid);
  __syncthreads();
  T n = T(0);
  if constexpr (sizeof(T) == 4) {
    // all four waves sum each chunk; the next chunk's loads are in flight meanwhile
    __shared__ VnExchange xch;
    int step = 0;
    VnPrefetch<T, 256> pf;
    for (size_t c = 0; c < nch; ++c) {
      const int m = (int)min((size_t)CH, len - c * CH);
      if (c + 1 < nch) pf.load(e + (c + 1) * CH, (int)min((size_t)CH, len - (c + 1) * CH), tid);
      n = vlad_exact_chain4(n, reinterpret_cast<const float*>(sq[c & 1]), m, c == 0 ? min(m, 1024) : 0, tid, xch,
                            step);
      // sq[(c + 1) & 1] was last read while summing chunk c - 1, before the
      // barriers of chunk c's steps (chunk c >= 1 has at least one step)
      if (c + 1 < nch) pf.store(sq[(c + 1) & 1], tid);
      __syncthreads();
    }
  } else {
    for (size_t c = 0; c < nch; ++c) {
      if (tid < 64) {
        if (tid == 0) n = vlad_norm_chain<T>(n, sq[c & 1], (int)min((size_t)CH, len - c * CH));
      } else if (c + 1 < nch) {
        const size_t c0 = (c + 1) * CH;
        vlad_norm_stage<T, 192>(e + c0, (int)min((size_t)CH, len - c0), sq[(c + 1) & 1], tid - 64);
      }
      __syncthreads();
    }
  }
  if (tid == 0) {
    n = dev_sqrt<T>(n);
    nsh = (T)(((double)n > 1e-12) ? (double)n : 1e-12);
  }
  __syncthreads();
  const T nn = nsh;
  constexpr int U = 16;
  for (size_t i0 = 0; i0 < len; i0 += (size_t)256 * U) {
    T v[U];
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const size_t i = i0 + tid + 256 * j;
      v[j] = i < len ? e[i] : T(0);
    }
#pragma unroll
    for (int j = 0; j < U; ++j) {
      const size_t i = i0 + tid + 256 * j;
      if (i < len) e[i] = v[j] / nn;
    }
  }
}

// Hard assignments (one cluster per descriptor): workgroup (image, block of
// CB clusters), thread d owns dimension d of every cluster of the block and
// adds the block's descriptors in index order -- the same sequence of
// roundings per (cluster, dimension) as vlad_slice_kernel with a one-hot q
// (vlad.c:181-188), without scanning every cluster for every row.  With
// q = 1 the reference's step enc = (TYPE)(enc + q * x), formed in double, is
// the plain TYPE sum: for float, rounding the exact double sum of two floats
// to float gives the correctly rounded float sum (53 >= 2 * 24 + 2 bits, so
// the double rounding is innocuous for addition).
// Per chunk of VH_CHUNK rows: the cluster indices are staged in LDS and each
// wave compacts the rows of its half that fall in the block into an LDS list
// (ballots, index order kept) and counts them per cluster (the mass: the
// reference adds q = 1 in double per row, which is the count exactly).  For
// blocks of up to 8 clusters the list is sorted by cluster (index order
// within a cluster), so that each cluster's sum is a plain chain of adds
// into a register; larger blocks add in index order into LDS accumulators.
// The listed rows are gathered into LDS in stages, every 16-byte load of a
// stage in flight together.  Real assignments are skewed (a few clusters
// hold many rows), so the workgroup of the largest clusters sets the time:
// it pays one memory round trip per stage, not per row.
// Finalisation as vlad_slice_kernel and vlad_component_norm_kernel
// (vlad.c:191-227), in one workgroup.  D <= 128, CB <= 256.
// The host picks CB so that even one image spreads over enough workgroups.
constexpr int VH_CHUNK = 2048;

// Stage rows s0.. of the block's list (list0[0..n0) then list1) into xs as
// pieces of P (16 bytes, or one value): 16 loads per thread issued together.
template <typename T, typename P>
__device__ __forceinline__ void vh_gather(const T* __restrict__ x, long long c0, int D, const uint32_t* list0, int n0,
                                          const uint32_t* list1, int s0, int npieces, T* xs, int tid) {
  constexpr int GU = 16;
  const int DW = D / (int)(sizeof(P) / sizeof(T));
  for (int q0 = 0; q0 < npieces; q0 += 128 * GU) {
    P v[GU];
#pragma unroll
    for (int g = 0; g < GU; ++g) {
      const int q = q0 + tid + 128 * g;
      if (q < npieces) {
        const int r = q / DW, c = q - r * DW;
        const int p = s0 + r;
        const uint32_t ent = p < n0 ? list0[p] : list1[p - n0];
        v[g] = reinterpret_cast<const P*>(x + (c0 + (long long)(ent >> 8)) * D)[c];
      }
    }
#pragma unroll
    for (int g = 0; g < GU; ++g) {
      const int q = q0 + tid + 128 * g;
      if (q < npieces) reinterpret_cast<P*>(xs)[q] = v[g];
    }
  }
}

template <typename T, int CB>
__global__ __launch_bounds__(128) void vlad_hard_kernel(const T* __restrict__ x, const uint64_t* __restrict__ offsets,
                                                        int numImages, int D, int K, const T* __restrict__ means,
                                                        const uint32_t* __restrict__ hard, int flags,
                                                        T* __restrict__ enc, int numKB) {
#pragma clang fp contract(off)
  static_assert(CB <= 256, "cluster block must fit the 8-bit list field");
  constexpr bool REG = CB <= 8;                          // accumulate in registers
  constexpr int XS = (REG ? 32768 : 8192) / sizeof(T);  // staged values
  typedef unsigned int piece16 __attribute__((ext_vector_type(4)));
  extern __shared__ __align__(16) unsigned char smem_raw[];
  T* acc = reinterpret_cast<T*>(smem_raw);                             // [CB][D]
  double* mass = reinterpret_cast<double*>(smem_raw + (((size_t)CB * D * sizeof(T) + 15) & ~size_t(15)));
  T* nrm = reinterpret_cast<T*>(mass + CB);                           // [CB]
  __shared__ uint32_t idxs[VH_CHUNK];
  __shared__ uint32_t list[2][VH_CHUNK / 2];  // (row in chunk) << 8 | cluster in block
  uint32_t* const lflat = &list[0][0];        // REG: one sorted list of VH_CHUNK
  __shared__ int cnt[2];
  __shared__ unsigned counts[CB];
  __shared__ int wcnt[2][REG ? CB : 1];  // rows per (wave, cluster) of the chunk
  __shared__ __align__(16) T xs[XS];
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
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int k0 = kb * CB;
  if (!REG)
    for (int i = tid; i < CB * D; i += blockDim.x) acc[i] = T(0);
  for (int i = tid; i < CB; i += blockDim.x) counts[i] = 0u;
  T a[REG ? CB : 1];
#pragma unroll
  for (int c = 0; c < (REG ? CB : 1); ++c) a[c] = T(0);
  const long long rBegin = (long long)offsets[img], rEnd = (long long)offsets[img + 1];
  const int d = tid;
  const bool vec16 = (D * (int)sizeof(T)) % 16 == 0 && ((uintptr_t)x & 15) == 0;
  const int SR = XS / D;  // rows per stage
  const unsigned long long lt = (1ull << lane) - 1;
  for (long long c0 = rBegin; c0 < rEnd; c0 += VH_CHUNK) {
    const int nr = (int)min((long long)VH_CHUNK, rEnd - c0);
    __syncthreads();  // the previous chunk's lists are consumed
    {
      constexpr int PER = VH_CHUNK / 128;
      uint32_t v[PER];
#pragma unroll
      for (int j = 0; j < PER; ++j) {
        const int i = tid + 128 * j;
        v[j] = i < nr ? hard[c0 + i] : 0xffffffffu;
      }
#pragma unroll
      for (int j = 0; j < PER; ++j) idxs[tid + 128 * j] = v[j];
    }
    __syncthreads();
    int cst[REG ? CB + 1 : 1];  // cluster c's rows: sorted list [cst[c], cst[c+1])
    int ntot;
    if constexpr (REG) {
      // counts per (wave, cluster), then each wave writes its rows of cluster
      // c after the rows of c of the clusters before and of wave 0
      int m[CB];
#pragma unroll
      for (int c = 0; c < CB; ++c) m[c] = 0;
      for (int g = 0; g < VH_CHUNK / 128; ++g) {
        const uint32_t kk = idxs[wave * (VH_CHUNK / 2) + g * 64 + lane];
        const int kl = (int)kk - k0;
        const bool in = kk != 0xffffffffu && kl >= 0 && kl < CB;
#pragma unroll
        for (int c = 0; c < CB; ++c) m[c] += __popcll(__ballot(in && kl == c));
      }
      if (lane == 0) {
#pragma unroll
        for (int c = 0; c < CB; ++c) wcnt[wave][c] = m[c];
      }
      __syncthreads();
      int base[CB];
      cst[0] = 0;
#pragma unroll
      for (int c = 0; c < CB; ++c) {
        const int w0 = wcnt[0][c], w1 = wcnt[1][c];
        base[c] = cst[c] + (wave ? w0 : 0);
        cst[c + 1] = cst[c] + w0 + w1;
      }
      ntot = cst[CB];
      for (int g = 0; g < VH_CHUNK / 128; ++g) {
        const int r = wave * (VH_CHUNK / 2) + g * 64 + lane;
        const uint32_t kk = idxs[r];
        const int kl = (int)kk - k0;
        const bool in = kk != 0xffffffffu && kl >= 0 && kl < CB;
#pragma unroll
        for (int c = 0; c < CB; ++c) {
          const unsigned long long mk = __ballot(in && kl == c);
          if (in && kl == c) lflat[base[c] + __popcll(mk & lt)] = (uint32_t)r << 8 | (uint32_t)c;
          base[c] += __popcll(mk);
        }
      }
      if (tid < CB) counts[tid] += (unsigned)(wcnt[0][tid] + wcnt[1][tid]);
    } else {
      int n = 0;
      for (int g = 0; g < VH_CHUNK / 128; ++g) {
        const int r = wave * (VH_CHUNK / 2) + g * 64 + lane;
        const uint32_t kk = idxs[r];
        const int kl = (int)kk - k0;
        const bool in = kk != 0xffffffffu && kl >= 0 && kl < CB;
        const unsigned long long mask = __ballot(in);
        if (in) {
          list[wave][n + __popcll(mask & lt)] = ((uint32_t)r << 8) | (uint32_t)kl;
          atomicAdd(&counts[kl], 1u);
        }
        n += __popcll(mask);
      }
      if (lane == 0) cnt[wave] = n;
      cst[0] = 0;
      ntot = 0;
    }
    __syncthreads();
    const int n0 = REG ? ntot : cnt[0];
    if (!REG) ntot = n0 + cnt[1];
    const uint32_t* l0 = REG ? lflat : list[0];
    const uint32_t* l1 = REG ? lflat : list[1];  // REG: every listed row is in l0
    for (int s0 = 0; s0 < ntot; s0 += SR) {
      const int ns = min(SR, ntot - s0);
      if (vec16)
        vh_gather<T, piece16>(x, c0, D, l0, n0, l1, s0, ns * D / (16 / (int)sizeof(T)), xs, tid);
      else
        vh_gather<T, T>(x, c0, D, l0, n0, l1, s0, ns * D, xs, tid);
      __syncthreads();
      if (d < D) {
        if constexpr (REG) {
#pragma unroll
          for (int c = 0; c < CB; ++c) {
            const int lo = max(cst[c], s0), hi = min(cst[c + 1], s0 + ns);
            T t = a[c];
#pragma unroll 8
            for (int r = lo; r < hi; ++r) t = t + xs[(r - s0) * D + d];
            a[c] = t;
          }
        } else {
          for (int r = 0; r < ns; ++r) {
            const int p = s0 + r;
            const uint32_t ent = p < n0 ? l0[p] : l1[p - n0];
            const int kl = (int)(ent & 255u);
            acc[kl * D + d] = acc[kl * D + d] + xs[r * D + d];
          }
        }
      }
      __syncthreads();
    }
  }
  if constexpr (REG) {
    if (d < D) {
#pragma unroll
      for (int c = 0; c < CB; ++c) acc[c * D + d] = a[c];
    }
  }
  for (int i = tid; i < CB; i += blockDim.x) mass[i] = (double)counts[i];
  __syncthreads();
  // vlad.c:191-213 per (cluster, dimension), then the per-cluster norms; the
  // block's means are contiguous and 8 of a thread's are loaded together
  const int nval = min(CB, K - k0) * D;
  for (int base = 0; base < nval; base += 128 * 8) {
    T mv[8];
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = base + tid + 128 * u;
      mv[u] = i < nval ? means[(size_t)k0 * D + i] : T(0);
    }
#pragma unroll
    for (int u = 0; u < 8; ++u) {
      const int i = base + tid + 128 * u;
      if (i >= nval) continue;
      const int c = i / D;
      T e = acc[i];
      const T mu = mv[u];
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
static void launch_vlad_hard_cb(T* enc, const T* means, int D, int K, const T* x, const uint64_t* offsets,
                                int numImages, const uint32_t* hard, int flags) {
  const int nkb = (K + CB - 1) / CB;
  const int padImages = numImages >= 8 ? ((numImages + 7) / 8) * 8 : numImages;
  VLHIP_REQUIRE((size_t)padImages * nkb < (size_t(1) << 31), "GPU VLAD encoder: too many (image, cluster block) pairs");
  const size_t sh = (((size_t)CB * D * sizeof(T) + 15) & ~size_t(15)) + CB * sizeof(double) + CB * sizeof(T);
  auto kern = vlad_hard_kernel<T, CB>;
  VLHIP_CHECK(hipFuncSetAttribute((const void*)kern, hipFuncAttributeMaxDynamicSharedMemorySize, (int)sh));
  VLHIP_LAUNCH("vlad_hard", kern, dim3((unsigned)((size_t)padImages * nkb)), dim3(128), sh, x, offsets, numImages, D,
               K, means, hard, flags, enc, nkb);
}

// Blocks of 16 clusters when the images alone give the device >= 1024
// workgroups (4096 images of 3000 rows, K = 512: 3.05 ms, against 3.35 ms with
// blocks of 32 and 10.6 ms with blocks of 8, whose register path sorts every
// chunk); fewer clusters per block (down to one) for small batches, since every
// workgroup walks all of its image's cluster indices.
template <typename T>
static void launch_vlad_hard(T* enc, const T* means, int D, int K, const T* x, const uint64_t* offsets, int numImages,
                             const uint32_t* hard, int flags) {
  int cb = 16;
  while (cb > 1 && (size_t)numImages * (size_t)((K + cb - 1) / cb) < 1024) cb >>= 1;
  switch (cb) {
    case 32: launch_vlad_hard_cb<T, 32>(enc, means, D, K, x, offsets, numImages, hard, flags); break;
    case 16: launch_vlad_hard_cb<T, 16>(enc, means, D, K, x, offsets, numImages, hard, flags); break;
    case 8: launch_vlad_hard_cb<T, 8>(enc, means, D, K, x, offsets, numImages, hard, flags); break;
    case 4: launch_vlad_hard_cb<T, 4>(enc, means, D, K, x, offsets, numImages, hard, flags); break;
    case 2: launch_vlad_hard_cb<T, 2>(enc, means, D, K, x, offsets, numImages, hard, flags); break;
    default: launch_vlad_hard_cb<T, 1>(enc, means, D, K, x, offsets, numImages, hard, flags); break;
  }
}

template <typename T>
static void launch_vlad(T* enc, const T* means, int D, int K, const T* x, const uint64_t* offsets, int numImages,
                        const T* assign, const uint32_t* hard, int flags) {
  if (hard && D <= 128) {
    launch_vlad_hard<T>(enc, means, D, K, x, offsets, numImages, hard, flags);
  } else {
    const int numKB = (K + VS_COMPS - 1) / VS_COMPS, numDS = (D + VS_DIMS - 1) / VS_DIMS;
    const int padImages = numImages >= 8 ? ((numImages + 7) / 8) * 8 : numImages;
    VLHIP_REQUIRE((size_t)padImages * numKB * numDS < (size_t(1) << 31) && (size_t)K * numImages < (size_t(1) << 31),
                  "GPU VLAD encoder: too many (image, cluster block, dimension slice) workgroups");
    const unsigned blocks = (unsigned)((size_t)padImages * numKB * numDS);
    if (hard)
      VLHIP_LAUNCH("vlad_slice", (vlad_slice_kernel<T, true>), dim3(blocks), dim3(256), 0, x, offsets, numImages, D,
                   K, means, assign, hard, flags, enc, numKB, numDS);
    else
      VLHIP_LAUNCH("vlad_slice", (vlad_slice_kernel<T, false>), dim3(blocks), dim3(256), 0, x, offsets, numImages, D,
                   K, means, assign, hard, flags, enc, numKB, numDS);
    if (flags & 1)
      VLHIP_LAUNCH("vlad_component_norm", (vlad_component_norm_kernel<T>), dim3((unsigned)((size_t)K * numImages)),
                   dim3(64), 0, enc, D, K);
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
  VLHIP_REQUIRE(D < (size_t(1) << 24), "GPU VLAD encoder: dimension too large");
  VLHIP_REQUIRE(K < (size_t(1) << 31) && numImages < (size_t(1) << 31), "GPU VLAD encoder: K or image count too large");
  if (numImages == 0) return;
  launch_vlad<T>(enc, means, (int)D, (int)K, x, offsets, (int)numImages, assign, hard, flags);
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
