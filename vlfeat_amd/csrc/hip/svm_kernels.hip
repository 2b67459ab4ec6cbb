// svm_kernels.hip -- linear SVM solvers (vl/svm.c) on the GPU.
//
// SDCA (svm.c:1931-2016) and SGD (svm.c:2018-2142) visit one data point per
// iteration: an inner product of the point with the model, a scalar update,
// and (when the update is not zero) model += multiplier * point.  The
// iterations are sequential, so the parallelism is inside each one: a
// persistent cooperative launch runs a whole stretch of iterations, and each
// of its G workgroups owns a 4096-value slice of the model, which stays in its
// registers for the whole launch.  Per iteration every workgroup forms its
// slice's partial inner product (fixed order), publishes it in a tagged slot
// and reads the G partials back in workgroup order, so every workgroup holds
// the same inner product, computes the same scalar update and applies it to
// its slice; the point of the next iteration (known from the permutation) is
// loaded while the partials are exchanged.  Slots are 64-bit words {tag,
// 32 bits of the partial} written and read with device-scope relaxed atomics
// (a reader that sees the tag sees the bits); a ring of SVM_RING slots per
// workgroup suffices because no workgroup can be more than one iteration
// ahead of another.  Every wait is bounded (SVM_SPIN_MAX polls, or another
// workgroup's abort), after which the launch stops and the host reports it.
//
// The scalar arithmetic restates svm.c:1602-1858 and the solver loops with
// the same double operations (no contraction), so the result differs from the
// reference's only through the summation order of the inner products (the
// reference sums a point's products one by one in double).
#include "vlhip_runtime.h"
#include "vlhip_device.h"
#include "vlhip.h"

#include <algorithm>
#include <cmath>

namespace vlhip {

constexpr int SVM_THREADS = 256;
constexpr int SVM_E = 16;                       // model values per thread
constexpr int SVM_SLICE = SVM_THREADS * SVM_E;  // model values per workgroup
constexpr int SVM_RING = 4;
constexpr int SVM_MAX_GROUPS = 1024;
constexpr long long SVM_SPIN_MAX = 1LL << 22;

enum { SVM_SOLVER_SGD = 1, SVM_SOLVER_SDCA = 2 };
enum { SVM_HINGE = 0, SVM_HINGE2, SVM_L1, SVM_L2, SVM_LOGISTIC };

#pragma clang fp contract(off)

// svm.c:1610-1620, 1654-1661, 1693-1700, 1737-1740, 1772-1779
__device__ double svm_loss_derivative(int loss, double inner, double label) {
  switch (loss) {
    case SVM_HINGE: return (label * inner < 1.0) ? -label : 0.0;
    case SVM_HINGE2: return (label * inner < 1.0) ? 2 * (inner - label) : 0.0;
    case SVM_L1: return (label > inner) ? -1.0 : 1.0;
    case SVM_L2: return -2 * (label - inner);
    default: {
      const double z = label * inner;
      const double t = 1 / (1 + exp(-z));
      return label * (t - 1);
    }
  }
}

// svm.c:1637-1643, 1677-1683, 1716-1726, 1752-1757, 1800-1858
__device__ double svm_dca_update(int loss, double alpha, double inner, double norm2, double label) {
  switch (loss) {
    case SVM_HINGE: {
      const double palpha = (label - inner) / norm2 + alpha;
      return label * fmax(0.0, fmin(1.0, label * palpha)) - alpha;
    }
    case SVM_HINGE2: {
      const double palpha = (label - inner - 0.5 * alpha) / (norm2 + 0.5) + alpha;
      return label * fmax(0.0, label * palpha) - alpha;
    }
    case SVM_L1: {
      if (fabs(alpha) <= 1) {
        const double palpha = (label - inner) / norm2 + alpha;
        return fmax(-1.0, fmin(1.0, palpha)) - alpha;
      }
      return INFINITY;
    }
    case SVM_L2: return (label - inner - 0.5 * alpha) / (norm2 + 0.5);
    default: {
      // five bisection steps for the zero of the derivative on (0, 1), then up
      // to two Newton steps away from the ends
      double beta1 = 0, beta2 = 1, beta = 0.5;
      for (int t = 0; t < 5; ++t) {
        const double df = log(beta) - log(1 - beta);
        const double der = norm2 * beta + label * (inner - norm2 * alpha) + df;
        if (der >= 0) beta2 = beta;
        else beta1 = beta;
        beta = 0.5 * (beta1 + beta2);
      }
      const double eps = 2.220446049250313e-16;
      for (int t = 0; (t < 2) & (beta > eps) & (beta < 1 - eps); ++t) {
        const double df = log(beta) - log(1 - beta);
        const double ddf = 1 / (beta * (1 - beta));
        const double der = norm2 * beta + label * (inner - norm2 * alpha) + df;
        const double dder = norm2 + ddf;
        beta -= der / dder;
        beta = fmax(0.0, fmin(1.0, beta));
      }
      return label * beta - alpha;
    }
  }
}

// Fixed-order sum over the workgroup (butterfly within each wave, then the
// waves in order); every thread receives the result.
__device__ __forceinline__ double svm_block_sum(double v, double* red) {
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
#pragma unroll
  for (int o = 32; o > 0; o >>= 1) v += __shfl_xor(v, o, 64);
  __syncthreads();  // red may still be read from the previous sum
  if (lane == 0) red[w] = v;
  __syncthreads();
  double r = 0.0;
#pragma unroll
  for (int i = 0; i < SVM_THREADS / 64; ++i) r += red[i];
  return r;
}

__device__ __forceinline__ unsigned long long svm_word(unsigned tag, unsigned bits) {
  return ((unsigned long long)tag << 32) | bits;
}

struct SvmArgs {
  const void* data;
  long long D, N;
  const double* labels;
  const double* weights;
  double* model;
  double* alpha;
  const double* norm2;
  double* scores;
  double* prevScores;
  const long long* perm;
  unsigned long long begin, end;
  double lambda, biasMultiplier, biasLearningRate;
  long long sgdT0;
  double bias, factor, biasFactor;
  double* stateOut;
  unsigned long long* slots;  // [SVM_RING][G][2]
  unsigned* abortFlag;
  int loss;
};

template <typename T, int SOLVER>
__global__ __launch_bounds__(SVM_THREADS) void svm_run_kernel(SvmArgs a) {
  __shared__ double red[SVM_THREADS / 64];
  const int tid = threadIdx.x, wg = blockIdx.x, G = gridDim.x;
  const T* x = reinterpret_cast<const T*>(a.data);
  const long long D = a.D, N = a.N;
  const long long d0 = (long long)wg * SVM_SLICE;
  double w[SVM_E];
#pragma unroll
  for (int e = 0; e < SVM_E; ++e) {
    const long long d = d0 + tid + (long long)SVM_THREADS * e;
    w[e] = d < D ? a.model[d] : 0.0;
  }
  double bias = a.bias, factor = a.factor, biasFactor = a.biasFactor;
  T xv[SVM_E];
  auto load_row = [&](long long i) {
#pragma unroll
    for (int e = 0; e < SVM_E; ++e) {
      const long long d = d0 + tid + (long long)SVM_THREADS * e;
      xv[e] = d < D ? x[i * D + d] : T(0);
    }
  };
  long long inext = a.perm[a.begin % (unsigned long long)N];
  load_row(inext);
  bool failed = false;
  for (unsigned long long t = a.begin; t < a.end; ++t) {
    const long long i = inext;
    T xc[SVM_E];
#pragma unroll
    for (int e = 0; e < SVM_E; ++e) xc[e] = xv[e];
    if (t + 1 < a.end) {
      inext = a.perm[(t + 1) % (unsigned long long)N];
      load_row(inext);
    }
    // the point's scalars are read before this workgroup publishes its partial:
    // workgroup 0 writes alpha[i] only after it has every partial
    double p = a.weights ? a.weights[i] : 1.0;
    const double label = a.labels[i];
    double alphai = 0.0, n2 = 0.0;
    if (SOLVER == SVM_SOLVER_SDCA) {
      alphai = a.alpha[i];
      n2 = a.norm2[i];
    } else {
      p = fmax(0.0, p);
    }
    const bool need = SOLVER == SVM_SOLVER_SGD || p > 0;
    double dot = 0.0;
    if (need) {
      double part = 0.0;
#pragma unroll
      for (int e = 0; e < SVM_E; ++e) part += (double)xc[e] * w[e];
      part = svm_block_sum(part, red);
      if (G > 1) {
        const unsigned tag = (unsigned)(t + 1);
        unsigned long long* ring = a.slots + (size_t)(t % SVM_RING) * G * 2;
        if (tid == 0) {
          const unsigned long long bits = (unsigned long long)__double_as_longlong(part);
          __hip_atomic_store(ring + 2 * wg, svm_word(tag, (unsigned)(bits >> 32)), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
          __hip_atomic_store(ring + 2 * wg + 1, svm_word(tag, (unsigned)bits), __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT);
        }
        // thread j sums the partials of workgroups j, j + 256, ... in order
        double v = 0.0;
        for (int j = tid; j < G; j += SVM_THREADS) {
          unsigned long long hi, lo;
          long long spins = 0;
          for (;;) {
            hi = __hip_atomic_load(ring + 2 * j, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            lo = __hip_atomic_load(ring + 2 * j + 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            if ((unsigned)(hi >> 32) == tag && (unsigned)(lo >> 32) == tag) break;
            if (++spins > SVM_SPIN_MAX ||
                __hip_atomic_load(a.abortFlag, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0u) {
              failed = true;
              break;
            }
            __builtin_amdgcn_s_sleep(1);
          }
          if (failed) break;
          v += __longlong_as_double((long long)(((hi & 0xffffffffull) << 32) | (lo & 0xffffffffull)));
        }
        if (__syncthreads_or(failed)) {
          if (tid == 0) __hip_atomic_store(a.abortFlag, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          failed = true;
          break;
        }
        dot = svm_block_sum(v, red);
      } else {
        dot = part;
      }
    }
    double mult = 0.0;
    if (SOLVER == SVM_SOLVER_SDCA) {
      // svm.c:1982-1997
      double delta = 0.0;
      if (need) {
        const double inner = dot + bias * a.biasMultiplier;
        delta = p * svm_dca_update(a.loss, alphai / p, inner, p * n2, label);
      }
      if (delta != 0) {
        if (wg == 0 && tid == 0) a.alpha[i] = alphai + delta;
        mult = delta / ((double)N * a.lambda);
        bias += a.biasMultiplier * mult;
      }
    } else {
      // svm.c:2078-2107
      const double inner = factor * dot + biasFactor * (a.biasMultiplier * bias);
      const double gradient = p * svm_loss_derivative(a.loss, inner, label);
      if (wg == 0 && tid == 0) {
        a.prevScores[i] = a.scores[i];
        a.scores[i] = inner;
      }
      const double rate = 1.0 / (a.lambda * (double)(t + (unsigned long long)a.sgdT0));
      const double biasRate = rate * a.biasLearningRate;
      factor *= (1.0 - a.lambda * rate);
      biasFactor *= (1.0 - a.lambda * biasRate);
      if (gradient != 0) {
        mult = -gradient * rate / factor;
        bias += a.biasMultiplier * (-gradient * biasRate / biasFactor);
      }
    }
    if (mult != 0.0) {
#pragma unroll
      for (int e = 0; e < SVM_E; ++e) w[e] += (double)xc[e] * mult;
    }
  }
#pragma unroll
  for (int e = 0; e < SVM_E; ++e) {
    const long long d = d0 + tid + (long long)SVM_THREADS * e;
    if (d < D) a.model[d] = w[e];
  }
  if (wg == 0 && tid == 0 && !failed) {
    a.stateOut[0] = bias;
    a.stateOut[1] = factor;
    a.stateOut[2] = biasFactor;
  }
}

// scores[k] = <x_k, model> + biasTerm, one workgroup per point (fixed order).
template <typename T>
__global__ __launch_bounds__(SVM_THREADS) void svm_scores_kernel(const T* __restrict__ x, const double* __restrict__ model,
                                                                 long long D, long long N, double biasTerm,
                                                                 double* __restrict__ scores) {
  __shared__ double red[SVM_THREADS / 64];
  const long long k = blockIdx.x;
  if (k >= N) return;
  double s = 0.0;
  for (long long d = threadIdx.x; d < D; d += SVM_THREADS) s += (double)x[k * D + d] * model[d];
  s = svm_block_sum(s, red);
  if (threadIdx.x == 0) scores[k] = s + biasTerm;
}

// out[k] = <x_k, x_k> (each value squared in double), one workgroup per point.
template <typename T>
__global__ __launch_bounds__(SVM_THREADS) void svm_square_norms_kernel(const T* __restrict__ x, long long D, long long N,
                                                                       double* __restrict__ out) {
  __shared__ double red[SVM_THREADS / 64];
  const long long k = blockIdx.x;
  if (k >= N) return;
  double s = 0.0;
  for (long long d = threadIdx.x; d < D; d += SVM_THREADS) {
    const double v = (double)x[k * D + d];
    s += v * v;
  }
  s = svm_block_sum(s, red);
  if (threadIdx.x == 0) out[k] = s;
}

__global__ void svm_scale_kernel(double* __restrict__ v, long long n, double factor) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) v[i] *= factor;
}

__global__ void svm_zero_kernel(unsigned long long* __restrict__ p, long long n) {
  const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ull;
}

#pragma clang fp contract(on)

template <typename T, int SOLVER>
static void svm_launch(SvmArgs& a, int G) {
  void* args[] = {(void*)&a};
  auto kern = svm_run_kernel<T, SOLVER>;
  int perCu = 0;
  VLHIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&perCu, (const void*)kern, SVM_THREADS, 0));
  VLHIP_REQUIRE((long long)perCu * ctx().num_cus >= G,
                "the SVM solver's workgroups do not fit the device at once (model dimension too large)");
  hipEvent_t ts = nullptr;
  const int ti = timing_begin("svm_run", &ts);
  VLHIP_CHECK(hipLaunchCooperativeKernel((const void*)kern, dim3((unsigned)G), dim3(SVM_THREADS), args, 0,
                                         ctx().stream));
  if (ti >= 0) timing_end(ti, ts);
}

}  // namespace vlhip

using namespace vlhip;

// Workspace: the exchange slots [SVM_RING][G][2], then the abort flag and
// the state written back by the launch (bias, factor, biasFactor).
static size_t svm_slot_bytes(size_t dimension) {
  const size_t G = (dimension + SVM_SLICE - 1) / SVM_SLICE;
  return (size_t)SVM_RING * G * 2 * sizeof(unsigned long long);
}

extern "C" size_t vlhip_svm_workspace_bytes(size_t dimension) { return svm_slot_bytes(dimension) + 256; }

extern "C" void vlhip_svm_run(VlhipSvmRun* r) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  VLHIP_REQUIRE(r->dtype == 1 || r->dtype == 2, "vlhip_svm_run: unknown data type");
  VLHIP_REQUIRE(r->solver == SVM_SOLVER_SGD || r->solver == SVM_SOLVER_SDCA, "vlhip_svm_run: unknown solver");
  VLHIP_REQUIRE(r->loss >= SVM_HINGE && r->loss <= SVM_LOGISTIC, "vlhip_svm_run: unknown loss");
  VLHIP_REQUIRE(r->dimension >= 1 && r->numData >= 1, "vlhip_svm_run: empty problem");
  VLHIP_REQUIRE(r->end < 0xffffffffull, "vlhip_svm_run: iteration numbers must stay below 2^32 - 1");
  if (r->begin >= r->end) return;
  VLHIP_REQUIRE(r->end - r->begin <= r->numData && r->begin / r->numData == (r->end - 1) / r->numData,
                "vlhip_svm_run: a run must stay within one epoch");
  const long long G = (long long)((r->dimension + SVM_SLICE - 1) / SVM_SLICE);
  VLHIP_REQUIRE(G <= SVM_MAX_GROUPS, "vlhip_svm_run: model dimension above 4,194,304");
  SvmArgs a;
  a.data = r->data;
  a.D = (long long)r->dimension;
  a.N = (long long)r->numData;
  a.labels = r->labels;
  a.weights = r->weights;
  a.model = r->model;
  a.alpha = r->alpha;
  a.norm2 = r->norm2;
  a.scores = r->scores;
  a.prevScores = r->previousScores;
  a.perm = (const long long*)r->permutation;
  a.begin = r->begin;
  a.end = r->end;
  a.lambda = r->lambda;
  a.biasMultiplier = r->biasMultiplier;
  a.biasLearningRate = r->biasLearningRate;
  a.sgdT0 = (long long)r->sgdT0;
  a.bias = r->state[0];
  a.factor = r->state[1];
  a.biasFactor = r->state[2];
  a.loss = r->loss;
  unsigned char* ws = (unsigned char*)r->workspace;
  a.slots = (unsigned long long*)ws;
  a.abortFlag = (unsigned*)(ws + svm_slot_bytes(r->dimension));
  a.stateOut = (double*)(ws + svm_slot_bytes(r->dimension) + 64);
  if (r->dtype == 1) {
    if (r->solver == SVM_SOLVER_SDCA) svm_launch<float, SVM_SOLVER_SDCA>(a, (int)G);
    else svm_launch<float, SVM_SOLVER_SGD>(a, (int)G);
  } else {
    if (r->solver == SVM_SOLVER_SDCA) svm_launch<double, SVM_SOLVER_SDCA>(a, (int)G);
    else svm_launch<double, SVM_SOLVER_SGD>(a, (int)G);
  }
  unsigned ab = 0;
  download(&ab, a.abortFlag, sizeof(unsigned));
  VLHIP_REQUIRE(ab == 0, "the SVM solver's exchange between workgroups timed out");
  download(r->state, a.stateOut, 3 * sizeof(double));
}

extern "C" void vlhip_svm_reset_workspace(void* workspace, size_t dimension) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  const long long words = (long long)(vlhip_svm_workspace_bytes(dimension) / 8);
  VLHIP_LAUNCH("svm_zero", svm_zero_kernel, dim3((unsigned)((words + 255) / 256)), dim3(256), 0,
               (unsigned long long*)workspace, words);
}

extern "C" void vlhip_svm_scores(int dtype, double* scores, void const* data, double const* model, size_t dimension,
                                 size_t numData, double biasTerm) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (numData == 0) return;
  VLHIP_REQUIRE(numData < (size_t)1 << 31, "vlhip_svm_scores: too many points");
  if (dtype == 1)
    VLHIP_LAUNCH("svm_scores", svm_scores_kernel<float>, dim3((unsigned)numData), dim3(SVM_THREADS), 0,
                 (const float*)data, model, (long long)dimension, (long long)numData, biasTerm, scores);
  else if (dtype == 2)
    VLHIP_LAUNCH("svm_scores", svm_scores_kernel<double>, dim3((unsigned)numData), dim3(SVM_THREADS), 0,
                 (const double*)data, model, (long long)dimension, (long long)numData, biasTerm, scores);
  else
    fatal("vlhip_svm_scores", __FILE__, __LINE__, "unknown data type");
}

extern "C" void vlhip_svm_square_norms(int dtype, double* out, void const* data, size_t dimension, size_t numData) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (numData == 0) return;
  VLHIP_REQUIRE(numData < (size_t)1 << 31, "vlhip_svm_square_norms: too many points");
  if (dtype == 1)
    VLHIP_LAUNCH("svm_square_norms", svm_square_norms_kernel<float>, dim3((unsigned)numData), dim3(SVM_THREADS), 0,
                 (const float*)data, (long long)dimension, (long long)numData, out);
  else if (dtype == 2)
    VLHIP_LAUNCH("svm_square_norms", svm_square_norms_kernel<double>, dim3((unsigned)numData), dim3(SVM_THREADS), 0,
                 (const double*)data, (long long)dimension, (long long)numData, out);
  else
    fatal("vlhip_svm_square_norms", __FILE__, __LINE__, "unknown data type");
}

extern "C" void vlhip_svm_scale(double* v, size_t n, double factor) {
  ensure_init();
  std::lock_guard<std::recursive_mutex> lk(ctx().mu);
  if (n == 0) return;
  VLHIP_LAUNCH("svm_scale", svm_scale_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, v, (long long)n,
               factor);
}
