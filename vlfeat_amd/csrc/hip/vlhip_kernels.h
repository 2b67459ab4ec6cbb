// vlhip_kernels.h -- tile constants and launchers shared between the GMM,
// Fisher, VLAD and k-means translation units.
#pragma once

#include <cstddef>
#include <cstdint>

namespace vlhip {

// E-step register tiling: 64 descriptors x 64 components per sweep step,
// 16x16 threads each owning a 4x4 tile, D staged through LDS in chunks.
constexpr int EM_ROWS = 64;
constexpr int EM_COLS = 64;
constexpr int EM_DC = 16;

// M-step moments: 32 components x 64-row chunks per workgroup step.
constexpr int MS_COMPS = 32;
constexpr int MS_ROWS = 64;

template <typename T>
struct GmmPrep {
  T* inv;         // K x D, 1 / sigma^2
  T* isig;        // K x D, sqrt(1 / sigma^2)
  double* cterm;  // K, log prior - D/2 log 2pi - 1/2 log det
};

template <typename T>
GmmPrep<T> gmm_prepare(const T* priors, const T* covs, size_t D, size_t K);

// Runs the E-step over N rows; q (N x K) optional, row statistics optional.
// Returns the FP64 log-likelihood (sum over rows).
template <typename T>
double run_estep(const T* x, size_t N, size_t D, size_t K, const T* means, const GmmPrep<T>& prep,
                 T* q, T* row_m, T* row_s, int* row_arg, const char* name, bool want_ll);

// Exact float quantizer screened on the matrix cores (kmeans_mfma.hip);
// returns false (nothing launched) when the shape is not supported.  ws: a
// device workspace of quantize_screened_bytes(D, K) bytes, or nullptr for the
// library's own grow-only table.
size_t quantize_screened_bytes(size_t D, size_t K);
bool quantize_screened(uint32_t* assign, float* dist, const float* x, const float* centers, size_t N, size_t D,
                       size_t K, int lanes, void* ws);

// Divides every image's encoding by max(||enc||, 1e-12) (fisher.c:474-485).
void launch_fisher_normalize(float* enc, const double* normpart, int numKB, size_t numImages, size_t len);

// Learned-gate setting of the matrix-core encoder (nullptr: dense).
struct GateUse {
  void* gate;
  int M;       // candidates per descriptor (<= 16)
  float need;  // predicted mass a descriptor needs to skip the dense E-step
};

// Split-16 matrix-core Fisher encoder (fisher_mfma.hip); fp16 selects fp16
// operands, otherwise bf16.
void fisher_split16(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                    const float* x, const uint64_t* offsets, size_t numImages, int flags,
                    unsigned long long* numTerms, bool fp16,
                    const GateUse* gate);

// FP32 matrix-core Fisher encoder (fisher_mfma.hip): plain FP32 operands on
// v_mfma_f32_16x16x4_f32, D <= 128.
void fisher_f32mfma(float* enc, const float* means, const float* covs, const float* priors, size_t D, size_t K,
                    const float* x, const uint64_t* offsets, size_t numImages, int flags,
                    unsigned long long* numTerms);

// VLAD of a batch of images (vlad_kernels.hip); exactly one of assign / hard.
template <typename T>
void vlad_t(T* enc, const T* means, size_t D, size_t K, const T* x, const uint64_t* offsets, size_t numImages,
            const T* assign, const uint32_t* hard, int flags);

// Exact nearest-centre quantizer (kmeans_kernels.hip); dist may be nullptr;
// ws as for quantize_screened.
template <typename T>
void quantize_t(uint32_t* assign, T* dist, const T* x, const T* centers, size_t N, size_t D, size_t K, int lanes,
                void* ws = nullptr);

// Pre-encoding projection (project_kernels.hip): out = P (x - c), each row
// optionally renormalised to unit L2 norm; float, outDimension <= 128.
void project_t(float* out, const float* x, size_t N, size_t Din, size_t D, const float* P, const float* c,
               int renorm);

}  // namespace vlhip
