/* vlhip.h -- device-resident C ABI of the MI355X GMM / Fisher / VLAD engine.
 *
 * The VLFeat-compatible entry points (vl/gmm.h, vl/fisher.h, vl/vlad.h,
 * vl/kmeans.h) take host buffers, as the reference does.  The functions below
 * are the layer those entry points call: every array argument is a DEVICE
 * pointer on the current GPU unless its name ends in _host, sizes are element
 * counts, and dtype is VL_TYPE_FLOAT (1) or VL_TYPE_DOUBLE (2).  Results are
 * complete when a function returns (the library synchronises its stream before
 * returning a host value, and otherwise leaves work queued in stream order;
 * vlhip_synchronize waits for it).
 *
 * Each entry point cites the reference routine whose semantics it computes.
 * None of them allocates host memory visible to the caller; device scratch is
 * owned by the library and reused across calls.  Errors (bad arguments, HIP
 * failures, no GPU) print a message through the VLFeat printf hook and abort(),
 * like the reference's abort() on an unknown type (gmm.c:1514).
 */
#ifndef VLHIP_H
#define VLHIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
#  define VLHIP_API extern "C" __attribute__((visibility("default")))
#else
#  define VLHIP_API extern __attribute__((visibility("default")))
#endif

/* Precision / algorithm of the Fisher encoder (mode argument). */
#define VLHIP_MODE_FP32     0  /* direct-form FP32 (FP64 for double data): the parity path */
#define VLHIP_MODE_BF16X3   1  /* compensated split-BF16 MFMA contractions, FP32 accumulation */
#define VLHIP_MODE_GATED    2  /* learned gate + exact top-M posteriors, dense fallback */
#define VLHIP_MODE_F16X3    3  /* compensated split-FP16 MFMA contractions (power-of-two scaled) */
#define VLHIP_MODE_FP32_MFMA 4 /* plain FP32 operands on the FP32 matrix cores (float data, D <= 128) */

/* ---- runtime ------------------------------------------------------------ */
VLHIP_API int    vlhip_available (void) ;                 /* 1 when a GPU can be used (never aborts) */
VLHIP_API void   vlhip_init (void) ;                      /* create the context on device 0 */
VLHIP_API void   vlhip_set_stream (void * hipStream) ;    /* NULL = legacy null stream */
VLHIP_API void * vlhip_malloc (size_t bytes) ;
VLHIP_API void   vlhip_free (void * ptr) ;
VLHIP_API void   vlhip_memcpy_h2d (void * dst, void const * src_host, size_t bytes) ;
VLHIP_API void   vlhip_memcpy_d2h (void * dst_host, void const * src, size_t bytes) ;
VLHIP_API void   vlhip_memcpy_d2d (void * dst, void const * src, size_t bytes) ;  /* stream-ordered */
/* Stream-ordered copies that return without waiting: the host buffer must stay
 * valid (h2d) or untouched (d2h) until vlhip_synchronize or a synchronous call. */
VLHIP_API void   vlhip_memcpy_h2d_async (void * dst, void const * src_host, size_t bytes) ;
VLHIP_API void   vlhip_memcpy_d2h_async (void * dst_host, void const * src, size_t bytes) ;
VLHIP_API void   vlhip_memset (void * dst, int value, size_t bytes) ;
VLHIP_API void   vlhip_synchronize (void) ;
VLHIP_API int    vlhip_num_compute_units (void) ;

/* Library-owned, grow-only device buffers for staging host arrays.  A buffer
 * keeps its contents until the next vlhip_buffer call for the same index with
 * a larger size.  Callers that stage data and then run kernels hold the
 * library lock across the sequence so concurrent host threads cannot reuse
 * the same buffer in between. */
#define VLHIP_BUF_DATA    0
#define VLHIP_BUF_PARAMS  1
#define VLHIP_BUF_POST    2
#define VLHIP_BUF_ENC     3
#define VLHIP_BUF_AUX     4
#define VLHIP_BUF_AUX2    5
VLHIP_API void * vlhip_buffer (int which, size_t bytes) ;
VLHIP_API void   vlhip_lock (void) ;
VLHIP_API void   vlhip_unlock (void) ;

/* Per-kernel timing with hipEvents recorded on the library stream around each
 * launch.  vlhip_timing_get returns the accumulated milliseconds and launch
 * count of kernel `index` (0 <= index < vlhip_timing_count()). */
VLHIP_API void   vlhip_timing_enable (int on) ;
VLHIP_API void   vlhip_timing_reset (void) ;
VLHIP_API int    vlhip_timing_count (void) ;
VLHIP_API int    vlhip_timing_get (int index, char * name, size_t nameCapacity,
                                   double * total_ms, unsigned long long * count) ;

/* ---- GMM (vl/gmm.c) ----------------------------------------------------- */

/* E-step, gmm.c:711-798.  posteriors (N x K, row-major, may be NULL when only
 * the log-likelihood is wanted) receives the normalised posteriors; returns the
 * log-likelihood summed over the data in FP64. */
VLHIP_API double vlhip_gmm_posteriors (int dtype, void * posteriors, size_t numClusters, size_t numData,
                                       void const * priors, void const * means, size_t dimension,
                                       void const * covariances, void const * data) ;

/* M-step sufficient statistics, gmm.c:1064-1167: for every component k with
 * posterior p_ik >= threshold,  S0[k] = sum p,  S1[k,:] = sum p x,
 * S2[k,:] = sum p (x - oldMeans[k,:])^2, and mass[k] = sum of all p_ik
 * (gmm.c:836-844).  out_host is a HOST array of 2K + 2KD doubles laid out
 * [S0 (K) | mass (K) | S1 (K x D) | S2 (K x D)], reduced in a fixed order. */
VLHIP_API void vlhip_gmm_moments (int dtype, double * out_host,
                                  void const * posteriors, void const * data, void const * oldMeans,
                                  size_t numData, size_t dimension, size_t numClusters, double threshold) ;

/* One EM step on the matrix cores for float data, dimension <= 128 (mode
 * VLHIP_MODE_BF16X3 or VLHIP_MODE_F16X3).  Replaces, for the EM loop of
 * gmm.c:1220-1308, the E-step of vlhip_gmm_posteriors (gmm.c:711-798) and
 * the moments of vlhip_gmm_moments (gmm.c:1064-1167) with compensated 16-bit
 * contractions accumulated in FP32 and reduced in FP64; the posterior matrix
 * is never materialised.  Returns the log-likelihood of the parameters
 * (FP64 sum); when out_host is not NULL also fills it, as vlhip_gmm_moments
 * does, with the M-step moments of these posteriors against oldMeans = means
 * (threshold 0.01 / numClusters).  All array arguments are device pointers
 * except out_host. */
VLHIP_API double vlhip_gmm_em_step (int mode, double * out_host, void const * priors, void const * means,
                                    void const * covariances, size_t dimension, size_t numClusters,
                                    void const * data, size_t numData) ;

/* Empty-mode split, gmm.c:968-990: datum i moves its mass p_ij + p_ii to
 * component iCl when data[i, dim] < mu, otherwise to jCl. */
VLHIP_API void vlhip_gmm_split (int dtype, void * posteriors, void const * data, size_t numData,
                                size_t dimension, size_t numClusters, size_t iCl, size_t jCl,
                                size_t dim, double mu) ;

/* One-hot posteriors from hard assignments, gmm.c:1361-1364. */
VLHIP_API void vlhip_gmm_onehot (int dtype, void * posteriors, uint32_t const * assignments,
                                 size_t numData, size_t numClusters) ;

/* ---- Fisher (vl/fisher.c) ------------------------------------------------ */

/* Fisher vectors of numImages images, fisher.c:363-488 per image.  Image i
 * owns descriptor rows [offsets[i], offsets[i+1]) of data; enc receives
 * numImages x 2*D*K values.  numTerms (device, may be NULL) receives one
 * count per image.  mode is a VLHIP_MODE_* value. */
VLHIP_API void vlhip_fisher_encode_batch (int dtype, void * enc,
                                          void const * means, void const * covariances, void const * priors,
                                          size_t dimension, size_t numClusters,
                                          void const * data, uint64_t const * offsets, size_t numImages,
                                          int flags, int mode, unsigned long long * numTerms) ;

/* Choice between the two matrix-core Fisher encoders (F16X3 / BF16X3 modes):
 * the two-pass encoder (first pass: row log-sum-exp over all components;
 * second pass: posteriors and moments per component block) and the fused
 * single-pass encoder (the component blocks of an image exchange per-row
 * partial log-sum-exps, so every logit is computed once; one cooperative
 * launch, float data, dimension <= 80, 129 to 2048 components).  mode 0:
 * always two-pass; 1 (default): fused when the batch has at least one image
 * per stream of the launch and the mixture fits one component block (K <=
 * 256), where it is the faster of the two; 2: fused whenever the shape
 * allows.  Both are as accurate against an FP64 evaluation; their encodings
 * differ by rounding.  An extension of this library, no reference
 * counterpart. */
VLHIP_API void vlhip_fisher_set_fused (int mode) ;

/* Gate used by VLHIP_MODE_GATED (a vlhip_gate_new handle, NULL to clear):
 * per descriptor its `candidates` (<= 16) most probable components; groups of
 * 32 consecutive descriptors of an image whose predicted mass on those
 * candidates reaches `need` for every descriptor get exact posteriors over the
 * candidates only (normalised over them), all others go through the dense
 * compensated FP16 E-step.  No counterpart in the reference. */
VLHIP_API void vlhip_fisher_set_gate (void * gate, int candidates, float need) ;
/* Groups of 32 descriptors seen by VLHIP_MODE_GATED encodings since the last
 * reset, and how many of them the gate served (host values). */
VLHIP_API void vlhip_fisher_gate_stats (unsigned long long * groups, unsigned long long * served, int reset) ;

/* Encode plans for batches of uneven images (no counterpart in the
 * reference, whose callers loop over images).  A plan fixes the dimension, the
 * number of components, the images' descriptor counts (offsets_host, HOST
 * array of numImages+1 values starting at 0) and the flags; it prepares the
 * image-aligned group tables and device buffers once and, on the first run,
 * picks the second pass's component tiling (16, 32 components per wave x 4, 8
 * waves per workgroup) and component-block schedule (XCD-aware image-major,
 * linear, XCD-aware component-major) by timing candidates on the plan's own
 * shapes: every tiling with the first schedule, then the other schedules with
 * the fastest tiling (choices cached per shape).  vlhip_encode_plan_run computes the compensated-FP16
 * Fisher vectors (fisher.c:363-488 per image) and, when vladFlags >= 0, the
 * VLAD of every image with each descriptor assigned to its nearest mean
 * (kmeans.c:655-702, vlad.c:163-242); its kernel sequence is captured into a
 * hipGraph on the first run and replayed while the pointers stay the same.
 * Float data, dimension <= 128.  Runs are synchronous. */
VLHIP_API void * vlhip_encode_plan_new (size_t dimension, size_t numClusters, uint64_t const * offsets_host,
                                        size_t numImages, int fisherFlags, int vladFlags) ;
VLHIP_API void   vlhip_encode_plan_run (void * plan, void const * data, void const * means,
                                        void const * covariances, void const * priors, void * fisherEnc,
                                        void * vladEnc, unsigned long long * numTerms) ;
/* Load stage of a plan (apps/recognition/encodeImage.m:91-97): from the next
 * run on, the data given to vlhip_encode_plan_run have inDimension columns and
 * every descriptor is projected, y = projection (x - center), and, when
 * renormalize is set, scaled to unit L2 norm (floor 1e-12) before it is
 * encoded; see vlhip_project.  projection (dimension x inDimension, row-major)
 * and center (inDimension) are device arrays copied into the plan. */
VLHIP_API void   vlhip_encode_plan_set_projection (void * plan, size_t inDimension, void const * projection,
                                                   void const * center, int renormalize) ;
/* chosen tiling (0: 8 waves x 32 components per workgroup, 1: 4 x 32,
 * 2: 8 x 16, 3: 4 x 16), the milliseconds of the chosen candidate and of
 * the default tiling with the first schedule (-1 if not timed: a choice
 * reused from the cache), batches */
VLHIP_API void   vlhip_encode_plan_info (void * plan, int * variant, double * variantMs, size_t * numBatches) ;
/* The autotuner's candidates in the order timed: returns their number and,
 * for 0 <= index < number, candidate index's components per wave, waves per
 * workgroup, schedule (0, 1, 2 as above) and milliseconds per run. */
VLHIP_API int    vlhip_encode_plan_candidates (void * plan, int index, int * componentsPerWave, int * waves,
                                               int * schedule, double * ms) ;
VLHIP_API void   vlhip_encode_plan_delete (void * plan) ;

/* ---- pre-encoding projection (apps/recognition) ---------------------------- */

/* PCA / whitening projection of float descriptors before encoding, as
 * apps/recognition/encodeImage.m:91-97 applies it (learned by
 * trainEncoder.m:150-178, whitening folded into the rows of projection):
 * out[i] = projection (data[i] - center), then, when renormalize is set,
 * out[i] * (1 / max(||out[i]||, 1e-12)).  projection is outDimension x
 * inDimension (row-major), center has inDimension entries; FP32 products and
 * sums (matrix cores), outDimension <= 128.  No counterpart in the C library
 * of the reference, whose application does this in MATLAB. */
VLHIP_API void vlhip_project (void * out, void const * data, size_t numData, size_t inDimension,
                              size_t outDimension, void const * projection, void const * center, int renormalize) ;

/* ---- homogeneous kernel map (vl/homkermap.c) -------------------------------- */

/* The map of every value of in (n values, dtype) into out (n x (2*order+1)
 * values, row i for in[i]), homkermap.c:532-571: table holds the map object's
 * rows of 2*order+1 doubles for numSubdivisions mantissas of every exponent
 * minExponent .. maxExponent (vl_homogeneouskernelmap_new).  Bitwise the
 * reference's vl_homogeneouskernelmap_evaluate_{f,d}. */
VLHIP_API void vlhip_homkermap_evaluate (int dtype, void * out, void const * in, size_t n, double const * table,
                                         int order, int numSubdivisions, int minExponent, int maxExponent) ;

/* One step of Elkan's k-means refinement (kmeans.c:1162-1745) over device
 * arrays, one thread per datum, the reference's operations in its order:
 *   step 0  initial assignment and bounds (table: K x K center distances);
 *   step 1  bounds moved by the centers' moves (aux: K move distances);
 *   step 2  reassignment guided by the bounds (table: K x K center
 *           distances, aux: each center's distance to its nearest other
 *           center); *moved (device, cleared by the caller) becomes 1 when a
 *           datum changes center;
 *   step 3  distances of every datum to its center (distances, numData).
 * upperBounds (numData), upperBoundIsExact (numData bytes) and lowerBounds
 * (numData x numCenters) persist between steps; all distances are squared L2
 * in the host's lane order (lanes 8, 4, 2 or 1), so the steps reproduce the
 * reference's bounds, assignments and distances bit for bit. */
VLHIP_API void vlhip_kmeans_elkan (int dtype, int step, void const * data, void const * centers, void const * table,
                                   void const * aux, size_t numData, size_t dimension, size_t numCenters, int lanes,
                                   uint32_t * assignments, void * upperBounds, unsigned char * upperBoundIsExact,
                                   void * lowerBounds, unsigned * moved, void * distances) ;

/* ---- SVM solvers (vl/svm.c) ----------------------------------------------- */

/* One run of SDCA (svm.c:1931-2016) or SGD (svm.c:2018-2142) iterations
 * [begin, end) inside one epoch, on device arrays: point i of the epoch's
 * iteration t is permutation[t % numData]; model, alpha (SDCA), scores and
 * previousScores (SGD) are updated in place; state holds (bias, factor,
 * biasFactor) on entry and receives them on return (host array; SGD's model
 * is factor * model and its bias biasFactor * bias, as in the reference).
 * norm2[i] (SDCA) = (<x_i, x_i> + biasMultiplier^2) / (lambda numData).
 * workspace: vlhip_svm_workspace_bytes(dimension) bytes, cleared by
 * vlhip_svm_reset_workspace before the first run of a training.  The runs of
 * one workspace must not overlap.  One persistent cooperative launch of
 * ceil(dimension / 4096) workgroups; dimension <= 4,194,304. */
typedef struct VlhipSvmRun {
  int dtype ;                      /* 1 float, 2 double points */
  int solver ;                     /* 1 SGD, 2 SDCA (VlSvmSolverType) */
  int loss ;                       /* VlSvmLossType */
  void const * data ;              /* numData x dimension points */
  size_t dimension, numData ;
  double const * labels ;
  double const * weights ;         /* NULL: every weight 1 */
  double * model ;
  double * alpha ;
  double const * norm2 ;
  double * scores, * previousScores ;
  int64_t const * permutation ;
  uint64_t begin, end ;
  double lambda, biasMultiplier, biasLearningRate ;
  int64_t sgdT0 ;                  /* max(2, ceil(1 / lambda)), svm.c:2026 */
  double state [3] ;
  void * workspace ;
} VlhipSvmRun ;
VLHIP_API size_t vlhip_svm_workspace_bytes (size_t dimension) ;
VLHIP_API void vlhip_svm_reset_workspace (void * workspace, size_t dimension) ;
VLHIP_API void vlhip_svm_run (VlhipSvmRun * run) ;
/* scores[k] = <x_k, model> + biasTerm for every point (svm.c:1885-1886) */
VLHIP_API void vlhip_svm_scores (int dtype, double * scores, void const * data, double const * model,
                                 size_t dimension, size_t numData, double biasTerm) ;
/* out[k] = <x_k, x_k> in double (svm.c:1943-1950) */
VLHIP_API void vlhip_svm_square_norms (int dtype, double * out, void const * data, size_t dimension,
                                       size_t numData) ;
/* v *= factor elementwise (svm.c:2115-2116) */
VLHIP_API void vlhip_svm_scale (double * v, size_t n, double factor) ;

/* ---- VLAD (vl/vlad.c) ---------------------------------------------------- */

/* VLAD of numImages images, vlad.c:163-242 per image.  Exactly one of
 * assignments (dense N x K soft assignments, dtype) and hardAssignments (N
 * cluster indices) is non-NULL. */
VLHIP_API void vlhip_vlad_encode_batch (int dtype, void * enc, void const * means,
                                        size_t dimension, size_t numClusters,
                                        void const * data, uint64_t const * offsets, size_t numImages,
                                        void const * assignments, uint32_t const * hardAssignments,
                                        int flags) ;

/* ---- k-means (vl/kmeans.c) ----------------------------------------------- */

/* Exact L2 quantizer, kmeans.c:655-702: argmin over centers with strict '<'
 * (first index wins), distances evaluated in the host's vector-lane order
 * (lanes = 8 for the AVX path, 4/2 for SSE2, 1 scalar; mathop.c:472-550) so the
 * assignments match the reference bit for bit.  distances may be NULL. */
VLHIP_API void vlhip_kmeans_quantize (int dtype, uint32_t * assignments, void * distances,
                                      void const * data, void const * centers,
                                      size_t numData, size_t dimension, size_t numCenters, int lanes) ;

/* vlhip_kmeans_quantize without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code under
 * hipGraph capture, serving loops, harnesses that forbid hipMalloc).  Same
 * arguments and results as vlhip_kmeans_quantize; the matrix-core screen's
 * centre table lives in a static device array (2.0 MiB per GPU, held until the
 * process ends), so float data with dimension <= 128 takes at most 4096
 * centers (more aborts with a message); other shapes use the exhaustive
 * kernel, which needs no work buffer.  Every call on a GPU shares that GPU's
 * table: calls that overlap on one GPU race. */
VLHIP_API void vlhip_kmeans_quantize_noalloc (int dtype, uint32_t * assignments, void * distances,
                                              void const * data, void const * centers,
                                              size_t numData, size_t dimension, size_t numCenters, int lanes) ;

/* Statistics of the quantizer's matrix-core screen, counted while kernel
 * timing is on (vlhip_timing_enable): counts[0] rows screened, counts[1]
 * exact distances evaluated for listed candidates (a center, or every center
 * one lane of the screen saw when two of them were near-ties), counts[2] rows
 * scanned exactly over all centers (no finite screen value, or more list
 * entries than the screen keeps).  A
 * diagnostic with no counterpart in the reference.  reset != 0 clears them. */
VLHIP_API void vlhip_kmeans_screen_stats (uint64_t * counts, int reset) ;

/* Squared L2 distance of every datum to one point, same lane order
 * (used by k-means++ seeding, kmeans.c:610-616). */
VLHIP_API void vlhip_l2_to_point (int dtype, void * distances, void const * data, void const * point,
                                  size_t numData, size_t dimension, int lanes) ;

/* Center update of Lloyd and ANN refinement, kmeans.c:899-931 (Lloyd) and
 * 1075-1107 (ANN), L2 distance: center c becomes the mean of the data
 * members[starts[c] .. starts[c+1]) (data indices in increasing order, so the
 * float sums are the reference's, bit for bit), or datum restartRow[c] when
 * the cluster is empty.  starts has numCenters + 1 entries. */
VLHIP_API void vlhip_kmeans_update_centers (int dtype, void * centers, void const * data,
                                            uint32_t const * members, uint64_t const * starts,
                                            uint64_t const * restartRow, size_t numCenters, size_t dimension) ;

/* Keep-previous rule of an ANN assignment update, kmeans.c:750-763: datum i
 * keeps center previous[i], at its distance to that center (same lane order as
 * the quantizer), unless distances[i] (from the new search) is strictly
 * smaller.  assignments and distances are updated in place. */
VLHIP_API void vlhip_kmeans_ann_merge (int dtype, uint32_t * assignments, void * distances,
                                       uint32_t const * previous, void const * data, void const * centers,
                                       size_t numData, size_t dimension, int lanes) ;

/* ---- learned gate of the Fisher encoder (no counterpart in the reference) --
 * A perceptron dimension -> hidden1 -> hidden2 -> numClusters with GELU
 * activations (BF16 matrix products, FP32 master weights) trained on the GPU
 * to predict the exact posteriors of a GMM (soft-target cross-entropy, Adam).
 * Parameters are laid out [W1 (D x H1) | b1 | W2 (H1 x H2) | b2 | W3 (H2 x K)
 * | b3], weights row-major with the input index first; the input is
 * standardised per dimension, (x - mean) * istd. */
VLHIP_API void * vlhip_gate_new (size_t dimension, size_t numClusters, size_t hidden1, size_t hidden2,
                                 unsigned long long seed) ;
VLHIP_API void   vlhip_gate_delete (void * gate) ;
VLHIP_API size_t vlhip_gate_num_params (void * gate) ;
VLHIP_API void   vlhip_gate_get_params (void * gate, float * params_host, float * mean_host, float * istd_host) ;
VLHIP_API void   vlhip_gate_set_params (void * gate, float const * params_host, float const * mean_host,
                                        float const * istd_host) ;
/* mean and 1 / (unbiased standard deviation) of numData FP32 descriptors */
VLHIP_API void   vlhip_gate_set_standardization_from_data (void * gate, void const * data, size_t numData) ;
/* mean loss of one batch (FP32 data rows x D, FP32 targets rows x K) and, when
 * grads_host is not NULL, its gradient (vlhip_gate_num_params floats) */
VLHIP_API double vlhip_gate_gradients (void * gate, void const * data, void const * targets, size_t rows,
                                       float * grads_host) ;
/* steps Adam steps (lr, betas 0.9 / 0.999) on consecutive batches of numData
 * FP32 descriptors against the GMM's exact posteriors (gmm.c:711-798);
 * returns the last batch's loss, loss_trace_host (optional) every step's */
VLHIP_API double vlhip_gate_train (void * gate, void const * data, size_t numData, void const * means,
                                   void const * covariances, void const * priors, size_t steps, size_t batch,
                                   float lr, float * loss_trace_host) ;
/* per descriptor: the M most probable components of the gate (idx, rows x M,
 * -1 when screened out), their predicted mass (mass), and conf = mass >= need */
VLHIP_API void   vlhip_gate_select (void * gate, void const * data, size_t rows, int M, float need, int * idx,
                                    float * mass, unsigned char * conf) ;
VLHIP_API void   vlhip_gate_logits (void * gate, void const * data, size_t rows, float * logits_host) ;

#endif /* VLHIP_H */
