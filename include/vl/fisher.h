/* fisher.h -- Fisher vector encoding (drop-in for VLFeat 0.9.21 vl/fisher.h).
 *
 * vl_fisher_encode replaces vl/fisher.h:48-54 (implementation fisher.c:548-579
 * and 363-488): same argument order, same flags, same return value (number of
 * (component, datum) terms with prior >= 1e-6 and posterior >= 1e-6), same
 * abort() on an unknown data type.  The work runs on the GPU; enc is a
 * caller-allocated host buffer of 2*dimension*numClusters elements laid out
 * [u_0 .. u_{K-1}, v_0 .. v_{K-1}].
 */
#ifndef VL_FISHER_H
#define VL_FISHER_H

#include "generic.h"

#define VL_FISHER_FLAG_SQUARE_ROOT  (0x1 << 0)
#define VL_FISHER_FLAG_NORMALIZED   (0x1 << 1)
#define VL_FISHER_FLAG_IMPROVED     (VL_FISHER_FLAG_NORMALIZED|VL_FISHER_FLAG_SQUARE_ROOT)
#define VL_FISHER_FLAG_FAST         (0x1 << 2)

VL_EXPORT vl_size vl_fisher_encode
(void * enc, vl_type dataType,
 void const * means, vl_size dimension, vl_size numClusters,
 void const * covariances,
 void const * priors,
 void const * data, vl_size numData,
 int flags) ;

/* Batched form (no counterpart in the reference, whose callers loop over
 * images, e.g. apps/recognition/encodeImage.m:60-67): numImages images whose
 * descriptors are stored back to back in data, image i owning rows
 * [offsets[i], offsets[i+1]).  enc receives numImages consecutive encodings of
 * 2*dimension*numClusters elements; numTerms (optional) one count per image.
 * Each image's result equals vl_fisher_encode on that image alone. */
VL_EXPORT void vl_fisher_encode_batch
(void * enc, vl_type dataType,
 void const * means, vl_size dimension, vl_size numClusters,
 void const * covariances,
 void const * priors,
 void const * data, vl_uint64 const * offsets, vl_size numImages,
 int flags, vl_size * numTerms) ;

/* Arithmetic of the float encoder for the calling thread (no counterpart in
 * the reference, which has one float path).  VL_FISHER_PRECISION_FP32, the
 * default, follows the reference's float arithmetic and is the parity path;
 * the SPLIT16 settings run both contractions on the GPU's matrix cores with
 * compensated 16-bit operands (relative error against an FP64 evaluation of
 * order 1e-5 for FP16, 1e-4 for BF16).  Double data always uses FP64. */
#define VL_FISHER_PRECISION_FP32     0
#define VL_FISHER_PRECISION_BF16X3   1
#define VL_FISHER_PRECISION_GATED    2
#define VL_FISHER_PRECISION_F16X3    3
/* Plain FP32 operands on the FP32 matrix cores (exact FP32 products, FP32
 * accumulation, the posteriors never leave registers), D <= 128. */
#define VL_FISHER_PRECISION_FP32_MFMA 4
VL_EXPORT void vl_fisher_set_precision (int precision) ;
VL_EXPORT int vl_fisher_get_precision (void) ;

/* Learned gate of VL_FISHER_PRECISION_GATED (no counterpart in the
 * reference): a handle from vlhip_gate_new (vlhip.h), trained on the GMM's
 * posteriors.  Descriptors are taken in groups of 32 per image; a group whose
 * every descriptor has predicted mass >= need on the gate's `candidates`
 * (<= 16) most probable components gets exact posteriors over those
 * candidates only; all other groups use the dense F16X3 E-step.  The gate is
 * shared by all threads of the process; NULL clears it. */
VL_EXPORT void vl_fisher_set_gate (void * gate, int candidates, float need) ;

#endif /* VL_FISHER_H */
