/* svm_private.h -- what the SVM solver needs from the dataset and
 * homogeneous-kernel-map modules (library-internal). */
#ifndef VL_SVM_PRIVATE_H
#define VL_SVM_PRIVATE_H

#include "vl/svmdataset.h"
#include "vl/homkermap.h"

#define VL_HIDDEN __attribute__((visibility("hidden")))

/* 1 when inner / acc are the dataset's own routines */
VL_HIDDEN int svm_dataset_is_builtin (VlSvmDataset const * self, VlSvmInnerProductFunction inner,
                                      VlSvmAccumulateFunction acc) ;
VL_HIDDEN void svm_dataset_describe (VlSvmDataset const * self, vl_type * dataType, void const ** data,
                                     vl_size * dimension, vl_size * numData, VlHomogeneousKernelMap const ** hom) ;
/* the map of n values of dIn into dOut (n x (2 order + 1)), both device
 * arrays of dataType, bitwise vl_homogeneouskernelmap_evaluate_{f,d} */
VL_HIDDEN void homkermap_evaluate_device (VlHomogeneousKernelMap const * self, vl_type dataType, void * dOut,
                                          void const * dIn, vl_size n) ;

#endif
