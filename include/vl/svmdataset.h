/* svmdataset.h -- data sets of the SVM solver (drop-in for VLFeat 0.9.21
 * vl/svmdataset.h).
 *
 * Same declarations as vl/svmdataset.h:37-76: a VlSvmDataset wraps a dense
 * float or double array of numData points of `dimension` values (point i at
 * data + i*dimension), optionally expanded on the fly by a homogeneous kernel
 * map (then each value becomes 2*order+1 features, laid out value by value).
 * The inner-product and accumulate functions returned here are the host
 * routines of svmdataset.c:331-401: sum over the point's features of
 * feature * model in double, in feature order, and model += feature * multiplier.
 * vl_svm_train recognises them and runs the solver on the GPU instead of
 * calling them.
 */
#ifndef VL_SVMDATASET_H
#define VL_SVMDATASET_H

#include "generic.h"
#include "homkermap.h"

struct VlSvm_ ;

/* vl/svmdataset.h:44-48: the solver's data abstraction */
typedef double (*VlSvmInnerProductFunction)(const void *data,
                                            vl_uindex element,
                                            double *model) ;
typedef void (*VlSvmAccumulateFunction)(const void *data,
                                        vl_uindex element,
                                        double *model,
                                        double multiplier) ;
typedef void (*VlSvmDiagnosticFunction) (struct VlSvm_ *svm, void *data) ;
typedef double (*VlSvmLossFunction) (double inner, double label) ;
typedef double (*VlSvmDcaUpdateFunction) (double alpha, double inner, double norm2, double label) ;

typedef struct VlSvmDataset_ VlSvmDataset ;

/* vl/svmdataset.h:55-76 */
VL_EXPORT VlSvmDataset* vl_svmdataset_new (vl_type dataType, void *data, vl_size dimension, vl_size numData) ;
VL_EXPORT void vl_svmdataset_delete (VlSvmDataset * dataset) ;
VL_EXPORT void vl_svmdataset_set_homogeneous_kernel_map (VlSvmDataset * self,
                                                         VlHomogeneousKernelMap * hom) ;
VL_EXPORT void* vl_svmdataset_get_data (VlSvmDataset const *self) ;
VL_EXPORT vl_size vl_svmdataset_get_num_data (VlSvmDataset const *self) ;
VL_EXPORT vl_size vl_svmdataset_get_dimension (VlSvmDataset const *self) ;
/* declared by the reference without a definition; defined here as the map
 * object and the number of features per value (1 without a map) */
VL_EXPORT void* vl_svmdataset_get_map (VlSvmDataset const *self) ;
VL_EXPORT vl_size vl_svmdataset_get_mapDim (VlSvmDataset const *self) ;
VL_EXPORT VlSvmAccumulateFunction vl_svmdataset_get_accumulate_function (VlSvmDataset const *self) ;
VL_EXPORT VlSvmInnerProductFunction vl_svmdataset_get_inner_product_function (VlSvmDataset const * self) ;
VL_EXPORT VlHomogeneousKernelMap * vl_svmdataset_get_homogeneous_kernel_map (VlSvmDataset const * self) ;

#endif /* VL_SVMDATASET_H */
