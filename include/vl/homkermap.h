/* homkermap.h -- homogeneous kernel map (drop-in for VLFeat 0.9.21 vl/homkermap.h).
 *
 * The explicit feature map of the additive intersection, chi2 and
 * Jensen-Shannon kernels that the recognition application applies to its
 * encodings before the linear SVM (apps/recognition/traintest.m:99).  Same
 * types, enum values and entry points as vl/homkermap.h:22-84: the map
 * object is built on the host (its interpolation table, vl/homkermap.c:
 * 326-426); vl_homogeneouskernelmap_evaluate_{d,f} map one value, as the
 * reference does, and the array form below maps whole matrices on the GPU.
 */
#ifndef VL_HOMKERMAP_H
#define VL_HOMKERMAP_H

#include "generic.h"

typedef enum {
  VlHomogeneousKernelIntersection = 0,
  VlHomogeneousKernelChi2,
  VlHomogeneousKernelJS
} VlHomogeneousKernelType ;

typedef enum {
  VlHomogeneousKernelMapWindowUniform = 0,
  VlHomogeneousKernelMapWindowRectangular = 1
} VlHomogeneousKernelMapWindowType ;

struct _VlHomogeneousKernelMap ;
typedef struct _VlHomogeneousKernelMap VlHomogeneousKernelMap ;

/* homkermap.h:45-51: period < 0 selects the reference's default period */
VL_EXPORT VlHomogeneousKernelMap *
vl_homogeneouskernelmap_new (VlHomogeneousKernelType kernelType,
                             double gamma,
                             vl_size order,
                             double period,
                             VlHomogeneousKernelMapWindowType windowType) ;
VL_EXPORT void
vl_homogeneouskernelmap_delete (VlHomogeneousKernelMap * self) ;

/* homkermap.h:57-67: the 2*order+1 values of the map of x, written to
 * destination[0], destination[stride], ... */
VL_EXPORT void
vl_homogeneouskernelmap_evaluate_d (VlHomogeneousKernelMap const * self,
                                    double * destination,
                                    vl_size stride,
                                    double x) ;
VL_EXPORT void
vl_homogeneouskernelmap_evaluate_f (VlHomogeneousKernelMap const * self,
                                    float * destination,
                                    vl_size stride,
                                    double x) ;

/* Array form (no counterpart in the reference, whose MEX front-end loops
 * over the elements, toolbox/misc/vl_homkermap.c): destination[i*(2*order+1)
 * + j] is value j of the map of data[i], i < numData, equal to
 * vl_homogeneouskernelmap_evaluate_{f,d}(self, destination + i*(2*order+1),
 * 1, data[i]) bit for bit; dataType is VL_TYPE_FLOAT or VL_TYPE_DOUBLE for
 * both arrays (host arrays; the map runs on the GPU). */
VL_EXPORT void
vl_homogeneouskernelmap_evaluate_array (VlHomogeneousKernelMap const * self,
                                        vl_type dataType,
                                        void * destination,
                                        void const * data,
                                        vl_size numData) ;

/* homkermap.h:73-84 */
VL_EXPORT vl_size
vl_homogeneouskernelmap_get_order (VlHomogeneousKernelMap const * self) ;
VL_EXPORT vl_size
vl_homogeneouskernelmap_get_dimension (VlHomogeneousKernelMap const * self) ;
VL_EXPORT VlHomogeneousKernelType
vl_homogeneouskernelmap_get_kernel_type (VlHomogeneousKernelMap const * self) ;
VL_EXPORT VlHomogeneousKernelMapWindowType
vl_homogeneouskernelmap_get_window_type (VlHomogeneousKernelMap const * self) ;

#endif /* VL_HOMKERMAP_H */
