/* svmdataset.c -- data sets of the SVM solver (vl/svmdataset.c).
 *
 * A dataset is a view of the caller's dense float or double array (it owns
 * only the buffer of one mapped value).  The host inner-product and
 * accumulate routines follow svmdataset.c:331-401: the point's features are
 * visited in order, each converted to double and multiplied by the model
 * value, the products added one by one in double (model += feature *
 * multiplier for the accumulation); with a homogeneous kernel map the
 * 2*order+1 features of each value come from vl_homogeneouskernelmap_evaluate
 * into the dataset's buffer first.  vl_svm_train recognises these routines
 * (svm_dataset_is_builtin below) and runs its solver on the GPU.
 */
#include "vl/svmdataset.h"
#include "svm_private.h"

#include <assert.h>
#include <string.h>

struct VlSvmDataset_ {
  vl_type dataType ;
  void * data ;
  vl_size numData ;
  vl_size dimension ;
  VlHomogeneousKernelMap * hom ;
  void * homBuffer ;         /* 2*order+1 values of the data type */
  vl_size homDimension ;
} ;

VL_EXPORT VlSvmDataset *
vl_svmdataset_new (vl_type dataType, void * data, vl_size dimension, vl_size numData)
{
  VlSvmDataset * self ;
  assert(dataType == VL_TYPE_DOUBLE || dataType == VL_TYPE_FLOAT) ;
  assert(data) ;
  self = (VlSvmDataset*) vl_calloc(1, sizeof(VlSvmDataset)) ;
  if (self == NULL) return NULL ;
  self->dataType = dataType ;
  self->data = data ;
  self->dimension = dimension ;
  self->numData = numData ;
  return self ;
}

VL_EXPORT void
vl_svmdataset_delete (VlSvmDataset * self)
{
  if (self->homBuffer) vl_free(self->homBuffer) ;
  vl_free(self) ;
}

VL_EXPORT void * vl_svmdataset_get_data (VlSvmDataset const * self) { return self->data ; }
VL_EXPORT vl_size vl_svmdataset_get_num_data (VlSvmDataset const * self) { return self->numData ; }

VL_EXPORT vl_size
vl_svmdataset_get_dimension (VlSvmDataset const * self)
{
  return self->hom ? self->dimension * vl_homogeneouskernelmap_get_dimension(self->hom) : self->dimension ;
}

VL_EXPORT void * vl_svmdataset_get_map (VlSvmDataset const * self) { return self->hom ; }
VL_EXPORT vl_size vl_svmdataset_get_mapDim (VlSvmDataset const * self) { return self->hom ? self->homDimension : 1 ; }

VL_EXPORT VlHomogeneousKernelMap *
vl_svmdataset_get_homogeneous_kernel_map (VlSvmDataset const * self)
{
  assert(self) ;
  return self->hom ;
}

VL_EXPORT void
vl_svmdataset_set_homogeneous_kernel_map (VlSvmDataset * self, VlHomogeneousKernelMap * hom)
{
  assert(self) ;
  self->hom = hom ;
  self->homDimension = 0 ;
  if (self->homBuffer) {
    vl_free(self->homBuffer) ;
    self->homBuffer = NULL ;
  }
  if (hom) {
    self->homDimension = vl_homogeneouskernelmap_get_dimension(hom) ;
    self->homBuffer = vl_calloc(self->homDimension, vl_get_type_size(self->dataType)) ;
  }
}

/* ---- host data functions ----------------------------------------------- */

#define SVMDS_PLAIN(SFX, T)                                                            \
static double                                                                          \
svmds_inner_##SFX (const void * ds, vl_uindex element, double * model)                 \
{                                                                                      \
  VlSvmDataset const * self = (VlSvmDataset const*) ds ;                               \
  T const * x = (T const*) self->data + self->dimension * element ;                    \
  double s = 0 ;                                                                       \
  vl_uindex d ;                                                                        \
  for (d = 0 ; d < self->dimension ; ++d) s += x[d] * model[d] ;                       \
  return s ;                                                                           \
}                                                                                      \
static void                                                                            \
svmds_accumulate_##SFX (const void * ds, vl_uindex element, double * model, double m)  \
{                                                                                      \
  VlSvmDataset const * self = (VlSvmDataset const*) ds ;                               \
  T const * x = (T const*) self->data + self->dimension * element ;                    \
  vl_uindex d ;                                                                        \
  for (d = 0 ; d < self->dimension ; ++d) model[d] += x[d] * m ;                       \
}
SVMDS_PLAIN(f, float)
SVMDS_PLAIN(d, double)

static double
svmds_inner_hom_f (const void * ds, vl_uindex element, double * model)
{
  VlSvmDataset const * self = (VlSvmDataset const*) ds ;
  float const * x = (float const*) self->data + self->dimension * element ;
  float * buf = (float*) self->homBuffer ;
  double s = 0 ;
  vl_uindex d, j ;
  for (d = 0 ; d < self->dimension ; ++d) {
    vl_homogeneouskernelmap_evaluate_f(self->hom, buf, 1, x[d]) ;
    for (j = 0 ; j < self->homDimension ; ++j) s += buf[j] * *model++ ;
  }
  return s ;
}

static double
svmds_inner_hom_d (const void * ds, vl_uindex element, double * model)
{
  VlSvmDataset const * self = (VlSvmDataset const*) ds ;
  double const * x = (double const*) self->data + self->dimension * element ;
  double * buf = (double*) self->homBuffer ;
  double s = 0 ;
  vl_uindex d, j ;
  for (d = 0 ; d < self->dimension ; ++d) {
    vl_homogeneouskernelmap_evaluate_d(self->hom, buf, 1, x[d]) ;
    for (j = 0 ; j < self->homDimension ; ++j) s += buf[j] * *model++ ;
  }
  return s ;
}

static void
svmds_accumulate_hom_f (const void * ds, vl_uindex element, double * model, double m)
{
  VlSvmDataset const * self = (VlSvmDataset const*) ds ;
  float const * x = (float const*) self->data + self->dimension * element ;
  float * buf = (float*) self->homBuffer ;
  vl_uindex d, j ;
  for (d = 0 ; d < self->dimension ; ++d) {
    vl_homogeneouskernelmap_evaluate_f(self->hom, buf, 1, x[d]) ;
    for (j = 0 ; j < self->homDimension ; ++j) *model++ += buf[j] * m ;
  }
}

static void
svmds_accumulate_hom_d (const void * ds, vl_uindex element, double * model, double m)
{
  VlSvmDataset const * self = (VlSvmDataset const*) ds ;
  double const * x = (double const*) self->data + self->dimension * element ;
  double * buf = (double*) self->homBuffer ;
  vl_uindex d, j ;
  for (d = 0 ; d < self->dimension ; ++d) {
    vl_homogeneouskernelmap_evaluate_d(self->hom, buf, 1, x[d]) ;
    for (j = 0 ; j < self->homDimension ; ++j) *model++ += buf[j] * m ;
  }
}

VL_EXPORT VlSvmAccumulateFunction
vl_svmdataset_get_accumulate_function (VlSvmDataset const * self)
{
  if (self->dataType == VL_TYPE_FLOAT) return self->hom ? svmds_accumulate_hom_f : svmds_accumulate_f ;
  return self->hom ? svmds_accumulate_hom_d : svmds_accumulate_d ;
}

VL_EXPORT VlSvmInnerProductFunction
vl_svmdataset_get_inner_product_function (VlSvmDataset const * self)
{
  if (self->dataType == VL_TYPE_FLOAT) return self->hom ? svmds_inner_hom_f : svmds_inner_f ;
  return self->hom ? svmds_inner_hom_d : svmds_inner_d ;
}

/* ---- for the solver ------------------------------------------------------ */

int
svm_dataset_is_builtin (VlSvmDataset const * self, VlSvmInnerProductFunction inner, VlSvmAccumulateFunction acc)
{
  return self && inner == vl_svmdataset_get_inner_product_function(self) &&
         acc == vl_svmdataset_get_accumulate_function(self) ;
}

void
svm_dataset_describe (VlSvmDataset const * self, vl_type * dataType, void const ** data,
                      vl_size * dimension, vl_size * numData, VlHomogeneousKernelMap const ** hom)
{
  *dataType = self->dataType ;
  *data = self->data ;
  *dimension = self->dimension ;
  *numData = self->numData ;
  *hom = self->hom ;
}
