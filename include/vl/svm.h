/* svm.h -- linear SVM solvers (drop-in for VLFeat 0.9.21 vl/svm.h).
 *
 * Same types, enum values and entry points as vl/svm.h:32-196
 * (implementation svm.c:957-2172): vl_svm_new (double data),
 * vl_svm_new_with_dataset, vl_svm_new_with_abstract_data, the getters and
 * setters, vl_svm_set_loss, vl_svm_train (SDCA, SGD, or evaluation only) and
 * the five predefined losses.  Defaults are the reference's: hinge loss,
 * bias multiplier 1, epsilon 1e-2, bias learning rate 0.01, diagnostic
 * frequency = numData, maxNumIterations = max(numData, ceil(10 / lambda)).
 *
 * Where the solver runs: with a dataset of float or double points (with or
 * without a homogeneous kernel map) and one of the predefined losses, every
 * iteration of SDCA or SGD runs on the GPU (one persistent kernel per run of
 * iterations between diagnostics or epoch starts); the per-epoch
 * permutations are drawn on the host from vl_get_rand() in the reference's
 * order, and the statistics are formed on the host from scores computed on
 * the GPU.  With abstract data or user-set loss / data functions the solver
 * calls those functions from the host, as the reference does.
 */
#ifndef VL_SVM_H
#define VL_SVM_H

#include "generic.h"
#include "svmdataset.h"

struct VlSvm_ ;
typedef struct VlSvm_ VlSvm ;

typedef enum
{
  VlSvmSolverNone = 0,
  VlSvmSolverSgd = 1,
  VlSvmSolverSdca
} VlSvmSolverType ;

typedef enum
{
  VlSvmLossHinge = 0,
  VlSvmLossHinge2 = 1,
  VlSvmLossL1,
  VlSvmLossL2,
  VlSvmLossLogistic
} VlSvmLossType ;

typedef enum
{
  VlSvmStatusTraining = 1,
  VlSvmStatusConverged,
  VlSvmStatusMaxNumIterationsReached
} VlSvmSolverStatus ;

typedef struct VlSvmStatistics_ {
  VlSvmSolverStatus status ;
  vl_size iteration ;
  vl_size epoch ;
  double objective ;
  double regularizer ;
  double loss ;
  double dualObjective ;
  double dualLoss ;
  double dualityGap ;
  double scoresVariation ;
  double elapsedTime ;
} VlSvmStatistics ;

/* vl/svm.h:92-110 */
VL_EXPORT VlSvm * vl_svm_new (VlSvmSolverType type,
                              double const * data,
                              vl_size dimension,
                              vl_size numData,
                              double const * labels,
                              double lambda) ;
VL_EXPORT VlSvm * vl_svm_new_with_dataset (VlSvmSolverType type,
                                           VlSvmDataset * dataset,
                                           double const * labels,
                                           double lambda) ;
VL_EXPORT VlSvm * vl_svm_new_with_abstract_data (VlSvmSolverType type,
                                                 void * data,
                                                 vl_size dimension,
                                                 vl_size numData,
                                                 double const * labels,
                                                 double lambda) ;
VL_EXPORT void vl_svm_delete (VlSvm * self) ;

/* vl/svm.h:116-130 */
VL_EXPORT VlSvmStatistics const * vl_svm_get_statistics (VlSvm const *self) ;
VL_EXPORT double const * vl_svm_get_model (VlSvm const *self) ;
VL_EXPORT double vl_svm_get_bias (VlSvm const *self) ;
VL_EXPORT vl_size vl_svm_get_dimension (VlSvm *self) ;
VL_EXPORT vl_size vl_svm_get_num_data (VlSvm *self) ;
VL_EXPORT double vl_svm_get_epsilon (VlSvm const *self) ;
VL_EXPORT double vl_svm_get_bias_learning_rate (VlSvm const *self) ;
VL_EXPORT vl_size vl_svm_get_max_num_iterations (VlSvm const *self) ;
VL_EXPORT vl_size vl_svm_get_diagnostic_frequency (VlSvm const *self) ;
VL_EXPORT VlSvmSolverType vl_svm_get_solver (VlSvm const *self) ;
VL_EXPORT double vl_svm_get_bias_multiplier (VlSvm const *self) ;
VL_EXPORT double vl_svm_get_lambda (VlSvm const *self) ;
/* defined (not declared) by the reference, svm.c:1288 */
VL_EXPORT void vl_svm_set_lambda (VlSvm *self, double lambda) ;
VL_EXPORT vl_size vl_svm_get_iteration_number (VlSvm const *self) ;
VL_EXPORT double const * vl_svm_get_scores (VlSvm const *self) ;
VL_EXPORT double const * vl_svm_get_weights (VlSvm const *self) ;

/* vl/svm.h:135-151 */
VL_EXPORT void vl_svm_set_epsilon (VlSvm *self, double epsilon) ;
VL_EXPORT void vl_svm_set_bias_learning_rate (VlSvm *self, double rate) ;
VL_EXPORT void vl_svm_set_max_num_iterations (VlSvm *self, vl_size maxNumIterations) ;
VL_EXPORT void vl_svm_set_diagnostic_frequency (VlSvm *self, vl_size f) ;
VL_EXPORT void vl_svm_set_bias_multiplier (VlSvm *self, double b) ;
VL_EXPORT void vl_svm_set_model (VlSvm *self, double const *model) ;
VL_EXPORT void vl_svm_set_bias (VlSvm *self, double b) ;
VL_EXPORT void vl_svm_set_iteration_number (VlSvm *self, vl_uindex n) ;
VL_EXPORT void vl_svm_set_weights (VlSvm *self, double const *weights) ;
VL_EXPORT void vl_svm_set_diagnostic_function (VlSvm *self, VlSvmDiagnosticFunction f, void *data) ;
VL_EXPORT void vl_svm_set_loss_function (VlSvm *self, VlSvmLossFunction f) ;
VL_EXPORT void vl_svm_set_loss_derivative_function (VlSvm *self, VlSvmLossFunction f) ;
VL_EXPORT void vl_svm_set_conjugate_loss_function (VlSvm *self, VlSvmLossFunction f) ;
VL_EXPORT void vl_svm_set_dca_update_function (VlSvm *self, VlSvmDcaUpdateFunction f) ;
VL_EXPORT void vl_svm_set_data_functions (VlSvm *self, VlSvmInnerProductFunction inner, VlSvmAccumulateFunction acc) ;
VL_EXPORT void vl_svm_set_loss (VlSvm *self, VlSvmLossType loss) ;

/* vl/svm.h:156 */
VL_EXPORT void vl_svm_train (VlSvm * self) ;

/* vl/svm.h:160-196: the predefined losses.  The reference header names the
 * first two parameters (label, inner); they are (inner, label) as the solver
 * calls them (svm.c:1602-1858) */
VL_EXPORT double vl_svm_hinge_loss (double inner, double label) ;
VL_EXPORT double vl_svm_hinge_loss_derivative (double inner, double label) ;
VL_EXPORT double vl_svm_hinge_conjugate_loss (double u, double label) ;
VL_EXPORT double vl_svm_hinge_dca_update (double alpha, double inner, double norm2, double label) ;
VL_EXPORT double vl_svm_hinge2_loss (double inner, double label) ;
VL_EXPORT double vl_svm_hinge2_loss_derivative (double inner, double label) ;
VL_EXPORT double vl_svm_hinge2_conjugate_loss (double u, double label) ;
VL_EXPORT double vl_svm_hinge2_dca_update (double alpha, double inner, double norm2, double label) ;
VL_EXPORT double vl_svm_l1_loss (double inner, double label) ;
VL_EXPORT double vl_svm_l1_loss_derivative (double inner, double label) ;
VL_EXPORT double vl_svm_l1_conjugate_loss (double u, double label) ;
VL_EXPORT double vl_svm_l1_dca_update (double alpha, double inner, double norm2, double label) ;
VL_EXPORT double vl_svm_l2_loss (double inner, double label) ;
VL_EXPORT double vl_svm_l2_loss_derivative (double inner, double label) ;
VL_EXPORT double vl_svm_l2_conjugate_loss (double u, double label) ;
VL_EXPORT double vl_svm_l2_dca_update (double alpha, double inner, double norm2, double label) ;
VL_EXPORT double vl_svm_logistic_loss (double inner, double label) ;
VL_EXPORT double vl_svm_logistic_loss_derivative (double inner, double label) ;
VL_EXPORT double vl_svm_logistic_conjugate_loss (double u, double label) ;
VL_EXPORT double vl_svm_logistic_dca_update (double alpha, double inner, double norm2, double label) ;

#endif /* VL_SVM_H */
