/* svm.c -- linear SVM solvers (vl/svm.c): SDCA, SGD and evaluation.
 *
 * The object, its defaults, setters, the five predefined losses and the
 * statistics follow svm.c:957-1858.  Training (svm.c:1931-2172):
 *
 *  - a dataset of float or double points (with or without a homogeneous
 *    kernel map) together with one of the predefined loss sets is trained on
 *    the GPU: the points (mapped on the GPU when the dataset has a map), the
 *    labels, weights, model and dual variables are copied to the device once;
 *    the host draws each epoch's permutation from vl_get_rand() exactly when
 *    the reference draws it (vl_rand_permute_indexes at every iteration t
 *    with t % numData == 0) and the iterations between two such draws or two
 *    diagnostics run in one persistent launch (vlhip_svm_run); at every
 *    diagnostic the model (and the duals) come back to the host, the scores
 *    of all points are formed on the GPU and the statistics on the host in
 *    the reference's order before the diagnostic function is called;
 *  - abstract data, or loss / data functions set by the caller, are trained
 *    by the host loop below, which calls those functions as the reference
 *    does (they are host code).
 *
 * The GPU sums each inner product in another order than the reference's
 * one-by-one double sum; everything else is the same double arithmetic.
 */
#include "vl/svm.h"
#include "vl/mathop.h"
#include "vl/random.h"
#include "vlhip.h"
#include "svm_private.h"
#include "host_fatal.h"

#include <assert.h>
#include <math.h>
#include <string.h>

struct VlSvm_ {
  VlSvmSolverType solver ;
  vl_size dimension ;
  double * model ;
  double bias ;                /* the model's bias is bias * biasMultiplier */
  double biasMultiplier ;
  double lambda ;
  void const * data ;
  vl_size numData ;
  double const * labels ;
  double const * weights ;
  VlSvmDataset * dataset ;     /* the dataset behind data, if created from one */
  VlSvmDataset * ownDataset ;
  VlSvmDiagnosticFunction diagnosticFn ;
  void * diagnosticFnData ;
  vl_size diagnosticFrequency ;
  VlSvmLossFunction lossFn ;
  VlSvmLossFunction conjugateLossFn ;
  VlSvmLossFunction lossDerivativeFn ;
  VlSvmDcaUpdateFunction dcaUpdateFn ;
  VlSvmInnerProductFunction innerProductFn ;
  VlSvmAccumulateFunction accumulateFn ;
  vl_size iteration ;
  vl_size maxNumIterations ;
  double epsilon ;
  VlSvmStatistics statistics ;
  double * scores ;
  double biasLearningRate ;    /* SGD */
  double * alpha ;             /* SDCA duals */
} ;

/* ---- construction ------------------------------------------------------- */

VL_EXPORT VlSvm *
vl_svm_new_with_abstract_data (VlSvmSolverType solver, void * data, vl_size dimension, vl_size numData,
                               double const * labels, double lambda)
{
  VlSvm * self ;
  assert(dimension >= 1) ;
  assert(numData >= 1) ;
  assert(labels) ;
  self = (VlSvm*) vl_calloc(1, sizeof(VlSvm)) ;
  if (!self) return NULL ;
  self->solver = solver ;
  self->dimension = dimension ;
  self->biasMultiplier = 1.0 ;
  self->lambda = lambda ;
  self->data = data ;
  self->numData = numData ;
  self->labels = labels ;
  self->diagnosticFrequency = numData ;
  self->lossFn = vl_svm_hinge_loss ;
  self->conjugateLossFn = vl_svm_hinge_conjugate_loss ;
  self->lossDerivativeFn = vl_svm_hinge_loss_derivative ;
  self->dcaUpdateFn = vl_svm_hinge_dca_update ;
  /* svm.c:1052: max(numData, ceil(10 / lambda)), the ceiling taken in float */
  self->maxNumIterations = (vl_size) VL_MAX((double) numData, (double)(long) ceilf((float)(10.0 / lambda))) ;
  self->epsilon = 1e-2 ;
  self->biasLearningRate = 0.01 ;
  self->model = (double*) vl_calloc(dimension, sizeof(double)) ;
  self->scores = (double*) vl_calloc(numData, sizeof(double)) ;
  if (solver == VlSvmSolverSdca) self->alpha = (double*) vl_calloc(numData, sizeof(double)) ;
  if (!self->model || !self->scores || (solver == VlSvmSolverSdca && !self->alpha)) {
    vl_free(self->model) ;
    vl_free(self->scores) ;
    vl_free(self->alpha) ;
    vl_free(self) ;
    return NULL ;
  }
  return self ;
}

VL_EXPORT VlSvm *
vl_svm_new_with_dataset (VlSvmSolverType solver, VlSvmDataset * dataset, double const * labels, double lambda)
{
  VlSvm * self = vl_svm_new_with_abstract_data(solver, dataset, vl_svmdataset_get_dimension(dataset),
                                               vl_svmdataset_get_num_data(dataset), labels, lambda) ;
  if (!self) return NULL ;
  self->dataset = dataset ;
  vl_svm_set_data_functions(self, vl_svmdataset_get_inner_product_function(dataset),
                            vl_svmdataset_get_accumulate_function(dataset)) ;
  return self ;
}

VL_EXPORT VlSvm *
vl_svm_new (VlSvmSolverType solver, double const * data, vl_size dimension, vl_size numData,
            double const * labels, double lambda)
{
  VlSvmDataset * dataset = vl_svmdataset_new(VL_TYPE_DOUBLE, (void*) data, dimension, numData) ;
  VlSvm * self = vl_svm_new_with_dataset(solver, dataset, labels, lambda) ;
  if (self) self->ownDataset = dataset ;
  else vl_svmdataset_delete(dataset) ;
  return self ;
}

VL_EXPORT void
vl_svm_delete (VlSvm * self)
{
  vl_free(self->model) ;
  vl_free(self->alpha) ;
  vl_free(self->scores) ;
  if (self->ownDataset) vl_svmdataset_delete(self->ownDataset) ;
  vl_free(self) ;
}

/* ---- parameters (svm.c:1123-1549) ---------------------------------------- */

VL_EXPORT void vl_svm_set_epsilon (VlSvm * self, double e) { assert(e >= 0) ; self->epsilon = e ; }
VL_EXPORT double vl_svm_get_epsilon (VlSvm const * self) { return self->epsilon ; }
VL_EXPORT void vl_svm_set_bias_learning_rate (VlSvm * self, double r) { assert(r > 0) ; self->biasLearningRate = r ; }
VL_EXPORT double vl_svm_get_bias_learning_rate (VlSvm const * self) { return self->biasLearningRate ; }
VL_EXPORT void vl_svm_set_bias_multiplier (VlSvm * self, double b) { assert(b >= 0) ; self->biasMultiplier = b ; }
VL_EXPORT double vl_svm_get_bias_multiplier (VlSvm const * self) { return self->biasMultiplier ; }
VL_EXPORT void vl_svm_set_iteration_number (VlSvm * self, vl_uindex n) { self->iteration = n ; }
VL_EXPORT vl_size vl_svm_get_iteration_number (VlSvm const * self) { return self->iteration ; }
VL_EXPORT void vl_svm_set_max_num_iterations (VlSvm * self, vl_size n) { self->maxNumIterations = n ; }
VL_EXPORT vl_size vl_svm_get_max_num_iterations (VlSvm const * self) { return self->maxNumIterations ; }
VL_EXPORT void vl_svm_set_diagnostic_frequency (VlSvm * self, vl_size f) { assert(f > 0) ; self->diagnosticFrequency = f ; }
VL_EXPORT vl_size vl_svm_get_diagnostic_frequency (VlSvm const * self) { return self->diagnosticFrequency ; }
VL_EXPORT VlSvmSolverType vl_svm_get_solver (VlSvm const * self) { return self->solver ; }
VL_EXPORT void vl_svm_set_lambda (VlSvm * self, double lambda) { assert(lambda >= 0) ; self->lambda = lambda ; }
VL_EXPORT double vl_svm_get_lambda (VlSvm const * self) { return self->lambda ; }
VL_EXPORT void vl_svm_set_weights (VlSvm * self, double const * w) { self->weights = w ; }
VL_EXPORT double const * vl_svm_get_weights (VlSvm const * self) { return self->weights ; }
VL_EXPORT vl_size vl_svm_get_dimension (VlSvm * self) { return self->dimension ; }
VL_EXPORT vl_size vl_svm_get_num_data (VlSvm * self) { return self->numData ; }
VL_EXPORT double const * vl_svm_get_model (VlSvm const * self) { return self->model ; }
VL_EXPORT VlSvmStatistics const * vl_svm_get_statistics (VlSvm const * self) { return &self->statistics ; }
VL_EXPORT double const * vl_svm_get_scores (VlSvm const * self) { return self->scores ; }
VL_EXPORT double vl_svm_get_bias (VlSvm const * self) { return self->bias * self->biasMultiplier ; }

VL_EXPORT void
vl_svm_set_model (VlSvm * self, double const * model)
{
  assert(model) ;
  memcpy(self->model, model, sizeof(double) * self->dimension) ;
}

VL_EXPORT void
vl_svm_set_bias (VlSvm * self, double b)
{
  /* the stored bias is in units of the bias multiplier (svm.c:1406-1412) */
  if (self->biasMultiplier) self->bias = b / self->biasMultiplier ;
}

VL_EXPORT void
vl_svm_set_diagnostic_function (VlSvm * self, VlSvmDiagnosticFunction f, void * data)
{
  self->diagnosticFn = f ;
  self->diagnosticFnData = data ;
}

VL_EXPORT void
vl_svm_set_data_functions (VlSvm * self, VlSvmInnerProductFunction inner, VlSvmAccumulateFunction acc)
{
  assert(inner) ;
  assert(acc) ;
  self->innerProductFn = inner ;
  self->accumulateFn = acc ;
}

VL_EXPORT void vl_svm_set_loss_function (VlSvm * self, VlSvmLossFunction f) { self->lossFn = f ; }
VL_EXPORT void vl_svm_set_loss_derivative_function (VlSvm * self, VlSvmLossFunction f) { self->lossDerivativeFn = f ; }
VL_EXPORT void vl_svm_set_conjugate_loss_function (VlSvm * self, VlSvmLossFunction f) { self->conjugateLossFn = f ; }
VL_EXPORT void vl_svm_set_dca_update_function (VlSvm * self, VlSvmDcaUpdateFunction f) { self->dcaUpdateFn = f ; }

/* ---- the predefined losses (svm.c:1602-1858) ----------------------------- */

VL_EXPORT double vl_svm_hinge_loss (double inner, double label) { return VL_MAX(1 - label * inner, 0.0) ; }
VL_EXPORT double vl_svm_hinge_loss_derivative (double inner, double label) { return label * inner < 1.0 ? -label : 0.0 ; }
VL_EXPORT double
vl_svm_hinge_conjugate_loss (double u, double label)
{
  double z = label * u ;
  return (-1 <= z && z <= 0) ? label * u : VL_INFINITY_D ;
}
VL_EXPORT double
vl_svm_hinge_dca_update (double alpha, double inner, double norm2, double label)
{
  double palpha = (label - inner) / norm2 + alpha ;
  return label * VL_MAX(0, VL_MIN(1, label * palpha)) - alpha ;
}

VL_EXPORT double
vl_svm_hinge2_loss (double inner, double label)
{
  double z = VL_MAX(1 - label * inner, 0.0) ;
  return z * z ;
}
VL_EXPORT double
vl_svm_hinge2_loss_derivative (double inner, double label)
{
  return label * inner < 1.0 ? 2 * (inner - label) : 0 ;
}
VL_EXPORT double
vl_svm_hinge2_conjugate_loss (double u, double label)
{
  return label * u <= 0 ? (label + u / 4) * u : VL_INFINITY_D ;
}
VL_EXPORT double
vl_svm_hinge2_dca_update (double alpha, double inner, double norm2, double label)
{
  double palpha = (label - inner - 0.5 * alpha) / (norm2 + 0.5) + alpha ;
  return label * VL_MAX(0, label * palpha) - alpha ;
}

VL_EXPORT double vl_svm_l1_loss (double inner, double label) { return vl_abs_d(label - inner) ; }
VL_EXPORT double vl_svm_l1_loss_derivative (double inner, double label) { return label > inner ? -1.0 : 1.0 ; }
VL_EXPORT double vl_svm_l1_conjugate_loss (double u, double label) { return vl_abs_d(u) <= 1 ? label * u : VL_INFINITY_D ; }
VL_EXPORT double
vl_svm_l1_dca_update (double alpha, double inner, double norm2, double label)
{
  if (vl_abs_d(alpha) <= 1) {
    double palpha = (label - inner) / norm2 + alpha ;
    return VL_MAX(-1.0, VL_MIN(1.0, palpha)) - alpha ;
  }
  return VL_INFINITY_D ;
}

VL_EXPORT double vl_svm_l2_loss (double inner, double label) { double z = label - inner ; return z * z ; }
VL_EXPORT double vl_svm_l2_loss_derivative (double inner, double label) { return -2 * (label - inner) ; }
VL_EXPORT double vl_svm_l2_conjugate_loss (double u, double label) { return (label + u / 4) * u ; }
VL_EXPORT double
vl_svm_l2_dca_update (double alpha, double inner, double norm2, double label)
{
  return (label - inner - 0.5 * alpha) / (norm2 + 0.5) ;
}

VL_EXPORT double
vl_svm_logistic_loss (double inner, double label)
{
  double z = label * inner ;
  return z >= 0 ? log(1.0 + exp(-z)) : -z + log(exp(z) + 1.0) ;
}
VL_EXPORT double
vl_svm_logistic_loss_derivative (double inner, double label)
{
  double z = label * inner ;
  double t = 1 / (1 + exp(-z)) ;
  return label * (t - 1) ;
}
static double svm_xlogx (double x) { return x <= 1e-10 ? 0 : x * log(x) ; }
VL_EXPORT double
vl_svm_logistic_conjugate_loss (double u, double label)
{
  double z = label * u ;
  return (-1 <= z && z <= 0) ? svm_xlogx(-z) + svm_xlogx(1 + z) : VL_INFINITY_D ;
}
VL_EXPORT double
vl_svm_logistic_dca_update (double alpha, double inner, double norm2, double label)
{
  /* the dual coordinate beta = label (alpha + delta) in [0, 1] minimises
   * norm2/2 beta^2 + label (inner - norm2 alpha) beta + beta log beta +
   * (1 - beta) log(1 - beta): five bisection steps on the derivative, then
   * up to two Newton steps kept inside (0, 1) */
  double beta1 = 0, beta2 = 1, beta = 0.5, df, ddf, der, dder ;
  int t ;
  for (t = 0 ; t < 5 ; ++t) {
    df = log(beta) - log(1 - beta) ;
    der = norm2 * beta + label * (inner - norm2 * alpha) + df ;
    if (der >= 0) beta2 = beta ; else beta1 = beta ;
    beta = 0.5 * (beta1 + beta2) ;
  }
  for (t = 0 ; (t < 2) & (beta > VL_EPSILON_D) & (beta < 1 - VL_EPSILON_D) ; ++t) {
    df = log(beta) - log(1 - beta) ;
    ddf = 1 / (beta * (1 - beta)) ;
    der = norm2 * beta + label * (inner - norm2 * alpha) + df ;
    dder = norm2 + ddf ;
    beta -= der / dder ;
    beta = VL_MAX(0, VL_MIN(1, beta)) ;
  }
  return label * beta - alpha ;
}

VL_EXPORT void
vl_svm_set_loss (VlSvm * self, VlSvmLossType loss)
{
  switch (loss) {
    case VlSvmLossHinge:
      self->lossFn = vl_svm_hinge_loss ; self->lossDerivativeFn = vl_svm_hinge_loss_derivative ;
      self->conjugateLossFn = vl_svm_hinge_conjugate_loss ; self->dcaUpdateFn = vl_svm_hinge_dca_update ;
      break ;
    case VlSvmLossHinge2:
      self->lossFn = vl_svm_hinge2_loss ; self->lossDerivativeFn = vl_svm_hinge2_loss_derivative ;
      self->conjugateLossFn = vl_svm_hinge2_conjugate_loss ; self->dcaUpdateFn = vl_svm_hinge2_dca_update ;
      break ;
    case VlSvmLossL1:
      self->lossFn = vl_svm_l1_loss ; self->lossDerivativeFn = vl_svm_l1_loss_derivative ;
      self->conjugateLossFn = vl_svm_l1_conjugate_loss ; self->dcaUpdateFn = vl_svm_l1_dca_update ;
      break ;
    case VlSvmLossL2:
      self->lossFn = vl_svm_l2_loss ; self->lossDerivativeFn = vl_svm_l2_loss_derivative ;
      self->conjugateLossFn = vl_svm_l2_conjugate_loss ; self->dcaUpdateFn = vl_svm_l2_dca_update ;
      break ;
    case VlSvmLossLogistic:
      self->lossFn = vl_svm_logistic_loss ; self->lossDerivativeFn = vl_svm_logistic_loss_derivative ;
      self->conjugateLossFn = vl_svm_logistic_conjugate_loss ; self->dcaUpdateFn = vl_svm_logistic_dca_update ;
      break ;
    default:
      vl_host_fatal("vl_svm_set_loss: unknown loss %d\n", (int) loss) ;
  }
}

/* index of the predefined loss set in use (VlSvmLossType), -1 if any of the
 * four functions was set by the caller */
static int
svm_builtin_loss (VlSvm const * self)
{
  static const struct {
    VlSvmLossFunction loss, derivative, conjugate ;
    VlSvmDcaUpdateFunction update ;
  } sets [5] = {
    {vl_svm_hinge_loss, vl_svm_hinge_loss_derivative, vl_svm_hinge_conjugate_loss, vl_svm_hinge_dca_update},
    {vl_svm_hinge2_loss, vl_svm_hinge2_loss_derivative, vl_svm_hinge2_conjugate_loss, vl_svm_hinge2_dca_update},
    {vl_svm_l1_loss, vl_svm_l1_loss_derivative, vl_svm_l1_conjugate_loss, vl_svm_l1_dca_update},
    {vl_svm_l2_loss, vl_svm_l2_loss_derivative, vl_svm_l2_conjugate_loss, vl_svm_l2_dca_update},
    {vl_svm_logistic_loss, vl_svm_logistic_loss_derivative, vl_svm_logistic_conjugate_loss,
     vl_svm_logistic_dca_update}} ;
  int k ;
  for (k = 0 ; k < 5 ; ++k) {
    if (self->lossFn == sets[k].loss && self->lossDerivativeFn == sets[k].derivative &&
        self->conjugateLossFn == sets[k].conjugate && self->dcaUpdateFn == sets[k].update) return k ;
  }
  return -1 ;
}

/* ---- statistics (svm.c:1871-1909) ---------------------------------------- */

/* inner[k] (may be NULL): the scores <x_k, w> + bias b of every point,
 * computed elsewhere; otherwise they come from the data functions */
static void
svm_update_statistics (VlSvm * self, double const * inner)
{
  vl_uindex i, k ;
  VlSvmStatistics * st = &self->statistics ;
  memset(st, 0, sizeof(VlSvmStatistics)) ;
  st->regularizer = self->bias * self->bias ;
  for (i = 0 ; i < self->dimension ; ++i) st->regularizer += self->model[i] * self->model[i] ;
  st->regularizer *= self->lambda * 0.5 ;
  for (k = 0 ; k < self->numData ; ++k) {
    double p = self->weights ? self->weights[k] : 1.0 ;
    double s ;
    if (p <= 0) continue ;
    if (inner) {
      s = inner[k] ;
    } else {
      s = self->innerProductFn(self->data, k, self->model) ;
      s += self->bias * self->biasMultiplier ;
    }
    self->scores[k] = s ;
    st->loss += p * self->lossFn(s, self->labels[k]) ;
    if (self->solver == VlSvmSolverSdca) {
      st->dualLoss -= p * self->conjugateLossFn(- self->alpha[k] / p, self->labels[k]) ;
    }
  }
  st->loss /= self->numData ;
  st->objective = st->regularizer + st->loss ;
  if (self->solver == VlSvmSolverSdca) {
    st->dualLoss /= self->numData ;
    st->dualObjective = - st->regularizer + st->dualLoss ;
    st->dualityGap = st->objective - st->dualObjective ;
  }
}

/* status after a diagnostic at iteration t (svm.c:2003-2008, 2126-2131) */
static void
svm_set_status (VlSvm * self, vl_uindex t, double criterion, double startTime)
{
  VlSvmStatistics * st = &self->statistics ;
  st->elapsedTime = vl_get_cpu_time() - startTime ;
  st->iteration = t ;
  st->epoch = t / self->numData ;
  st->status = VlSvmStatusTraining ;
  if (criterion < self->epsilon) st->status = VlSvmStatusConverged ;
  else if (t + 1 == self->maxNumIterations) st->status = VlSvmStatusMaxNumIterationsReached ;
  if (self->diagnosticFn) self->diagnosticFn(self, self->diagnosticFnData) ;
}

static int
svm_is_diagnostic (VlSvm const * self, vl_uindex t)
{
  return (t + 1) % self->diagnosticFrequency == 0 || t + 1 == self->maxNumIterations ;
}

/* ---- host loop (abstract data or caller-set functions) ------------------- */

static void
svm_sdca_host (VlSvm * self)
{
  vl_size const n = self->numData ;
  double * norm2 = (double*) vl_calloc(n, sizeof(double)) ;
  vl_index * perm = (vl_index*) vl_calloc(n, sizeof(vl_index)) ;
  double * buffer = (double*) vl_calloc(self->dimension, sizeof(double)) ;
  double startTime = vl_get_cpu_time() ;
  VlRand * rand = vl_get_rand() ;
  vl_uindex i, t ;
  for (i = 0 ; i < n ; ++i) {
    double n2 ;
    perm[i] = (vl_index) i ;
    memset(buffer, 0, self->dimension * sizeof(double)) ;
    self->accumulateFn(self->data, i, buffer, 1) ;
    n2 = self->innerProductFn(self->data, i, buffer) + self->biasMultiplier * self->biasMultiplier ;
    norm2[i] = n2 / (self->lambda * n) ;
  }
  vl_free(buffer) ;
  for (t = 0 ; ; ++t) {
    double p, delta = 0 ;
    if (t % n == 0) vl_rand_permute_indexes(rand, perm, n) ;
    i = (vl_uindex) perm[t % n] ;
    p = self->weights ? self->weights[i] : 1.0 ;
    if (p > 0) {
      double inner = self->innerProductFn(self->data, i, self->model) + self->bias * self->biasMultiplier ;
      delta = p * self->dcaUpdateFn(self->alpha[i] / p, inner, p * norm2[i], self->labels[i]) ;
    }
    if (delta != 0) {
      double m = delta / (n * self->lambda) ;
      self->alpha[i] += delta ;
      self->accumulateFn(self->data, i, self->model, m) ;
      self->bias += self->biasMultiplier * m ;
    }
    if (svm_is_diagnostic(self, t)) {
      svm_update_statistics(self, NULL) ;
      svm_set_status(self, t, self->statistics.dualityGap, startTime) ;
      if (self->statistics.status != VlSvmStatusTraining) break ;
    }
  }
  vl_free(norm2) ;
  vl_free(perm) ;
}

static void
svm_sgd_host (VlSvm * self)
{
  vl_size const n = self->numData ;
  vl_index * perm = (vl_index*) vl_calloc(n, sizeof(vl_index)) ;
  double * sc = (double*) vl_calloc(2 * n, sizeof(double)) ;
  double * prev = sc + n ;
  double factor = 1.0, biasFactor = 1.0 ;
  vl_index const t0 = VL_MAX(2, (vl_index) ceil(1.0 / self->lambda)) ;
  double startTime = vl_get_cpu_time() ;
  VlRand * rand = vl_get_rand() ;
  vl_uindex i, t, k ;
  for (i = 0 ; i < n ; ++i) { perm[i] = (vl_index) i ; prev[i] = - VL_INFINITY_D ; }
  for (t = 0 ; ; ++t) {
    double p, inner, gradient, rate, biasRate ;
    if (t % n == 0) vl_rand_permute_indexes(rand, perm, n) ;
    i = (vl_uindex) perm[t % n] ;
    p = self->weights ? self->weights[i] : 1.0 ;
    p = VL_MAX(0.0, p) ;
    inner = factor * self->innerProductFn(self->data, i, self->model) ;
    inner += biasFactor * (self->biasMultiplier * self->bias) ;
    gradient = p * self->lossDerivativeFn(inner, self->labels[i]) ;
    prev[i] = sc[i] ;
    sc[i] = inner ;
    rate = 1.0 / (self->lambda * (t + t0)) ;
    biasRate = rate * self->biasLearningRate ;
    factor *= (1.0 - self->lambda * rate) ;
    biasFactor *= (1.0 - self->lambda * biasRate) ;
    if (gradient != 0) {
      self->accumulateFn(self->data, i, self->model, - gradient * rate / factor) ;
      self->bias += self->biasMultiplier * (- gradient * biasRate / biasFactor) ;
    }
    if (svm_is_diagnostic(self, t)) {
      double var = 0 ;
      for (k = 0 ; k < self->dimension ; ++k) self->model[k] *= factor ;
      self->bias *= biasFactor ;
      factor = 1.0 ;
      biasFactor = 1.0 ;
      svm_update_statistics(self, NULL) ;
      for (k = 0 ; k < n ; ++k) { double d = sc[k] - prev[k] ; var += d * d ; }
      self->statistics.scoresVariation = sqrt(var) / n ;
      svm_set_status(self, t, self->statistics.scoresVariation, startTime) ;
      if (self->statistics.status != VlSvmStatusTraining) break ;
    }
  }
  vl_free(sc) ;
  vl_free(perm) ;
}

/* ---- GPU path ------------------------------------------------------------- */

typedef struct {
  vl_type dataType ;
  void * data ;              /* numData x dimension mapped points */
  double * labels, * weights, * model, * alpha, * norm2, * scores, * prev, * inner ;
  int64_t * perm ;
  void * workspace ;
  double * innerHost ;       /* host copy of the scores of all points */
} SvmDevice ;

static void
svm_device_open (VlSvm * self, SvmDevice * dv)
{
  vl_type dt ;
  void const * hdata ;
  vl_size dim0, n, es, D = self->dimension ;
  VlHomogeneousKernelMap const * hom ;
  memset(dv, 0, sizeof(*dv)) ;
  svm_dataset_describe(self->dataset, &dt, &hdata, &dim0, &n, &hom) ;
  es = vl_get_type_size(dt) ;
  dv->dataType = dt ;
  dv->data = vlhip_malloc(es * D * n) ;
  if (hom) {
    void * raw = vlhip_malloc(es * dim0 * n) ;
    vlhip_memcpy_h2d(raw, hdata, es * dim0 * n) ;
    homkermap_evaluate_device(hom, dt, dv->data, raw, dim0 * n) ;
    vlhip_synchronize() ;
    vlhip_free(raw) ;
  } else {
    vlhip_memcpy_h2d(dv->data, hdata, es * D * n) ;
  }
  dv->labels = (double*) vlhip_malloc(sizeof(double) * n) ;
  vlhip_memcpy_h2d(dv->labels, self->labels, sizeof(double) * n) ;
  if (self->weights) {
    dv->weights = (double*) vlhip_malloc(sizeof(double) * n) ;
    vlhip_memcpy_h2d(dv->weights, self->weights, sizeof(double) * n) ;
  }
  dv->model = (double*) vlhip_malloc(sizeof(double) * D) ;
  vlhip_memcpy_h2d(dv->model, self->model, sizeof(double) * D) ;
  dv->inner = (double*) vlhip_malloc(sizeof(double) * n) ;
  dv->innerHost = (double*) vl_malloc(sizeof(double) * n) ;
  dv->perm = (int64_t*) vlhip_malloc(sizeof(int64_t) * n) ;
  dv->workspace = vlhip_malloc(vlhip_svm_workspace_bytes(D)) ;
  vlhip_svm_reset_workspace(dv->workspace, D) ;
}

static void
svm_device_close (SvmDevice * dv)
{
  void * p [] = {dv->data, dv->labels, dv->weights, dv->model, dv->alpha, dv->norm2, dv->scores, dv->prev,
                 dv->inner, dv->perm, dv->workspace} ;
  size_t i ;
  for (i = 0 ; i < sizeof(p) / sizeof(p[0]) ; ++i) if (p[i]) vlhip_free(p[i]) ;
  vl_free(dv->innerHost) ;
}

/* the scores of all points from the model on the device (model downloaded
 * into self->model by the caller), then the statistics */
static void
svm_device_statistics (VlSvm * self, SvmDevice * dv)
{
  vlhip_svm_scores((int) dv->dataType, dv->inner, dv->data, dv->model, self->dimension, self->numData,
                   self->bias * self->biasMultiplier) ;
  vlhip_memcpy_d2h(dv->innerHost, dv->inner, sizeof(double) * self->numData) ;
  svm_update_statistics(self, dv->innerHost) ;
}

/* last iteration of the run that starts at t: the end of t's epoch or the
 * next diagnostic, whichever comes first */
static vl_uindex
svm_run_last (VlSvm const * self, vl_uindex t)
{
  vl_uindex const f = self->diagnosticFrequency ;
  vl_uindex last = (t / self->numData + 1) * self->numData - 1 ;
  vl_uindex diag = ((t + f) / f) * f - 1 ;
  if (diag < last) last = diag ;
  if (self->maxNumIterations > t && self->maxNumIterations - 1 < last) last = self->maxNumIterations - 1 ;
  return last ;
}

static void
svm_train_device (VlSvm * self, int loss)
{
  vl_size const n = self->numData ;
  SvmDevice dv ;
  VlhipSvmRun run ;
  vl_index * perm = (vl_index*) vl_calloc(n, sizeof(vl_index)) ;
  double startTime = vl_get_cpu_time() ;
  VlRand * rand = vl_get_rand() ;
  vl_uindex i, t = 0 ;
  int const sdca = self->solver == VlSvmSolverSdca ;

  svm_device_open(self, &dv) ;
  memset(&run, 0, sizeof(run)) ;
  run.dtype = (int) dv.dataType ;
  run.solver = (int) self->solver ;
  run.loss = loss ;
  run.data = dv.data ;
  run.dimension = self->dimension ;
  run.numData = n ;
  run.labels = dv.labels ;
  run.weights = dv.weights ;
  run.model = dv.model ;
  run.lambda = self->lambda ;
  run.biasMultiplier = self->biasMultiplier ;
  run.biasLearningRate = self->biasLearningRate ;
  run.sgdT0 = VL_MAX(2, (vl_index) ceil(1.0 / self->lambda)) ;
  run.state[0] = self->bias ;
  run.state[1] = 1.0 ;
  run.state[2] = 1.0 ;
  run.permutation = dv.perm ;
  run.workspace = dv.workspace ;
  for (i = 0 ; i < n ; ++i) perm[i] = (vl_index) i ;

  if (sdca) {
    /* svm.c:1943-1952: (<x_i, x_i> + b^2) / (lambda n) */
    double * h = (double*) vl_malloc(sizeof(double) * n) ;
    dv.alpha = (double*) vlhip_malloc(sizeof(double) * n) ;
    dv.norm2 = (double*) vlhip_malloc(sizeof(double) * n) ;
    vlhip_memcpy_h2d(dv.alpha, self->alpha, sizeof(double) * n) ;
    vlhip_svm_square_norms((int) dv.dataType, dv.norm2, dv.data, self->dimension, n) ;
    vlhip_memcpy_d2h(h, dv.norm2, sizeof(double) * n) ;
    for (i = 0 ; i < n ; ++i) h[i] = (h[i] + self->biasMultiplier * self->biasMultiplier) / (self->lambda * n) ;
    vlhip_memcpy_h2d(dv.norm2, h, sizeof(double) * n) ;
    vl_free(h) ;
    run.alpha = dv.alpha ;
    run.norm2 = dv.norm2 ;
  } else {
    double * h = (double*) vl_malloc(sizeof(double) * n) ;
    dv.scores = (double*) vlhip_malloc(sizeof(double) * n) ;
    dv.prev = (double*) vlhip_malloc(sizeof(double) * n) ;
    for (i = 0 ; i < n ; ++i) h[i] = 0.0 ;
    vlhip_memcpy_h2d(dv.scores, h, sizeof(double) * n) ;
    for (i = 0 ; i < n ; ++i) h[i] = - VL_INFINITY_D ;
    vlhip_memcpy_h2d(dv.prev, h, sizeof(double) * n) ;
    vl_free(h) ;
    run.scores = dv.scores ;
    run.previousScores = dv.prev ;
  }

  for (;;) {
    vl_uindex last ;
    if (t % n == 0) {
      vl_rand_permute_indexes(rand, perm, n) ;
      vlhip_memcpy_h2d(dv.perm, perm, sizeof(int64_t) * n) ;
    }
    last = svm_run_last(self, t) ;
    run.begin = t ;
    run.end = last + 1 ;
    vlhip_svm_run(&run) ;
    t = last + 1 ;
    if (!svm_is_diagnostic(self, last)) continue ;
    if (sdca) {
      self->bias = run.state[0] ;
      vlhip_memcpy_d2h(self->model, dv.model, sizeof(double) * self->dimension) ;
      vlhip_memcpy_d2h(self->alpha, dv.alpha, sizeof(double) * n) ;
      svm_device_statistics(self, &dv) ;
      svm_set_status(self, last, self->statistics.dualityGap, startTime) ;
    } else {
      /* svm.c:2113-2124: the scale factors are applied to the model and bias */
      double var = 0 ;
      double * sc = (double*) vl_malloc(sizeof(double) * 2 * n) ;
      vlhip_svm_scale(dv.model, self->dimension, run.state[1]) ;
      self->bias = run.state[0] * run.state[2] ;
      run.state[0] = self->bias ;
      run.state[1] = 1.0 ;
      run.state[2] = 1.0 ;
      vlhip_memcpy_d2h(self->model, dv.model, sizeof(double) * self->dimension) ;
      svm_device_statistics(self, &dv) ;
      vlhip_memcpy_d2h(sc, dv.scores, sizeof(double) * n) ;
      vlhip_memcpy_d2h(sc + n, dv.prev, sizeof(double) * n) ;
      for (i = 0 ; i < n ; ++i) { double d = sc[i] - sc[n + i] ; var += d * d ; }
      vl_free(sc) ;
      self->statistics.scoresVariation = sqrt(var) / n ;
      svm_set_status(self, last, self->statistics.scoresVariation, startTime) ;
    }
    if (self->statistics.status != VlSvmStatusTraining) break ;
  }
  svm_device_close(&dv) ;
  vl_free(perm) ;
}

static void
svm_evaluate (VlSvm * self, int device)
{
  double startTime = vl_get_cpu_time() ;
  if (device) {
    SvmDevice dv ;
    svm_device_open(self, &dv) ;
    svm_device_statistics(self, &dv) ;
    svm_device_close(&dv) ;
  } else {
    svm_update_statistics(self, NULL) ;
  }
  self->statistics.elapsedTime = vl_get_cpu_time() - startTime ;
  self->statistics.iteration = 0 ;
  self->statistics.epoch = 0 ;
  self->statistics.status = VlSvmStatusConverged ;
  if (self->diagnosticFn) self->diagnosticFn(self, self->diagnosticFnData) ;
}

VL_EXPORT void
vl_svm_train (VlSvm * self)
{
  int const loss = svm_builtin_loss(self) ;
  int const device = self->dataset != NULL && loss >= 0 &&
                     svm_dataset_is_builtin(self->dataset, self->innerProductFn, self->accumulateFn) ;
  switch (self->solver) {
    case VlSvmSolverSdca:
    case VlSvmSolverSgd:
      if (device) svm_train_device(self, loss) ;
      else if (self->solver == VlSvmSolverSdca) svm_sdca_host(self) ;
      else svm_sgd_host(self) ;
      break ;
    case VlSvmSolverNone:
      svm_evaluate(self, device) ;
      break ;
    default:
      vl_host_fatal("vl_svm_train: unknown solver %d\n", (int) self->solver) ;
  }
}
