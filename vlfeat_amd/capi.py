"""ctypes binding of the VLFeat C interface (vl/gmm.h, vl/fisher.h, vl/vlad.h,
vl/kmeans.h, vl/generic.h) and of the device ABI (vlhip.h).

``CLib(path)`` binds any shared object that exports the VLFeat 0.9.21 C API
with the reference's signatures -- this library's ``libvl.so``, and (for the
test harness) the reference build -- so the same calls drive both.  Arrays
are numpy, row-major, one datum per row (the memory layout of the reference's
column-major D x N MATLAB arrays).
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

VL_TYPE_FLOAT = 1
VL_TYPE_DOUBLE = 2

VlDistanceL2 = 1
VlKMeansLloyd, VlKMeansElkan, VlKMeansANN = 0, 1, 2
VlKMeansRandomSelection, VlKMeansPlusPlus = 0, 1
VlGMMKMeans, VlGMMRand, VlGMMCustom = 0, 1, 2

FISHER_SQUARE_ROOT = 1
FISHER_NORMALIZED = 2
FISHER_IMPROVED = 3
FISHER_FAST = 4

VLAD_NORMALIZE_COMPONENTS = 1
VLAD_SQUARE_ROOT = 2
VLAD_UNNORMALIZED = 4
VLAD_NORMALIZE_MASS = 8

MODE_FP32, MODE_BF16X3, MODE_GATED, MODE_F16X3, MODE_FP32_MFMA = 0, 1, 2, 3, 4

_p = C.c_void_p
_sz = C.c_uint64


class VlKMeans(C.Structure):
    """Field layout of the public VlKMeans structure (vl/kmeans.h:45-68)."""
    _fields_ = [
        ("dataType", C.c_uint32), ("dimension", _sz), ("numCenters", _sz), ("numTrees", _sz),
        ("maxNumComparisons", _sz), ("initialization", C.c_int), ("algorithm", C.c_int),
        ("distance", C.c_int), ("maxNumIterations", _sz), ("minEnergyVariation", C.c_double),
        ("numRepetitions", _sz), ("verbosity", C.c_int), ("centers", _p), ("centerDistances", _p),
        ("energy", C.c_double), ("floatVectorComparisonFn", _p), ("doubleVectorComparisonFn", _p),
    ]


class VlRand(C.Structure):
    _fields_ = [("mt", C.c_uint32 * 624), ("mti", C.c_uint32)]


_SIGS = {
    # generic.h / random.h
    "vl_set_num_threads": (None, [_sz]),
    "vl_get_max_threads": (_sz, []),
    "vl_set_simd_enabled": (None, [C.c_int]),
    "vl_get_simd_enabled": (C.c_int, []),
    "vl_cpu_has_avx": (C.c_int, []),
    "vl_get_rand": (_p, []),
    "vl_rand_seed": (None, [_p, C.c_uint32]),
    "vl_rand_seed_by_array": (None, [_p, _p, _sz]),
    "vl_rand_init": (None, [_p]),
    "vl_rand_uint32": (C.c_uint32, [_p]),
    # gmm.h
    "vl_gmm_new": (_p, [C.c_uint32, _sz, _sz]),
    "vl_gmm_delete": (None, [_p]),
    "vl_gmm_cluster": (C.c_double, [_p, _p, _sz]),
    "vl_gmm_em": (C.c_double, [_p, _p, _sz]),
    "vl_gmm_init_with_rand_data": (None, [_p, _p, _sz]),
    "vl_gmm_init_with_kmeans": (None, [_p, _p, _sz, _p]),
    "vl_gmm_set_initialization": (None, [_p, C.c_int]),
    "vl_gmm_set_max_num_iterations": (None, [_p, _sz]),
    "vl_gmm_set_num_repetitions": (None, [_p, _sz]),
    "vl_gmm_set_verbosity": (None, [_p, C.c_int]),
    "vl_gmm_set_covariance_lower_bound": (None, [_p, C.c_double]),
    "vl_gmm_set_covariance_lower_bounds": (None, [_p, _p]),
    "vl_gmm_set_kmeans_init_object": (None, [_p, _p]),
    "vl_gmm_get_kmeans_init_object": (_p, [_p]),
    "vl_gmm_set_means": (None, [_p, _p]),
    "vl_gmm_set_covariances": (None, [_p, _p]),
    "vl_gmm_set_priors": (None, [_p, _p]),
    "vl_gmm_get_means": (_p, [_p]),
    "vl_gmm_get_covariances": (_p, [_p]),
    "vl_gmm_get_priors": (_p, [_p]),
    "vl_gmm_get_posteriors": (_p, [_p]),
    "vl_gmm_get_loglikelihood": (C.c_double, [_p]),
    "vl_gmm_get_num_data": (_sz, [_p]),
    "vl_get_gmm_data_posteriors_f": (C.c_double, [_p, _sz, _sz, _p, _p, _sz, _p, _p]),
    "vl_get_gmm_data_posteriors_d": (C.c_double, [_p, _sz, _sz, _p, _p, _sz, _p, _p]),
    # fisher.h / vlad.h
    "vl_fisher_encode": (_sz, [_p, C.c_uint32, _p, _sz, _sz, _p, _p, _p, _sz, C.c_int]),
    "vl_vlad_encode": (None, [_p, C.c_uint32, _p, _sz, _sz, _p, _sz, _p, C.c_int]),
    # kmeans.h
    "vl_kmeans_new": (_p, [C.c_uint32, C.c_int]),
    "vl_kmeans_delete": (None, [_p]),
    "vl_kmeans_cluster": (C.c_double, [_p, _p, _sz, _sz, _sz]),
    "vl_kmeans_quantize": (None, [_p, _p, _p, _p, _sz]),
    "vl_kmeans_set_centers": (None, [_p, _p, _sz, _sz]),
    "vl_kmeans_init_centers_with_rand_data": (None, [_p, _p, _sz, _sz, _sz]),
    "vl_kmeans_init_centers_plus_plus": (None, [_p, _p, _sz, _sz, _sz]),
    "vl_kmeans_refine_centers": (C.c_double, [_p, _p, _sz]),
    # homkermap.h
    "vl_homogeneouskernelmap_new": (_p, [C.c_int, C.c_double, _sz, C.c_double, C.c_int]),
    "vl_homogeneouskernelmap_delete": (None, [_p]),
    "vl_homogeneouskernelmap_evaluate_d": (None, [_p, _p, _sz, C.c_double]),
    "vl_homogeneouskernelmap_evaluate_f": (None, [_p, _p, _sz, C.c_double]),
    "vl_homogeneouskernelmap_get_order": (_sz, [_p]),
    "vl_homogeneouskernelmap_get_dimension": (_sz, [_p]),
    "vl_homogeneouskernelmap_get_kernel_type": (C.c_int, [_p]),
    "vl_homogeneouskernelmap_get_window_type": (C.c_int, [_p]),
    # svm.h / svmdataset.h
    "vl_svmdataset_new": (_p, [C.c_uint32, _p, _sz, _sz]),
    "vl_svmdataset_delete": (None, [_p]),
    "vl_svmdataset_set_homogeneous_kernel_map": (None, [_p, _p]),
    "vl_svmdataset_get_dimension": (_sz, [_p]),
    "vl_svmdataset_get_inner_product_function": (_p, [_p]),
    "vl_svmdataset_get_accumulate_function": (_p, [_p]),
    "vl_svm_new": (_p, [C.c_int, _p, _sz, _sz, _p, C.c_double]),
    "vl_svm_new_with_dataset": (_p, [C.c_int, _p, _p, C.c_double]),
    "vl_svm_new_with_abstract_data": (_p, [C.c_int, _p, _sz, _sz, _p, C.c_double]),
    "vl_svm_delete": (None, [_p]),
    "vl_svm_train": (None, [_p]),
    "vl_svm_set_loss": (None, [_p, C.c_int]),
    "vl_svm_set_epsilon": (None, [_p, C.c_double]),
    "vl_svm_set_bias_multiplier": (None, [_p, C.c_double]),
    "vl_svm_set_bias_learning_rate": (None, [_p, C.c_double]),
    "vl_svm_set_max_num_iterations": (None, [_p, _sz]),
    "vl_svm_set_diagnostic_frequency": (None, [_p, _sz]),
    "vl_svm_set_weights": (None, [_p, _p]),
    "vl_svm_set_model": (None, [_p, _p]),
    "vl_svm_set_bias": (None, [_p, C.c_double]),
    "vl_svm_set_data_functions": (None, [_p, _p, _p]),
    "vl_svm_set_diagnostic_function": (None, [_p, _p, _p]),
    "vl_svm_get_model": (_p, [_p]),
    "vl_svm_get_bias": (C.c_double, [_p]),
    "vl_svm_get_scores": (_p, [_p]),
    "vl_svm_get_statistics": (_p, [_p]),
    "vl_svm_get_max_num_iterations": (_sz, [_p]),
    "vl_svm_get_dimension": (_sz, [_p]),
    "vl_svm_hinge_loss": (C.c_double, [C.c_double, C.c_double]),
    "vl_svm_logistic_dca_update": (C.c_double, [C.c_double, C.c_double, C.c_double, C.c_double]),
}


class VlSvmStatistics(C.Structure):
    """vl/svm.h:76-88"""
    _fields_ = [("status", C.c_int), ("iteration", _sz), ("epoch", _sz), ("objective", C.c_double),
                ("regularizer", C.c_double), ("loss", C.c_double), ("dualObjective", C.c_double),
                ("dualLoss", C.c_double), ("dualityGap", C.c_double), ("scoresVariation", C.c_double),
                ("elapsedTime", C.c_double)]


SVM_STAT_FIELDS = ("status", "iteration", "epoch", "objective", "regularizer", "loss", "dualObjective",
                   "dualLoss", "dualityGap", "scoresVariation")

# Entry points present only in this library (batched forms and the device ABI).
_EXT_SIGS = {
    "vl_fisher_encode_batch": (None, [_p, C.c_uint32, _p, _sz, _sz, _p, _p, _p, _p, _sz, C.c_int, _p]),
    "vl_fisher_set_precision": (None, [C.c_int]),
    "vl_gmm_set_precision": (None, [_p, C.c_int]),
    "vl_gmm_get_precision": (C.c_int, [_p]),
    "vlhip_gmm_em_step": (C.c_double, [C.c_int, _p, _p, _p, _p, C.c_size_t, C.c_size_t, _p, C.c_size_t]),
    "vlhip_gmm_moments": (None, [C.c_int, _p, _p, _p, _p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_double]),
    "vl_fisher_get_precision": (C.c_int, []),
    "vl_fisher_set_gate": (None, [_p, C.c_int, C.c_float]),
    "vlhip_fisher_set_gate": (None, [_p, C.c_int, C.c_float]),
    "vlhip_fisher_set_fused": (None, [C.c_int]),
    "vlhip_fisher_gate_stats": (None, [C.POINTER(C.c_ulonglong), C.POINTER(C.c_ulonglong), C.c_int]),
    "vl_vlad_encode_hard": (None, [_p, C.c_uint32, _p, _sz, _sz, _p, _sz, _p, C.c_int]),
    "vlhip_available": (C.c_int, []),
    "vlhip_init": (None, []),
    "vlhip_set_stream": (None, [_p]),
    "vlhip_malloc": (_p, [C.c_size_t]),
    "vlhip_free": (None, [_p]),
    "vlhip_memcpy_h2d": (None, [_p, _p, C.c_size_t]),
    "vlhip_memcpy_d2h": (None, [_p, _p, C.c_size_t]),
    "vlhip_memset": (None, [_p, C.c_int, C.c_size_t]),
    "vlhip_synchronize": (None, []),
    "vlhip_num_compute_units": (C.c_int, []),
    "vlhip_timing_enable": (None, [C.c_int]),
    "vlhip_timing_reset": (None, []),
    "vlhip_timing_count": (C.c_int, []),
    "vlhip_timing_get": (C.c_int, [C.c_int, C.c_char_p, C.c_size_t, C.POINTER(C.c_double),
                                   C.POINTER(C.c_ulonglong)]),
    "vlhip_gmm_posteriors": (C.c_double, [C.c_int, _p, C.c_size_t, C.c_size_t, _p, _p, C.c_size_t, _p, _p]),
    "vlhip_fisher_encode_batch": (None, [C.c_int, _p, _p, _p, _p, C.c_size_t, C.c_size_t, _p, _p,
                                         C.c_size_t, C.c_int, C.c_int, _p]),
    "vlhip_vlad_encode_batch": (None, [C.c_int, _p, _p, C.c_size_t, C.c_size_t, _p, _p, C.c_size_t,
                                       _p, _p, C.c_int]),
    "vlhip_kmeans_quantize": (None, [C.c_int, _p, _p, _p, _p, C.c_size_t, C.c_size_t, C.c_size_t, C.c_int]),
    "vlhip_kmeans_screen_stats": (None, [_p, C.c_int]),
    "vlhip_kmeans_quantize_noalloc": (None, [C.c_int, _p, _p, _p, _p, C.c_size_t, C.c_size_t, C.c_size_t,
                                             C.c_int]),
    "vlhip_encode_plan_new": (_p, [_sz, _sz, _p, _sz, C.c_int, C.c_int]),
    "vlhip_encode_plan_run": (None, [_p, _p, _p, _p, _p, _p, _p, _p]),
    "vlhip_encode_plan_info": (None, [_p, C.POINTER(C.c_int), C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    "vlhip_encode_plan_delete": (None, [_p]),
    "vlhip_encode_plan_candidates": (C.c_int, [_p, C.c_int, C.POINTER(C.c_int), C.POINTER(C.c_int),
                                               C.POINTER(C.c_int), C.POINTER(C.c_double)]),
    "vlhip_encode_plan_set_projection": (None, [_p, _sz, _p, _p, C.c_int]),
    "vl_homogeneouskernelmap_evaluate_array": (None, [_p, C.c_uint32, _p, _p, _sz]),
    "vlhip_homkermap_evaluate": (None, [C.c_int, _p, _p, _sz, _p, C.c_int, C.c_int, C.c_int, C.c_int]),
    "vlhip_memcpy_h2d_async": (None, [_p, _p, C.c_size_t]),
    "vlhip_memcpy_d2h_async": (None, [_p, _p, C.c_size_t]),
    "vlhip_project": (None, [_p, _p, _sz, _sz, _sz, _p, _p, C.c_int]),
    "vlhip_memcpy_d2d": (None, [_p, _p, C.c_size_t]),
    "vlhip_kmeans_update_centers": (None, [C.c_int, _p, _p, _p, _p, _p, _sz, _sz]),
    "vlhip_kmeans_ann_merge": (None, [C.c_int, _p, _p, _p, _p, _p, _sz, _sz, C.c_int]),
    "vlhip_gate_new": (_p, [_sz, _sz, _sz, _sz, C.c_ulonglong]),
    "vlhip_gate_delete": (None, [_p]),
    "vlhip_gate_num_params": (_sz, [_p]),
    "vlhip_gate_get_params": (None, [_p, _p, _p, _p]),
    "vlhip_gate_set_params": (None, [_p, _p, _p, _p]),
    "vlhip_gate_set_standardization_from_data": (None, [_p, _p, _sz]),
    "vlhip_gate_gradients": (C.c_double, [_p, _p, _p, _sz, _p]),
    "vlhip_gate_train": (C.c_double, [_p, _p, _sz, _p, _p, _p, _sz, _sz, C.c_float, _p]),
    "vlhip_gate_select": (None, [_p, _p, _sz, C.c_int, C.c_float, _p, _p, _p]),
    "vlhip_gate_logits": (None, [_p, _p, _sz, _p]),
}


def dtype_code(a: np.ndarray) -> int:
    if a.dtype == np.float32:
        return VL_TYPE_FLOAT
    if a.dtype == np.float64:
        return VL_TYPE_DOUBLE
    raise TypeError(f"unsupported dtype {a.dtype}; float32 or float64 expected")


def ptr(a: np.ndarray | None) -> int | None:
    return None if a is None else a.ctypes.data


class CLib:
    """One loaded library exporting the VLFeat C API."""

    def __init__(self, path: str | Path, extended: bool = False):
        self.path = str(path)
        self.lib = C.CDLL(self.path, mode=C.RTLD_LOCAL)
        sigs = dict(_SIGS)
        if extended:
            sigs.update(_EXT_SIGS)
        for name, (res, args) in sigs.items():
            fn = getattr(self.lib, name)
            fn.restype = res
            fn.argtypes = args
            setattr(self, name, fn)
        self.extended = extended

    # ---- generic -------------------------------------------------------
    def seed(self, s: int) -> None:
        self.vl_rand_seed(self.vl_get_rand(), s)

    def set_num_threads(self, n: int) -> None:
        self.vl_set_num_threads(n)

    # ---- SVM -----------------------------------------------------------
    def svm_train(self, data, labels, lam, *, solver=2, loss=0, weights=None, hom=None, bias_multiplier=None,
                  epsilon=None, max_iter=None, diag_freq=None, bias_lr=None, abstract=False, seed=None,
                  model=None, bias=None, diagnostics=False):
        """Train an SVM on the rows of data (float32 or float64) through the C
        API; returns a dict with model, bias, statistics, scores and, with
        diagnostics, the statistics seen at every diagnostic call.  hom:
        (kernel, gamma, order, period, window) of a homogeneous kernel map
        applied by the dataset.  abstract: train through
        vl_svm_new_with_abstract_data with the dataset's own data functions
        (the solver then calls them as caller-set functions)."""
        x = np.ascontiguousarray(data)
        n, d = x.shape
        y = np.ascontiguousarray(labels, np.float64)
        w = None if weights is None else np.ascontiguousarray(weights, np.float64)
        if seed is not None:
            self.seed(seed)
        ds = self.vl_svmdataset_new(dtype_code(x), ptr(x), d, n)
        hmap = None
        if hom is not None:
            hmap = self.vl_homogeneouskernelmap_new(*hom)
            self.vl_svmdataset_set_homogeneous_kernel_map(ds, hmap)
        if abstract:
            svm = self.vl_svm_new_with_abstract_data(solver, ds, self.vl_svmdataset_get_dimension(ds), n,
                                                     ptr(y), lam)
            self.vl_svm_set_data_functions(svm, self.vl_svmdataset_get_inner_product_function(ds),
                                           self.vl_svmdataset_get_accumulate_function(ds))
        else:
            svm = self.vl_svm_new_with_dataset(solver, ds, ptr(y), lam)
        self.vl_svm_set_loss(svm, loss)
        if bias_multiplier is not None:
            self.vl_svm_set_bias_multiplier(svm, bias_multiplier)
        if epsilon is not None:
            self.vl_svm_set_epsilon(svm, epsilon)
        if max_iter is not None:
            self.vl_svm_set_max_num_iterations(svm, max_iter)
        if diag_freq is not None:
            self.vl_svm_set_diagnostic_frequency(svm, diag_freq)
        if bias_lr is not None:
            self.vl_svm_set_bias_learning_rate(svm, bias_lr)
        if w is not None:
            self.vl_svm_set_weights(svm, ptr(w))
        D = int(self.vl_svm_get_dimension(svm))
        if model is not None:
            m0 = np.ascontiguousarray(model, np.float64)
            self.vl_svm_set_model(svm, ptr(m0))
        if bias is not None:
            self.vl_svm_set_bias(svm, bias)
        seen = []
        cb = None
        if diagnostics:
            CB = C.CFUNCTYPE(None, C.c_void_p, C.c_void_p)

            def on_diag(handle, _):
                st = C.cast(self.vl_svm_get_statistics(handle), C.POINTER(VlSvmStatistics)).contents
                seen.append({f: getattr(st, f) for f in SVM_STAT_FIELDS})
            cb = CB(on_diag)
            self.vl_svm_set_diagnostic_function(svm, C.cast(cb, C.c_void_p), None)
        self.vl_svm_train(svm)
        st = C.cast(self.vl_svm_get_statistics(svm), C.POINTER(VlSvmStatistics)).contents
        out = {
            "model": np.ctypeslib.as_array(C.cast(self.vl_svm_get_model(svm), C.POINTER(C.c_double)), (D,)).copy(),
            "bias": float(self.vl_svm_get_bias(svm)),
            "scores": np.ctypeslib.as_array(C.cast(self.vl_svm_get_scores(svm), C.POINTER(C.c_double)), (n,)).copy(),
            "statistics": {f: getattr(st, f) for f in SVM_STAT_FIELDS},
            "max_iter": int(self.vl_svm_get_max_num_iterations(svm)),
            "diagnostics": seen,
        }
        self.vl_svm_delete(svm)
        self.vl_svmdataset_delete(ds)
        if hmap is not None:
            self.vl_homogeneouskernelmap_delete(hmap)
        del cb
        return out

    # ---- GMM -----------------------------------------------------------
    def gmm_posteriors(self, data, means, covs, priors):
        K, D = means.shape
        N = data.shape[0]
        q = np.empty((N, K), dtype=data.dtype)
        fn = self.vl_get_gmm_data_posteriors_f if data.dtype == np.float32 else self.vl_get_gmm_data_posteriors_d
        ll = fn(ptr(q), K, N, ptr(priors), ptr(means), D, ptr(covs), ptr(data))
        return q, ll

    def gmm(self, data, K, *, init="custom", means=None, covs=None, priors=None, max_iter=50,
            reps=1, cov_bound=None, kmeans=None, verbosity=0, want_post=True, precision=0):
        """vl_gmm-style driver: returns (means, covs, priors, LL, posteriors).
        precision: VL_GMM_PRECISION_* of the EM iterations (0 = FP32 parity path)."""
        data = np.ascontiguousarray(data)
        N, D = data.shape
        t = dtype_code(data)
        g = self.vl_gmm_new(t, D, K)
        try:
            self.vl_gmm_set_max_num_iterations(g, max_iter)
            self.vl_gmm_set_num_repetitions(g, reps)
            self.vl_gmm_set_verbosity(g, verbosity)
            if precision:
                self.vl_gmm_set_precision(g, int(precision))
            if cov_bound is not None:
                if np.ndim(cov_bound) == 0:
                    self.vl_gmm_set_covariance_lower_bound(g, float(cov_bound))
                else:
                    b = np.ascontiguousarray(cov_bound, dtype=np.float64)
                    self.vl_gmm_set_covariance_lower_bounds(g, ptr(b))
            mode = {"kmeans": VlGMMKMeans, "rand": VlGMMRand, "custom": VlGMMCustom}[init]
            self.vl_gmm_set_initialization(g, mode)
            if init == "custom":
                self.vl_gmm_set_means(g, ptr(np.ascontiguousarray(means, dtype=data.dtype)))
                self.vl_gmm_set_covariances(g, ptr(np.ascontiguousarray(covs, dtype=data.dtype)))
                self.vl_gmm_set_priors(g, ptr(np.ascontiguousarray(priors, dtype=data.dtype)))
            if kmeans is not None:
                self.vl_gmm_set_kmeans_init_object(g, kmeans)
            ll = self.vl_gmm_cluster(g, ptr(data), N)
            es = data.dtype
            m = np.ctypeslib.as_array(C.cast(self.vl_gmm_get_means(g), C.POINTER(np.ctypeslib.as_ctypes_type(es))),
                                      shape=(K, D)).copy()
            c = np.ctypeslib.as_array(C.cast(self.vl_gmm_get_covariances(g), C.POINTER(np.ctypeslib.as_ctypes_type(es))),
                                      shape=(K, D)).copy()
            p = np.ctypeslib.as_array(C.cast(self.vl_gmm_get_priors(g), C.POINTER(np.ctypeslib.as_ctypes_type(es))),
                                      shape=(K,)).copy()
            q = None
            if want_post and N > 0:
                q = np.ctypeslib.as_array(C.cast(self.vl_gmm_get_posteriors(g),
                                                 C.POINTER(np.ctypeslib.as_ctypes_type(es))), shape=(N, K)).copy()
            return m, c, p, ll, q
        finally:
            self.vl_gmm_delete(g)

    # ---- encoders ------------------------------------------------------
    def fisher(self, data, means, covs, priors, flags=0):
        K, D = means.shape
        enc = np.empty(2 * K * D, dtype=data.dtype)
        n = self.vl_fisher_encode(ptr(enc), dtype_code(data), ptr(means), D, K, ptr(covs), ptr(priors),
                                  ptr(data), data.shape[0], flags)
        return enc, n

    def vlad(self, data, means, assignments, flags=0):
        K, D = means.shape
        enc = np.empty(K * D, dtype=data.dtype)
        self.vl_vlad_encode(ptr(enc), dtype_code(data), ptr(means), D, K, ptr(data), data.shape[0],
                            ptr(assignments), flags)
        return enc

    # ---- k-means -------------------------------------------------------
    def kmeans_new(self, dtype, algorithm=VlKMeansLloyd, init=VlKMeansRandomSelection, max_iter=100,
                   num_trees=3, max_comparisons=100, reps=1, verbosity=0):
        km = self.vl_kmeans_new(VL_TYPE_FLOAT if dtype == np.float32 else VL_TYPE_DOUBLE, VlDistanceL2)
        s = C.cast(km, C.POINTER(VlKMeans)).contents
        s.algorithm = algorithm
        s.initialization = init
        s.maxNumIterations = max_iter
        s.numTrees = num_trees
        s.maxNumComparisons = max_comparisons
        s.numRepetitions = reps
        s.verbosity = verbosity
        return km

    def kmeans_centers(self, km, dtype):
        s = C.cast(km, C.POINTER(VlKMeans)).contents
        ct = np.ctypeslib.as_ctypes_type(np.dtype(dtype))
        return np.ctypeslib.as_array(C.cast(s.centers, C.POINTER(ct)), shape=(s.numCenters, s.dimension)).copy()

    def homkermap(self, data, kernel=1, gamma=1.0, order=1, period=-1.0, window=0, array=False):
        """Map every value of data (float32 or float64) with a homogeneous kernel
        map; returns data.shape + (2*order+1,).  array=True uses the array entry
        point (this library only), otherwise the scalar one per value."""
        d = np.ascontiguousarray(data)
        h = self.vl_homogeneouskernelmap_new(kernel, gamma, order, period, window)
        try:
            F = int(self.vl_homogeneouskernelmap_get_dimension(h))
            out = np.empty(d.shape + (F,), dtype=d.dtype)
            flat = d.ravel()
            if array:
                self.vl_homogeneouskernelmap_evaluate_array(h, dtype_code(d), out.ctypes.data, flat.ctypes.data,
                                                            flat.size)
            else:
                fn = self.vl_homogeneouskernelmap_evaluate_f if d.dtype == np.float32 else \
                    self.vl_homogeneouskernelmap_evaluate_d
                base = out.ctypes.data
                for i, v in enumerate(flat.tolist()):
                    fn(h, base + i * F * d.itemsize, 1, v)
            return out
        finally:
            self.vl_homogeneouskernelmap_delete(h)

    def kmeans_quantize(self, km, data):
        N = data.shape[0]
        a = np.empty(N, dtype=np.uint32)
        dist = np.empty(N, dtype=data.dtype)
        self.vl_kmeans_quantize(km, ptr(a), ptr(dist), ptr(data), N)
        return a, dist
