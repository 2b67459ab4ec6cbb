"""MI355X-native VLFeat GMM / Fisher / VLAD / k-means library.

The product is the C shared library ``vlfeat_amd/lib/libvl.so`` (VLFeat's C
API in ``include/vl/*.h`` plus the device ABI in ``include/vlhip.h``).  This
package only loads it and offers numpy-level wrappers shaped like the
reference's MATLAB front-ends (``vl_gmm``, ``vl_fisher``, ``vl_vlad``,
``vl_kmeans``).  There is no Python or CPU fallback: if the library is not
built, or no GPU is present when a GPU entry point runs, the call fails.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .capi import (  # noqa: F401
    CLib, VL_TYPE_DOUBLE, VL_TYPE_FLOAT, FISHER_FAST, FISHER_IMPROVED, FISHER_NORMALIZED,
    FISHER_SQUARE_ROOT, VLAD_NORMALIZE_COMPONENTS, VLAD_NORMALIZE_MASS, VLAD_SQUARE_ROOT,
    VLAD_UNNORMALIZED, MODE_BF16X3, MODE_F16X3, MODE_FP32, MODE_FP32_MFMA, MODE_GATED, VlKMeansANN, VlKMeansElkan,
    VlKMeansLloyd, VlKMeansPlusPlus, VlKMeansRandomSelection, dtype_code, ptr,
)

LIB_PATH = Path(__file__).resolve().parent / "lib" / "libvl.so"
_LIB: CLib | None = None


def lib() -> CLib:
    """The loaded product library (raises if it has not been built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m vlfeat_amd.build`")
        _LIB = CLib(LIB_PATH, extended=True)
    return _LIB


def gpu_available() -> bool:
    return bool(lib().vlhip_available())


# ---- MATLAB-style front-ends -------------------------------------------------

def vl_gmm(data, K, **kw):
    """GMM by EM (toolbox/gmm/vl_gmm.c). Returns (means, covariances, priors, LL, posteriors)."""
    return lib().gmm(np.ascontiguousarray(data), K, **kw)


def vl_fisher(data, means, covariances, priors, flags=0):
    """Fisher vector of one image (toolbox/fisher/vl_fisher.c). Returns (enc, numTerms)."""
    d = np.ascontiguousarray(data)
    return lib().fisher(d, np.ascontiguousarray(means, d.dtype), np.ascontiguousarray(covariances, d.dtype),
                        np.ascontiguousarray(priors, d.dtype), flags)


class fisher_precision:
    """Context manager selecting the float Fisher arithmetic (MODE_FP32,
    MODE_FP32_MFMA, MODE_BF16X3, MODE_F16X3 or MODE_GATED) for the calling thread."""

    def __init__(self, mode):
        self.mode = mode

    def __enter__(self):
        self.prev = lib().vl_fisher_get_precision()
        lib().vl_fisher_set_precision(self.mode)
        return self

    def __exit__(self, *exc):
        lib().vl_fisher_set_precision(self.prev)


def set_fisher_gate(gate, candidates: int = 16, need: float = 0.99):
    """Gate used by MODE_GATED encodings (a vlfeat_amd.gate.Gate, or None)."""
    lib().vl_fisher_set_gate(None if gate is None else gate.handle, candidates, need)


def fisher_gate_stats(reset: bool = True):
    """(groups, served): descriptor groups seen by gated encodings and how many the gate served."""
    g, s = C.c_ulonglong(), C.c_ulonglong()
    lib().vlhip_fisher_gate_stats(C.byref(g), C.byref(s), int(reset))
    return g.value, s.value


def vl_fisher_batch(data, offsets, means, covariances, priors, flags=0):
    """Fisher vectors of many images stored back to back; returns (enc[numImages, 2KD], numTerms)."""
    d = np.ascontiguousarray(data)
    off = np.ascontiguousarray(offsets, dtype=np.uint64)
    n = len(off) - 1
    K, D = means.shape
    enc = np.empty((n, 2 * K * D), dtype=d.dtype)
    terms = np.empty(n, dtype=np.uint64)
    m = np.ascontiguousarray(means, d.dtype)
    c = np.ascontiguousarray(covariances, d.dtype)
    p = np.ascontiguousarray(priors, d.dtype)
    lib().vl_fisher_encode_batch(ptr(enc), dtype_code(d), ptr(m), D, K, ptr(c), ptr(p), ptr(d), ptr(off), n,
                                 flags, ptr(terms))
    return enc, terms


def vl_vlad(data, means, assignments, flags=0):
    """VLAD of one image with dense soft assignments (toolbox/vlad/vl_vlad.c)."""
    d = np.ascontiguousarray(data)
    return lib().vlad(d, np.ascontiguousarray(means, d.dtype), np.ascontiguousarray(assignments, d.dtype), flags)


def vl_vlad_hard(data, means, assignment, flags=0):
    """VLAD from hard assignments (cluster index per datum)."""
    d = np.ascontiguousarray(data)
    K, D = means.shape
    m = np.ascontiguousarray(means, d.dtype)
    a = np.ascontiguousarray(assignment, dtype=np.uint32)
    enc = np.empty(K * D, dtype=d.dtype)
    lib().vl_vlad_encode_hard(ptr(enc), dtype_code(d), ptr(m), D, K, ptr(d), d.shape[0], ptr(a), flags)
    return enc


def vl_kmeans_quantize(data, centers):
    """Exact nearest-center assignment and squared distance (vl_kmeans_quantize)."""
    L = lib()
    d = np.ascontiguousarray(data)
    km = L.kmeans_new(d.dtype)
    try:
        c = np.ascontiguousarray(centers, d.dtype)
        L.vl_kmeans_set_centers(km, ptr(c), c.shape[1], c.shape[0])
        return L.kmeans_quantize(km, d)
    finally:
        L.vl_kmeans_delete(km)


def vl_kmeans(data, K, algorithm=VlKMeansLloyd, init=VlKMeansRandomSelection, max_iter=100,
              num_trees=3, max_comparisons=100, reps=1):
    """k-means clustering (toolbox/kmeans/vl_kmeans.c); returns (centers, energy)."""
    L = lib()
    d = np.ascontiguousarray(data)
    km = L.kmeans_new(d.dtype, algorithm, init, max_iter, num_trees, max_comparisons, reps)
    try:
        e = L.vl_kmeans_cluster(km, ptr(d), d.shape[1], d.shape[0], K)
        return L.kmeans_centers(km, d.dtype), e
    finally:
        L.vl_kmeans_delete(km)


# ---- device timing ------------------------------------------------------------

def kernel_timings() -> dict[str, tuple[float, int]]:
    """Accumulated per-kernel milliseconds and launch counts (hipEvent timing)."""
    L = lib()
    out = {}
    name = C.create_string_buffer(128)
    for i in range(L.vlhip_timing_count()):
        ms = C.c_double()
        cnt = C.c_ulonglong()
        L.vlhip_timing_get(i, name, 128, C.byref(ms), C.byref(cnt))
        out[name.value.decode()] = (ms.value, cnt.value)
    return out
