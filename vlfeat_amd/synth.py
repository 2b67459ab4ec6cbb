"""Seeded synthetic descriptor generators for the benchmark workloads.

* ``mixture``   -- x ~ sum_k pi_k N(mu_k, diag s_k^2), mu ~ U[-1,1]^D,
                   s ~ U[0.1, 0.5], pi ~ Dirichlet(1) (normalised -log U).
* ``sift_like`` -- z ~ N(0, I_128), rectified, L2-normalised, clipped at 0.2,
                   renormalised (SIFT's normalisation), then optionally
                   projected to D dimensions with a fixed orthonormal matrix
                   (rows of the QR factor of a seeded Gaussian).
* ``lognormal_counts`` -- per-image descriptor counts, round(exp(ln m + s z))
                   clamped to [lo, hi].

The benchmark inputs come from the C generator (``libvlsynth.so``, header
``include/vlsynth.h``): ``c_sift_like``, ``c_mixture``, ``c_projection``,
``c_lognormal_counts``.  Every value depends only on (seed, row), from
per-row SplitMix64 streams, the polar method and a logarithm built from basic
IEEE operations, so a C caller of the same functions gets the same bits, on
any thread count and for any chunking of the rows.  The numpy versions (same
distributions, other draws) serve the tests; ``sift_like_torch`` generates
the distribution on the GPU.
"""
from __future__ import annotations

import ctypes as _C
from pathlib import Path as _Path

import numpy as np

_SYNTH = None


def _synth():
    """The C generator library, built by vlfeat_amd.build.build_synth."""
    global _SYNTH
    if _SYNTH is None:
        path = _Path(__file__).resolve().parent / "lib" / "libvlsynth.so"
        if not path.exists():
            raise RuntimeError(f"{path} is not built (vlfeat_amd.build.build_synth())")
        lib = _C.CDLL(str(path))
        lib.vlsynth_sift_like.argtypes = [_C.c_void_p, _C.c_size_t, _C.c_int, _C.c_uint64, _C.c_uint64]
        lib.vlsynth_sift_like.restype = None
        lib.vlsynth_projection.argtypes = [_C.c_void_p, _C.c_int, _C.c_int, _C.c_uint64]
        lib.vlsynth_projection.restype = None
        lib.vlsynth_mixture.argtypes = [_C.c_void_p, _C.c_size_t, _C.c_int, _C.c_int, _C.c_uint64, _C.c_uint64,
                                        _C.c_void_p, _C.c_void_p, _C.c_void_p]
        lib.vlsynth_mixture.restype = None
        lib.vlsynth_lognormal_counts.argtypes = [_C.c_void_p, _C.c_size_t, _C.c_double, _C.c_double, _C.c_int64,
                                                 _C.c_int64, _C.c_uint64]
        lib.vlsynth_lognormal_counts.restype = None
        for name in ("vlsynth_log", "vlsynth_exp"):
            getattr(lib, name).argtypes = [_C.c_double]
            getattr(lib, name).restype = _C.c_double
        lib.vlsynth_uniform.argtypes = [_C.c_uint64, _C.c_uint64, _C.c_uint64, _C.c_uint64]
        lib.vlsynth_uniform.restype = _C.c_double
        _SYNTH = lib
    return _SYNTH


def c_sift_like(n: int, dim: int = 80, seed: int = 1, row_offset: int = 0, out: np.ndarray | None = None) -> np.ndarray:
    """Rows [row_offset, row_offset + n) of the SIFT-like workload (C generator), float32 n x dim."""
    if out is None:
        out = np.empty((n, dim), np.float32)
    assert out.dtype == np.float32 and out.shape == (n, dim) and out.flags.c_contiguous
    _synth().vlsynth_sift_like(out.ctypes.data, n, dim, seed, row_offset)
    return out


def c_projection(dim_out: int = 80, dim_in: int = 128, seed: int = 1) -> np.ndarray:
    """The SIFT-like workload's projection (orthonormal rows), float64 dim_out x dim_in."""
    P = np.empty((dim_out, dim_in), np.float64)
    _synth().vlsynth_projection(P.ctypes.data, dim_out, dim_in, seed)
    return P


def c_mixture(n: int, dim: int = 64, k_true: int = 64, seed: int = 0, row_offset: int = 0):
    """(x float32 n x dim, means, sigmas, weights) of the mixture workload (C generator)."""
    x = np.empty((n, dim), np.float32)
    mu = np.empty((k_true, dim), np.float64)
    sg = np.empty((k_true, dim), np.float64)
    w = np.empty(k_true, np.float64)
    _synth().vlsynth_mixture(x.ctypes.data, n, dim, k_true, seed, row_offset, mu.ctypes.data, sg.ctypes.data,
                             w.ctypes.data)
    return x, mu, sg, w


def c_lognormal_counts(num_images: int, median: float = 3000, sigma: float = 0.5, lo: int = 200, hi: int = 20000,
                       seed: int = 4) -> np.ndarray:
    """Per-image descriptor counts of the uneven-image workload (C generator)."""
    c = np.empty(num_images, np.int64)
    _synth().vlsynth_lognormal_counts(c.ctypes.data, num_images, float(median), float(sigma), lo, hi, seed)
    return c


def mixture(n, dim=64, k_true=64, seed=0, dtype=np.float32):
    rng = np.random.default_rng(seed)
    mu = rng.uniform(-1, 1, size=(k_true, dim))
    s = rng.uniform(0.1, 0.5, size=(k_true, dim))
    w = -np.log(rng.uniform(size=k_true))
    w /= w.sum()
    comp = rng.choice(k_true, size=n, p=w)
    x = mu[comp] + s[comp] * rng.standard_normal((n, dim))
    return x.astype(dtype)


def projection(dim_out=80, dim_in=128, seed=1):
    rng = np.random.default_rng(seed + 1000)
    g = rng.standard_normal((dim_in, dim_in))
    qm, _ = np.linalg.qr(g)
    return qm[:dim_out].astype(np.float64)


def _sift_normalise(z):
    z = np.maximum(z, 0.0)
    z /= np.maximum(np.linalg.norm(z, axis=1, keepdims=True), 1e-12)
    z = np.minimum(z, 0.2)
    z /= np.maximum(np.linalg.norm(z, axis=1, keepdims=True), 1e-12)
    return z


def sift_like(n, dim=80, seed=1, dtype=np.float32, project=True):
    rng = np.random.default_rng(seed)
    z = _sift_normalise(rng.standard_normal((n, 128)))
    if project and dim != 128:
        z = z @ projection(dim, 128, seed).T
    return z.astype(dtype)


def lognormal_counts(num_images, median=3000, sigma=0.5, lo=200, hi=20000, seed=4):
    rng = np.random.default_rng(seed)
    n = np.round(np.exp(np.log(median) + sigma * rng.standard_normal(num_images)))
    return np.clip(n, lo, hi).astype(np.int64)


def sift_like_torch(n, dim=80, seed=1, device="cuda", chunk=1 << 21):
    """Same distribution as ``sift_like`` generated on the GPU (float32)."""
    import torch
    gen = torch.Generator(device=device)
    gen.manual_seed(seed)
    P = torch.as_tensor(projection(dim, 128, seed), dtype=torch.float32, device=device)
    out = torch.empty((n, dim), dtype=torch.float32, device=device)
    for s in range(0, n, chunk):
        e = min(n, s + chunk)
        z = torch.randn((e - s, 128), generator=gen, device=device)
        z = torch.clamp_min(z, 0)
        z = z / torch.clamp_min(z.norm(dim=1, keepdim=True), 1e-12)
        z = torch.clamp_max(z, 0.2)
        z = z / torch.clamp_min(z.norm(dim=1, keepdim=True), 1e-12)
        out[s:e] = z @ P.T if dim != 128 else z
    return out
