"""Seeded synthetic descriptor generators for the benchmark workloads.

* ``mixture``   -- x ~ sum_k pi_k N(mu_k, diag s_k^2), mu ~ U[-1,1]^D,
                   s ~ U[0.1, 0.5], pi ~ Dirichlet(1) (normalised -log U).
* ``sift_like`` -- z ~ N(0, I_128), rectified, L2-normalised, clipped at 0.2,
                   renormalised (SIFT's normalisation), then optionally
                   projected to D dimensions with a fixed orthonormal matrix
                   (rows of the QR factor of a seeded Gaussian).
* ``lognormal_counts`` -- per-image descriptor counts, round(exp(ln m + s z))
                   clamped to [lo, hi].

numpy versions serve tests and small runs; ``sift_like_torch`` generates the
same distribution on the GPU for the large benchmark inputs.
"""
from __future__ import annotations

import numpy as np


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
