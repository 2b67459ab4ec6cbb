"""Posterior sparsity of the Fisher benchmark workload (SIFT-like 80-d
descriptors, K=1024 diagonal GMM): how much of the dense work a gate could
skip while keeping improved Fisher vectors within 1e-3 of the exact ones.

Not a test (pytest does not collect it).  Run: python tests/posterior_sparsity.py
It fits the GMM by 10 FP64 EM iterations on 200,000 descriptors (numpy, about
a minute and a half on 8 cores), then, on 4 images of 3,000 descriptors,
prints the exact top-M posterior mass and the error of Fisher vectors built
from the top M posteriors only, from posteriors above a threshold, and the
fraction of (32-row x 64-component) tiles whose posteriors are all below a
threshold.  Everything is FP64 numpy; no GPU is used.
"""
import os
import sys
import time

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(os.path.abspath(__file__)), ".."))
from vlfeat_amd import synth  # noqa: E402

K, D = 1024, 80
rng = np.random.default_rng(0)


def logits(x, m, c, p):
    iv = 1.0 / c
    b = np.log(np.maximum(p, 1e-300)) - 0.5 * D * np.log(2 * np.pi) - 0.5 * np.log(c).sum(1) - 0.5 * (m * m * iv).sum(1)
    return b[None, :] - 0.5 * (x * x) @ iv.T + x @ (m * iv).T


def post(x, m, c, p):
    l = logits(x, m, c, p)
    mx = l.max(1, keepdims=True)
    e = np.exp(l - mx)
    s = e.sum(1, keepdims=True)
    return e / s, (mx + np.log(s)).ravel()


def fit(x, iters=10):
    N = x.shape[0]
    m = x[rng.choice(N, K, replace=False)].astype(np.float64)
    c = np.tile(np.var(x, 0), (K, 1)).astype(np.float64)
    p = np.full(K, 1.0 / K)
    for it in range(iters):
        S0 = np.zeros(K); S1 = np.zeros((K, D)); S2 = np.zeros((K, D)); ll = 0
        for s in range(0, N, 20000):
            xs = x[s:s + 20000].astype(np.float64)
            q, l = post(xs, m, c, p)
            ll += l.sum()
            q[q < 0.01 / K] = 0
            S0 += q.sum(0); S1 += q.T @ xs; S2 += q.T @ (xs * xs)
        ok = S0 > 1e-6 / K
        mn = np.where(ok[:, None], S1 / np.maximum(S0, 1e-300)[:, None], m)
        c = np.where(ok[:, None], S2 / np.maximum(S0, 1e-300)[:, None] - mn * mn, c)
        c = np.maximum(c, 1e-6)
        m = mn
        p = S0 / S0.sum()
        print(f"EM iter {it}: LL {ll:.6g}", flush=True)
    return m, c, p


def fisher(x, m, c, p, q, improved=True):
    n = x.shape[0]
    isg = 1 / np.sqrt(c)
    S0 = q.sum(0)
    S1 = q.T @ x
    S2 = q.T @ (x * x)
    u = (S1 - m * S0[:, None]) * isg / (n * np.sqrt(p))[:, None]
    v = ((S2 - 2 * m * S1 + m * m * S0[:, None]) / c - S0[:, None]) / (n * np.sqrt(2 * p))[:, None]
    f = np.concatenate([u.ravel(), v.ravel()])
    if improved:
        f = np.sign(f) * np.sqrt(np.abs(f))
        f /= max(np.linalg.norm(f), 1e-12)
    return f


def rel(a, b):
    return np.linalg.norm(a - b) / np.linalg.norm(b)


t0 = time.time()
x = synth.sift_like(200000, D, seed=1).astype(np.float64)
m, c, p = fit(x, 10)
print(f"fit {time.time() - t0:.0f}s; prior range {p.min():.2e}..{p.max():.2e}")
xe = synth.sift_like(8 * 3000, D, seed=5).astype(np.float64)
for img in range(4):
    xi = xe[img * 3000:(img + 1) * 3000]
    q, lse = post(xi, m, c, p)
    qe = np.where(q < 1e-6, 0, q)
    fv = fisher(xi, m, c, p, qe)
    cnt = (q >= 1e-6).sum(1)
    qs = -np.sort(-q, axis=1)
    cm = np.cumsum(qs, axis=1)
    print(f"img {img}: terms>=1e-6 per row: mean {cnt.mean():.0f} min {cnt.min()} max {cnt.max()}")
    for M in (16, 32, 64, 128, 256, 512):
        mass = cm[:, M - 1]
        # top-M truncation, renormalised over the kept components
        thr = qs[:, M - 1:M]
        qt = np.where(q >= thr, q, 0)
        qt = qt / qt.sum(1, keepdims=True)
        qt = np.where(qt < 1e-6, 0, qt)
        e = rel(fisher(xi, m, c, p, qt), fv)
        print(f"   top-{M:3d}: mass mean {mass.mean():.5f} min {mass.min():.5f} "
              f"tail>1e-4 rows {(1 - mass > 1e-4).mean():.3f}; FV err {e:.2e}")
    for tau in (1e-5, 1e-4, 1e-3):
        qt = np.where(q < tau, 0, q)
        e = rel(fisher(xi, m, c, p, qt), fv)
        print(f"   drop q<{tau:g} (no renorm): FV err {e:.2e}, kept/row {(q >= tau).sum(1).mean():.0f}")
    # tile skipping: 32-row groups x 64-component blocks
    # component order: k-means of the means (16 clusters) then 64-chunks
    from scipy.cluster.vq import kmeans2
    _, lab = kmeans2(m, 16, seed=1, minit="++")
    order = np.argsort(lab, kind="stable")
    qo = q[:, order]
    for name, rows in (("data order", np.arange(3000)),
                       ("sorted by argmax", np.argsort(np.argmax(qo, 1), kind="stable"))):
        qr = qo[rows]
        G = 3000 // 32
        t = qr[:G * 32].reshape(G, 32, 16, 64).max(axis=(1, 3))
        for tau in (1e-6, 1e-5, 1e-4):
            print(f"   tiles {name}: skippable at max q < {tau:g}: {(t < tau).mean():.3f}")
