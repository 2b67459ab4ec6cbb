"""Checkers for the GPU library -- TEST INFRASTRUCTURE ONLY.

Imported only by ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline
leg of ``bench.py``; never by the product.

* ``Oracle``  -- ctypes binding of ``liboracle.so``, this repository's FP64
  restatement of VLFeat 0.9.21 (``oracle.c``; each routine cites the reference
  lines it follows).
* ``reference()`` -- the reference library itself, ``_ref/libvlref.so``,
  compiled from the reference sources by ``oracle/Makefile`` (float and double
  paths, OpenMP).  Absent on machines where it was not built.

Pinning of each oracle routine (what checks it beyond its own formula):
  or_gmm_posteriors   vl_test_fisher.m's closed form; scipy Gaussian-mixture
                      log-likelihood; the reference's double path
  or_gmm_maximization one-hot invariant (per-cluster mean / variance);
                      the reference's double path (EM trajectories)
  or_gmm_restart      forced empty mode vs the reference's double path
  or_gmm_em           the reference's double path
  or_fisher_encode    vl_test_fisher.m's five closed-form variants
  or_vlad_encode      vl_test_vlad.m's known answers and closed forms
  or_kmeans_quantize  brute force (numpy argmin)
  or_mlp_loss_grads   central finite differences of its own loss; softmax
                      cross-entropy identities (uniform logits, one-hot targets)
"""
from __future__ import annotations

import ctypes as C
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ORACLE_SO = HERE / "liboracle.so"
REF_SO = HERE / "_ref" / "libvlref.so"
REF_DIR = Path("/root/reference/vl")

_p = C.c_void_p
_sz = C.c_size_t


def build(with_reference: bool = True) -> None:
    """Compile liboracle.so and, when the reference sources exist, _ref/libvlref.so."""
    targets = ["oracle"]
    if with_reference and REF_DIR.exists():
        targets.append("ref")
    subprocess.run(["make", "-s", "-C", str(HERE), "-j8", *targets], check=True)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    def __init__(self, path=ORACLE_SO):
        if not Path(path).exists():
            build(with_reference=False)
        L = C.CDLL(str(path), mode=C.RTLD_LOCAL)
        sig = {
            "or_gmm_posteriors": (C.c_double, [_p, _sz, _sz, _p, _p, _sz, _p, _p]),
            "or_gmm_maximization": (None, [_p, _p, _sz, _sz, _sz, _p, _p, _p, _p]),
            "or_gmm_restart": (_sz, [_p, _p, _sz, _sz, _sz, _p, _p, _p, _p]),
            "or_gmm_em": (C.c_double, [_p, _sz, _sz, _sz, _sz, _p, _p, _p, _p, _p]),
            "or_fisher_encode": (_sz, [_p, _p, _sz, _sz, _p, _p, _p, _sz, C.c_int]),
            "or_vlad_encode": (None, [_p, _p, _sz, _sz, _p, _sz, _p, C.c_int]),
            "or_kmeans_quantize": (None, [_p, _p, _p, _p, _sz, _sz, _sz]),
            "or_mlp_loss_grads": (C.c_double, [_p, _sz, _sz, _sz, _sz, _p, _p, _sz, _p]),
            "or_homkermap": (None, [_p, _p, _sz, C.c_int, C.c_double, _sz, C.c_double, C.c_int]),
        }
        for n, (r, a) in sig.items():
            f = getattr(L, n)
            f.restype = r
            f.argtypes = a
        self.L = L

    def posteriors(self, x, means, covs, priors):
        x, means, covs, priors = map(_f64, (x, means, covs, priors))
        K, D = means.shape
        q = np.empty((x.shape[0], K))
        ll = self.L.or_gmm_posteriors(q.ctypes.data, K, x.shape[0], priors.ctypes.data, means.ctypes.data, D,
                                      covs.ctypes.data, x.ctypes.data)
        return q, ll

    def maximization(self, q, x, means, covs, priors, lbs):
        q, x = _f64(q), _f64(x)
        m, c, p = _f64(means).copy(), _f64(covs).copy(), _f64(priors).copy()
        K, D = m.shape
        lb = _f64(np.broadcast_to(lbs, (D,)))
        self.L.or_gmm_maximization(q.ctypes.data, x.ctypes.data, x.shape[0], D, K, p.ctypes.data, m.ctypes.data,
                                   c.ctypes.data, lb.ctypes.data)
        return m, c, p

    def em(self, x, means, covs, priors, max_iter, lbs=1e-4):
        x = _f64(x)
        m, c, p = _f64(means).copy(), _f64(covs).copy(), _f64(priors).copy()
        K, D = m.shape
        lb = _f64(np.broadcast_to(lbs, (D,)))
        q = np.empty((x.shape[0], K))
        ll = self.L.or_gmm_em(x.ctypes.data, x.shape[0], D, K, max_iter, p.ctypes.data, m.ctypes.data,
                              c.ctypes.data, lb.ctypes.data, q.ctypes.data)
        return m, c, p, ll, q

    def fisher(self, x, means, covs, priors, flags=0):
        x, means, covs, priors = map(_f64, (x, means, covs, priors))
        K, D = means.shape
        enc = np.empty(2 * K * D)
        n = self.L.or_fisher_encode(enc.ctypes.data, means.ctypes.data, D, K, covs.ctypes.data, priors.ctypes.data,
                                    x.ctypes.data, x.shape[0], flags)
        return enc, n

    def vlad(self, x, means, assign, flags=0):
        x, means, assign = map(_f64, (x, means, assign))
        K, D = means.shape
        enc = np.empty(K * D)
        self.L.or_vlad_encode(enc.ctypes.data, means.ctypes.data, D, K, x.ctypes.data, x.shape[0],
                              assign.ctypes.data, flags)
        return enc

    def mlp_loss_grads(self, params, D, H1, H2, K, xin, q, want_grads=True):
        """Mean soft-target cross-entropy of the gate perceptron and its gradient (FP64)."""
        params, xin, q = _f64(params), _f64(xin), _f64(q)
        g = np.empty(params.shape[0]) if want_grads else None
        loss = self.L.or_mlp_loss_grads(params.ctypes.data, D, H1, H2, K, xin.ctypes.data, q.ctypes.data,
                                        xin.shape[0], None if g is None else g.ctypes.data)
        return loss, g

    def homkermap(self, x, kernel=1, gamma=1.0, order=1, period=-1.0, window=0):
        """Homogeneous kernel map of every value of x: x.shape + (2*order+1,), FP64."""
        xf = _f64(x).ravel()
        out = np.empty((xf.size, 2 * order + 1))
        self.L.or_homkermap(out.ctypes.data, xf.ctypes.data, xf.size, kernel, gamma, order, period, window)
        return out.reshape(np.shape(x) + (2 * order + 1,))

    def quantize(self, x, centers):
        x, centers = _f64(x), _f64(centers)
        a = np.empty(x.shape[0], dtype=np.uint32)
        d = np.empty(x.shape[0])
        self.L.or_kmeans_quantize(a.ctypes.data, d.ctypes.data, x.ctypes.data, centers.ctypes.data, x.shape[0],
                                  x.shape[1], centers.shape[0])
        return a, d


_ORACLE = None
_REF = None


def oracle() -> Oracle:
    global _ORACLE
    if _ORACLE is None:
        _ORACLE = Oracle()
    return _ORACLE


def reference():
    """The compiled reference library, or None when it is not available."""
    global _REF
    if _REF is None and REF_SO.exists():
        from vlfeat_amd.capi import CLib
        _REF = CLib(REF_SO, extended=False)
    return _REF
