"""Learned gate of the Fisher encoder (numpy-level wrapper of vlhip_gate_*).

The network, its training and its inference run in the product library
(gate_kernels.hip); this module stages numpy arrays in library-owned device
memory and calls the C ABI.  There is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import lib


class DeviceArray:
    """A device copy of a numpy array (freed with the object)."""

    def __init__(self, a: np.ndarray | None = None, nbytes: int | None = None):
        L = lib()
        self.nbytes = int(a.nbytes if a is not None else nbytes)
        self.ptr = L.vlhip_malloc(max(self.nbytes, 16))
        if a is not None and self.nbytes:
            a = np.ascontiguousarray(a)
            L.vlhip_memcpy_h2d(self.ptr, a.ctypes.data, self.nbytes)

    def to_host(self, dtype, shape) -> np.ndarray:
        out = np.empty(shape, dtype=dtype)
        lib().vlhip_memcpy_d2h(out.ctypes.data, self.ptr, out.nbytes)
        return out

    def __del__(self):
        try:
            lib().vlhip_free(self.ptr)
        except Exception:
            pass


class Gate:
    """Perceptron D -> H1 -> H2 -> K (GELU) predicting GMM posteriors."""

    def __init__(self, D: int, K: int, H1: int = 256, H2: int = 256, seed: int = 0):
        L = lib()
        L.vlhip_init()
        self.D, self.K, self.H1, self.H2 = D, K, H1, H2
        self.handle = L.vlhip_gate_new(D, K, H1, H2, seed)
        self.num_params = int(L.vlhip_gate_num_params(self.handle))

    def __del__(self):
        try:
            lib().vlhip_gate_delete(self.handle)
        except Exception:
            pass

    # ---- parameters -----------------------------------------------------
    def params(self):
        P = np.empty(self.num_params, np.float32)
        mean = np.empty(self.D, np.float32)
        istd = np.empty(self.D, np.float32)
        lib().vlhip_gate_get_params(self.handle, P.ctypes.data, mean.ctypes.data, istd.ctypes.data)
        return P, mean, istd

    def set_params(self, P=None, mean=None, istd=None):
        def p(a):
            return None if a is None else np.ascontiguousarray(a, np.float32)
        P, mean, istd = p(P), p(mean), p(istd)
        lib().vlhip_gate_set_params(self.handle, None if P is None else P.ctypes.data,
                                    None if mean is None else mean.ctypes.data,
                                    None if istd is None else istd.ctypes.data)

    def split(self, P):
        """Views (W1, b1, W2, b2, W3, b3) of a flat parameter vector."""
        D, H1, H2, K = self.D, self.H1, self.H2, self.K
        sizes = [D * H1, H1, H1 * H2, H2, H2 * K, K]
        shapes = [(D, H1), (H1,), (H1, H2), (H2,), (H2, K), (K,)]
        out, o = [], 0
        for n, s in zip(sizes, shapes):
            out.append(P[o:o + n].reshape(s))
            o += n
        return out

    def standardize_from(self, x: np.ndarray):
        d = DeviceArray(np.ascontiguousarray(x, np.float32))
        lib().vlhip_gate_set_standardization_from_data(self.handle, d.ptr, x.shape[0])

    # ---- training -------------------------------------------------------
    def gradients(self, x: np.ndarray, q: np.ndarray):
        dx = DeviceArray(np.ascontiguousarray(x, np.float32))
        dq = DeviceArray(np.ascontiguousarray(q, np.float32))
        g = np.empty(self.num_params, np.float32)
        loss = lib().vlhip_gate_gradients(self.handle, dx.ptr, dq.ptr, x.shape[0], g.ctypes.data)
        return loss, g

    def train(self, x, means, covs, priors, steps: int, batch: int = 4096, lr: float = 1e-3):
        """Adam steps on the exact posteriors of the GMM; returns the loss trace.
        x may be a numpy array or a DeviceArray (with rows given by x.rows)."""
        if isinstance(x, DeviceArray):
            dx, n = x, x.rows
        else:
            dx, n = DeviceArray(np.ascontiguousarray(x, np.float32)), x.shape[0]
        dm = DeviceArray(np.ascontiguousarray(means, np.float32))
        dc = DeviceArray(np.ascontiguousarray(covs, np.float32))
        dp = DeviceArray(np.ascontiguousarray(priors, np.float32))
        trace = np.empty(steps, np.float32)
        lib().vlhip_gate_train(self.handle, dx.ptr, n, dm.ptr, dc.ptr, dp.ptr, steps, batch, C.c_float(lr),
                               trace.ctypes.data)
        return trace

    # ---- inference ------------------------------------------------------
    def logits(self, x: np.ndarray) -> np.ndarray:
        dx = DeviceArray(np.ascontiguousarray(x, np.float32))
        z = np.empty((x.shape[0], self.K), np.float32)
        lib().vlhip_gate_logits(self.handle, dx.ptr, x.shape[0], z.ctypes.data)
        return z

    def select(self, x: np.ndarray, M: int = 16, need: float = 0.99):
        """(idx [rows, M] int32, mass [rows], conf [rows] bool)."""
        n = x.shape[0]
        dx = DeviceArray(np.ascontiguousarray(x, np.float32))
        di = DeviceArray(nbytes=4 * n * M)
        dmass = DeviceArray(nbytes=4 * n)
        dconf = DeviceArray(nbytes=n)
        lib().vlhip_gate_select(self.handle, dx.ptr, n, M, C.c_float(need), di.ptr, dmass.ptr, dconf.ptr)
        return (di.to_host(np.int32, (n, M)), dmass.to_host(np.float32, (n,)),
                dconf.to_host(np.uint8, (n,)).astype(bool))


def bf16_round(a: np.ndarray) -> np.ndarray:
    """Round float32 values to the nearest bfloat16 (ties to even), as float32."""
    u = np.ascontiguousarray(a, np.float32).view(np.uint32).astype(np.uint64)
    lsb = (u >> 16) & 1
    r = ((u + 0x7FFF + lsb) >> 16) << 16
    return r.astype(np.uint32).view(np.float32)
