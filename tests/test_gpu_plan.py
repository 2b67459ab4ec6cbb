"""Encode plans (batched uneven images, autotuned tiling, hipGraph replay):
every replay -- with the library's own eager work and synchronisations in
between -- equals the eager batched encoder (compensated FP16 Fisher vectors)
and the hard-assignment VLAD of the nearest means, and stays within the
FP64 oracle's tolerance."""
import ctypes as C

import numpy as np
import pytest

import vlfeat_amd as va
from oracle import oracle
from vlfeat_amd import synth
from vlfeat_amd.gate import DeviceArray

pytestmark = pytest.mark.gpu


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gmm_params(x, K, seed):
    rng = np.random.default_rng(seed)
    m = x[rng.choice(x.shape[0], K, replace=False)].astype(np.float32)
    c = (np.var(x, 0)[None, :] * rng.uniform(0.1, 0.5, (K, x.shape[1]))).astype(np.float32)
    p = rng.dirichlet(np.ones(K)).astype(np.float32)
    return m, c, p


@pytest.mark.parametrize("K,flags,vflags", [(192, 3, 0), (64, 7, 1), (300, 0, 2)])
def test_plan_replays_match_eager(K, flags, vflags):
    L = va.lib()
    rng = np.random.default_rng(K)
    counts = np.clip(np.round(np.exp(np.log(900) + 0.6 * rng.standard_normal(37))), 1, 4000).astype(np.int64)
    counts[3] = 0
    counts[5] = 1
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    N = int(off[-1])
    x = synth.sift_like(N, 80, seed=K)
    m, c, p = gmm_params(x, K, seed=K + 1)
    n = len(counts)

    with va.fisher_precision(va.MODE_F16X3):
        e_eager, t_eager = va.vl_fisher_batch(x, off, m, c, p, flags)
    assign, _ = va.vl_kmeans_quantize(x, m)
    v_eager = np.stack([va.vl_vlad_hard(x[off[i]:off[i + 1]], m, assign[off[i]:off[i + 1]], vflags)
                        for i in range(n)])

    dx, dm, dc, dp = (DeviceArray(a) for a in (x, m, c, p))
    denc = DeviceArray(nbytes=4 * n * 2 * K * 80)
    dvlad = DeviceArray(nbytes=4 * n * K * 80)
    dterms = DeviceArray(nbytes=8 * n)
    plan = L.vlhip_encode_plan_new(80, K, off.ctypes.data, n, flags, vflags)
    try:
        for rep in range(3):
            L.vlhip_encode_plan_run(plan, dx.ptr, dm.ptr, dc.ptr, dp.ptr, denc.ptr, dvlad.ptr, dterms.ptr)
            enc = denc.to_host(np.float32, (n, 2 * K * 80))
            vl = dvlad.to_host(np.float32, (n, K * 80))
            terms = dterms.to_host(np.uint64, (n,))
            variant = C.c_int()
            ms = (C.c_double * 2)()
            nb = C.c_uint64()
            L.vlhip_encode_plan_info(plan, C.byref(variant), ms, C.byref(nb))
            print(f"K={K} replay {rep}: tiling {variant.value} (candidates {ms[0]:.3f} / {ms[1]:.3f} ms), "
                  f"{nb.value} batch(es)")
            default_tiling = 0 if K > 128 else 1
            if variant.value == default_tiling:
                assert np.array_equal(enc, e_eager), rep
            else:  # a different component split only changes the order of the norm's partial sums
                assert max(rel(enc[i], e_eager[i]) for i in range(n) if counts[i]) < 1e-6
            np.testing.assert_array_equal(terms, t_eager)
            assert np.array_equal(vl, v_eager), rep
            # eager library work between replays
            va.vl_fisher(x[:500], m, c, p, 3)
            L.vlhip_synchronize()
    finally:
        L.vlhip_encode_plan_delete(plan)
    for i in (0, 1, 2):
        eo, _ = oracle().fisher(x[off[i]:off[i + 1]], m, c, p, flags)
        assert rel(enc[i], eo) < 2e-4


def test_plan_replay_survives_library_table_growth():
    # the plan's quantizer table is its own: a library call that grows the
    # library's quantizer table between replays (more centres, wider
    # descriptors) leaves the replayed encodings unchanged bit for bit
    L = va.lib()
    K, D = 96, 80
    counts = np.array([700, 1, 2500, 64, 1300], np.int64)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    n = len(counts)
    x = synth.sift_like(int(off[-1]), D, seed=21)
    m, c, p = gmm_params(x, K, seed=22)
    dx, dm, dc, dp = (DeviceArray(a) for a in (x, m, c, p))
    denc = DeviceArray(nbytes=4 * n * 2 * K * D)
    dvlad = DeviceArray(nbytes=4 * n * K * D)
    dterms = DeviceArray(nbytes=8 * n)
    plan = L.vlhip_encode_plan_new(D, K, off.ctypes.data, n, 3, 1)
    try:
        L.vlhip_encode_plan_run(plan, dx.ptr, dm.ptr, dc.ptr, dp.ptr, denc.ptr, dvlad.ptr, dterms.ptr)
        e0 = denc.to_host(np.float32, (n, 2 * K * D))
        v0 = dvlad.to_host(np.float32, (n, K * D))
        for Kb, Db in ((1000, 120), (2048, 128)):
            xb = synth.sift_like(5000, Db, seed=Kb, project=Db != 128)
            va.vl_kmeans_quantize(xb, xb[:Kb])
            L.vlhip_encode_plan_run(plan, dx.ptr, dm.ptr, dc.ptr, dp.ptr, denc.ptr, dvlad.ptr, dterms.ptr)
            assert np.array_equal(denc.to_host(np.float32, (n, 2 * K * D)), e0), Kb
            assert np.array_equal(dvlad.to_host(np.float32, (n, K * D)), v0), Kb
    finally:
        L.vlhip_encode_plan_delete(plan)
    assign, _ = va.vl_kmeans_quantize(x, m)
    for i in range(n):
        assert np.array_equal(v0[i], va.vl_vlad_hard(x[off[i]:off[i + 1]], m, assign[off[i]:off[i + 1]], 1)), i
