"""Parity at the real shapes of the benchmark configurations (K, D and image
sizes as the workloads define them; fewer images or descriptors where only
the count changes, so that the FP64 oracle finishes in seconds).

  config 2  FP32 drop-in, D=80, K=256: vl_gmm_cluster (30 EM iterations) on
            50,000 SIFT-like descriptors against the reference's float EM
            (means / covariances / priors within 1e-4 relative L2, LL 1e-5);
            FP32 Fisher (every flag set) and VLAD (one-hot assignments from the
            exact quantizer) of 3,000-descriptor images against the reference
            (Fisher 1e-4 relative L2, VLAD and assignments bit for bit).
  config 3  precision form, D=128, K=1024: one matrix-core EM step (F16X3 and
            BF16X3) on 65,536 descriptors against the FP64 oracle's posteriors:
            LL 1e-6 / 1e-5 relative, means / covariances / priors after the
            M-step within 1e-4 / 1e-3 relative L2 (max elementwise 1e-3 / 5e-3).
  config 4  batched uneven encoding, D=80, K=512: an encode plan over 300
            images with log-normal sizes (median 3,000, sigma 0.5, clamped to
            [200, 20,000]) equals the eager batched encoders (Fisher bit for
            bit at the same tiling, hard VLAD bit for bit), the Fisher vectors
            of three images are within 1e-3 of the FP64 oracle and their VLAD
            vectors bit for bit the reference's.
  config 5  the metric's encoder shape, D=80, K=1024, 3,000-descriptor
            images: FP32 GPU within 1e-4 of the reference float path; the
            compensated F16X3 encoder and the gated encoder within 1e-3 of the
            FP64 oracle on every one of 8 images (the maximum, not a median).
"""
import concurrent.futures as cf

import numpy as np
import pytest

import vlfeat_amd as va
from oracle import oracle, reference
from vlfeat_amd import synth
from vlfeat_amd.gate import DeviceArray

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.skipif(reference() is None, reason="reference library not built")


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def start(x, K, seed):
    """Seeded EM start: random data points, data variance, uniform priors."""
    rng = np.random.default_rng(seed)
    m = x[rng.choice(x.shape[0], K, replace=False)].copy()
    c = np.tile(np.var(x, 0), (K, 1)).astype(np.float32)
    p = np.full(K, 1.0 / K, np.float32)
    return m, c, p


def fit(x, K, iters, seed):
    m, c, p = start(x, K, seed)
    m, c, p, _, _ = va.vl_gmm(x, K, init="custom", means=m, covs=c, priors=p, max_iter=iters, want_post=False)
    return m, c, p


def oracle_fisher_many(xs, m, c, p, flags):
    """FP64 oracle encodings of several images on host threads."""
    O = oracle()
    with cf.ThreadPoolExecutor(8) as ex:
        return list(ex.map(lambda a: O.fisher(a, m, c, p, flags), xs))


def oracle_posteriors_threaded(x, m, c, p, parts=16):
    O = oracle()
    b = np.linspace(0, x.shape[0], parts + 1).astype(int)
    with cf.ThreadPoolExecutor(8) as ex:
        res = list(ex.map(lambda i: O.posteriors(x[b[i]:b[i + 1]], m, c, p), range(parts)))
    return np.concatenate([q for q, _ in res]), float(sum(l for _, l in res))


# ------------------------------------------------------------------ config 2

@pytest.fixture(scope="module")
def cfg2():
    x = synth.sift_like(50000, 80, seed=1)
    xe = synth.sift_like(4 * 3000, 80, seed=2)
    return x, xe


@needs_ref
def test_config2_gmm_cluster_fp32_vs_reference(cfg2):
    x, _ = cfg2
    K = 256
    m0, c0, p0 = start(x, K, 7)
    R = reference()
    R.set_num_threads(0)
    g = va.lib().gmm(x, K, init="custom", means=m0, covs=c0, priors=p0, max_iter=30, want_post=False)
    r = R.gmm(x, K, init="custom", means=m0, covs=c0, priors=p0, max_iter=30, want_post=False)
    errs = [rel(g[i], r[i]) for i in range(3)]
    print(f"config 2 EM (N=50000, K=256, 30 iterations): means {errs[0]:.2e}, covariances {errs[1]:.2e}, "
          f"priors {errs[2]:.2e}, LL {abs(g[3] - r[3]) / abs(r[3]):.2e}")
    assert max(errs) < 1e-4
    assert abs(g[3] - r[3]) / abs(r[3]) < 1e-5


@needs_ref
@pytest.mark.parametrize("flags", [0, 1, 2, 3, 7])
def test_config2_fisher_fp32_vs_reference(cfg2, flags):
    x, xe = cfg2
    m, c, p = fit(x, 256, 10, 3)
    R = reference()
    R.set_num_threads(0)
    for i in range(4):
        xi = xe[i * 3000:(i + 1) * 3000]
        e, n = va.vl_fisher(xi, m, c, p, flags)
        er, nr = R.fisher(xi, m, c, p, flags)
        assert rel(e, er) < 1e-4, (i, rel(e, er))
        # the term count differs from the reference's only for posteriors that
        # round to either side of 1e-6 (float E-step arithmetic in another order)
        assert abs(int(n) - int(nr)) <= max(2, nr // 10000), (n, nr)


@needs_ref
@pytest.mark.parametrize("vflags", [0, 1, 2, 8, 15])
def test_config2_vlad_fp32_vs_reference(cfg2, vflags):
    x, xe = cfg2
    m, _, _ = fit(x, 256, 5, 3)
    R = reference()
    R.set_num_threads(0)
    a, _ = va.vl_kmeans_quantize(xe, m)
    km = R.kmeans_new(np.float32)
    R.vl_kmeans_set_centers(km, m.ctypes.data, 80, 256)
    ar, _ = R.kmeans_quantize(km, xe)
    R.vl_kmeans_delete(km)
    assert np.array_equal(a, ar)
    for i in range(4):
        xi = xe[i * 3000:(i + 1) * 3000]
        oh = np.zeros((3000, 256), np.float32)
        oh[np.arange(3000), a[i * 3000:(i + 1) * 3000]] = 1
        v = va.vl_vlad(xi, m, oh, vflags)
        vr = R.vlad(xi, m, oh, vflags)
        assert np.array_equal(v, vr), i


# ------------------------------------------------------------------ config 3

@pytest.fixture(scope="module")
def cfg3():
    x = synth.sift_like(100000, 128, seed=1, project=False)
    m, c, p = fit(x, 1024, 3, 0)
    xs = x[:65536]
    qo, llo = oracle_posteriors_threaded(xs, m, c, p)
    return xs, m, c, p, qo, llo


def em_finalize(S0, S1, S2, m_old, K, lb=1e-4):
    """M-step finalisation of gmm.c:1169-1202 in FP64, applied alike to the
    GPU's and the oracle's moments."""
    upd = (S0 >= 1e-6 / K)[:, None]
    s0 = np.where(S0 > 0, S0, 1.0)[:, None]
    mu = np.where(upd, S1 / s0, S1)
    cv = np.maximum(np.where(upd, S2 / s0 - (mu - m_old) ** 2, S2), lb)
    return mu, cv, S0 / max(S0.sum(), 1e-12)


@pytest.mark.parametrize("mode,bounds", [(va.MODE_F16X3, (1e-6, 1e-4, 1e-3)), (va.MODE_BF16X3, (1e-5, 1e-3, 5e-3))],
                         ids=["f16x3", "bf16x3"])
def test_config3_em_step_d128_k1024_vs_fp64(cfg3, mode, bounds):
    xs, m, c, p, qo, llo = cfg3
    N, D, K = xs.shape[0], 128, 1024
    L = va.lib()
    dx, dm, dc, dp = (DeviceArray(a) for a in (xs, m, c, p))
    out = np.zeros(2 * K + 2 * K * D)
    ll = L.vlhip_gmm_em_step(mode, out.ctypes.data, dp.ptr, dm.ptr, dc.ptr, D, K, dx.ptr, N)
    KD = K * D
    m64 = m.astype(np.float64)
    got = em_finalize(out[:K], out[2 * K:2 * K + KD].reshape(K, D), out[2 * K + KD:].reshape(K, D), m64, K)
    xo = xs.astype(np.float64)
    qt = np.where(qo < 0.01 / K, 0.0, qo)
    S0 = qt.sum(0)
    S1 = qt.T @ xo
    S2 = qt.T @ (xo * xo) - 2 * m64 * S1 + m64 * m64 * S0[:, None]
    want = em_finalize(S0, S1, S2, m64, K)
    ll_err = abs(ll - llo) / abs(llo)
    errs = [rel(g, w) for g, w in zip(got, want)]
    maxrel = [float(np.max(np.abs(g - w) / np.maximum(np.abs(w), 1e-30))) for g, w in zip(got[:2], want[:2])]
    print(f"config 3 EM step mode {mode}: LL {ll_err:.2e}, means {errs[0]:.2e} (max {maxrel[0]:.2e}), "
          f"covariances {errs[1]:.2e} (max {maxrel[1]:.2e}), priors {errs[2]:.2e}, "
          f"mass {rel(out[K:2 * K], qo.sum(0)):.2e}")
    assert ll_err < bounds[0]
    assert max(errs) < bounds[1]
    assert max(maxrel) < bounds[2]
    assert rel(out[K:2 * K], qo.sum(0)) < bounds[1]


# ------------------------------------------------------------------ config 4

@pytest.fixture(scope="module")
def cfg4():
    counts = synth.lognormal_counts(300, median=3000, sigma=0.5, lo=200, hi=20000, seed=4)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    x = synth.sift_like(int(off[-1]), 80, seed=2)
    m, c, p = fit(x[:200000], 512, 5, 0)
    return x, off, m, c, p


def test_config4_plan_k512_lognormal_vs_eager_and_fp64(cfg4):
    import ctypes as C
    x, off, m, c, p = cfg4
    K, D, n = 512, 80, len(off) - 1
    L = va.lib()
    with va.fisher_precision(va.MODE_F16X3):
        e_eager, t_eager = va.vl_fisher_batch(x, off, m, c, p, 3)
    assign, _ = va.vl_kmeans_quantize(x, m)
    dx, dm, dc, dp = (DeviceArray(a) for a in (x, m, c, p))
    denc = DeviceArray(nbytes=4 * n * 2 * K * D)
    dvlad = DeviceArray(nbytes=4 * n * K * D)
    dterms = DeviceArray(nbytes=8 * n)
    plan = L.vlhip_encode_plan_new(D, K, off.ctypes.data, n, 3, 0)
    try:
        for rep in range(2):
            L.vlhip_encode_plan_run(plan, dx.ptr, dm.ptr, dc.ptr, dp.ptr, denc.ptr, dvlad.ptr, dterms.ptr)
        enc = denc.to_host(np.float32, (n, 2 * K * D))
        vl = dvlad.to_host(np.float32, (n, K * D))
        terms = dterms.to_host(np.uint64, (n,))
        variant = C.c_int()
        ms = (C.c_double * 2)()
        nb = C.c_uint64()
        L.vlhip_encode_plan_info(plan, C.byref(variant), ms, C.byref(nb))
    finally:
        L.vlhip_encode_plan_delete(plan)
    print(f"config 4 plan: {n} images, {int(off[-1])} descriptors, tiling {variant.value} "
          f"({ms[0]:.2f} / {ms[1]:.2f} ms), {nb.value} batch(es)")
    if variant.value == 0:  # the eager encoder's tiling at K > 128
        assert np.array_equal(enc, e_eager)
    else:
        assert max(rel(enc[i], e_eager[i]) for i in range(n)) < 1e-6
    np.testing.assert_array_equal(terms, t_eager)
    sel = [0, int(np.argmax(np.diff(off))), int(np.argmin(np.diff(off)))]
    for i in sel:
        xi = x[off[i]:off[i + 1]]
        v = va.vl_vlad_hard(xi, m, assign[off[i]:off[i + 1]], 0)
        assert np.array_equal(vl[i], v), i
    ref = oracle_fisher_many([x[off[i]:off[i + 1]] for i in sel], m, c, p, 3)
    errs = [rel(enc[i], r[0]) for i, r in zip(sel, ref)]
    print(f"config 4 Fisher vs FP64 oracle, images {sel} (sizes {[int(off[i + 1] - off[i]) for i in sel]}): "
          + ", ".join(f"{e:.2e}" for e in errs))
    assert max(errs) < 1e-3
    R = reference()
    if R is not None:
        for i in sel:
            xi = x[off[i]:off[i + 1]]
            oh = np.zeros((xi.shape[0], K), np.float32)
            oh[np.arange(xi.shape[0]), assign[off[i]:off[i + 1]]] = 1
            assert np.array_equal(vl[i], R.vlad(xi, m, oh, 0)), i


# ------------------------------------------------------------------ config 5

@pytest.fixture(scope="module")
def cfg5():
    x = synth.sift_like(200000, 80, seed=1)
    m, c, p = fit(x, 1024, 10, 0)
    xe = synth.sift_like(8 * 3000, 80, seed=5)
    off = (np.arange(9) * 3000).astype(np.uint64)
    ref = oracle_fisher_many([xe[int(off[i]):int(off[i + 1])] for i in range(8)], m, c, p, 3)
    return x, xe, off, m, c, p, ref


def test_config5_fp32_vs_reference(cfg5):
    _, xe, off, m, c, p, ref = cfg5
    with va.fisher_precision(va.MODE_FP32):
        e, t = va.vl_fisher_batch(xe, off, m, c, p, 3)
    errs = [rel(e[i], ref[i][0]) for i in range(8)]
    print(f"config 5 FP32 GPU vs FP64 oracle: max {max(errs):.2e}")
    assert max(errs) < 1e-4
    R = reference()
    if R is not None:
        R.set_num_threads(0)
        er = [rel(e[i], R.fisher(xe[int(off[i]):int(off[i + 1])], m, c, p, 3)[0]) for i in range(8)]
        print(f"config 5 FP32 GPU vs reference float: max {max(er):.2e}")
        assert max(er) < 1e-4


def test_config5_fp32_matrix_cores_vs_reference(cfg5):
    """The FP32 matrix-core encoder (the bench's FP32 comparator) at the
    metric's shape: every image within 1e-4 of the FP64 oracle and of the
    reference's float path."""
    _, xe, off, m, c, p, ref = cfg5
    with va.fisher_precision(va.MODE_FP32_MFMA):
        e, t = va.vl_fisher_batch(xe, off, m, c, p, 3)
    errs = [rel(e[i], ref[i][0]) for i in range(8)]
    print(f"config 5 FP32 matrix cores vs FP64 oracle: per image " + ", ".join(f"{v:.2e}" for v in errs))
    assert max(errs) < 1e-4
    for i in range(8):
        assert abs(int(t[i]) - int(ref[i][1])) <= max(2, int(ref[i][1]) // 10000)
    R = reference()
    if R is not None:
        R.set_num_threads(0)
        rf = [R.fisher(xe[int(off[i]):int(off[i + 1])], m, c, p, 3)[0] for i in range(8)]
        er = [rel(e[i], rf[i]) for i in range(8)]
        eref = [rel(rf[i], ref[i][0]) for i in range(8)]
        print(f"config 5 FP32 matrix cores vs reference float: max {max(er):.2e} "
              f"(reference vs FP64: max {max(eref):.2e})")
        # the reference's float L2 norm (a sequential float sum) is itself off
        # in scale by up to ~1e-4 at this size
        assert all(a < 1e-4 + b for a, b in zip(er, eref))


@pytest.mark.parametrize("mode", [va.MODE_F16X3, va.MODE_BF16X3], ids=["f16x3", "bf16x3"])
def test_config5_split16_vs_fp64(cfg5, mode):
    _, xe, off, m, c, p, ref = cfg5
    with va.fisher_precision(mode):
        e, t = va.vl_fisher_batch(xe, off, m, c, p, 3)
    errs = [rel(e[i], ref[i][0]) for i in range(8)]
    print(f"config 5 mode {mode} vs FP64 oracle: per image " + ", ".join(f"{v:.2e}" for v in errs))
    if mode == va.MODE_F16X3:
        assert max(errs) < 1e-3
    else:
        # 8-bit mantissas: the compensated BF16 products leave ~2^-16 relative
        # logit error, which the square root and normalisation amplify at K=1024
        assert max(errs) < 5e-3
