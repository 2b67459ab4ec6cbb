"""GPU parity: the HIP path (through the C ABI) against the reference's own
float/double implementation (compiled from source, when present) and the FP64
oracle, on identical seeded inputs.

Tolerances (stated per test): Fisher / VLAD vectors within 1e-4 relative L2 of
the reference float path; GMM means / variances within 1e-4 relative; hard
assignments bit-exact; double-precision paths within 1e-9 of the FP64 oracle.
"""
import numpy as np
import pytest

import vlfeat_amd as va
from oracle import oracle, reference
from vlfeat_amd import synth

pytestmark = pytest.mark.gpu
needs_ref = pytest.mark.skipif(reference() is None, reason="reference library not built")


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def gmm_params(x, K, seed=0, dtype=np.float32):
    rng = np.random.default_rng(seed)
    idx = rng.choice(x.shape[0], K, replace=False)
    means = x[idx].astype(dtype)
    covs = (np.var(x, 0)[None, :] * rng.uniform(0.1, 0.5, (K, x.shape[1]))).astype(dtype)
    pri = rng.dirichlet(np.ones(K)).astype(dtype)
    return means, covs, pri


@pytest.fixture(scope="module")
def cfg1():
    x = synth.mixture(6000, 64, 64, seed=0)
    return x, *gmm_params(x, 64)


@pytest.fixture(scope="module")
def cfg2():
    x = synth.sift_like(4000, 80, seed=1)
    return x, *gmm_params(x, 128, seed=2)


# ---------------------------------------------------------------- posteriors

@pytest.mark.parametrize("case", ["cfg1", "cfg2"])
def test_posteriors_float(case, request):
    x, m, c, p = request.getfixturevalue(case)
    q, ll = va.lib().gmm_posteriors(x, m, c, p)
    qo, llo = oracle().posteriors(x, m, c, p)
    assert rel(q, qo) < 1e-5
    assert abs(ll - llo) / abs(llo) < 1e-6
    np.testing.assert_allclose(q.sum(1), 1, atol=1e-5)
    R = reference()
    if R is not None:
        qr, llr = R.gmm_posteriors(x, m, c, p)
        assert rel(q, qr) < 1e-4
        assert abs(ll - llr) / abs(llr) < 1e-6


def test_posteriors_double(cfg1):
    x, m, c, p = [a.astype(np.float64) for a in cfg1]
    q, ll = va.lib().gmm_posteriors(x, m, c, p)
    qo, llo = oracle().posteriors(x, m, c, p)
    assert rel(q, qo) < 1e-12
    assert abs(ll - llo) / abs(llo) < 1e-12


def test_posteriors_tiny_prior_and_single_component():
    x = synth.mixture(500, 16, 4, seed=3)
    m, c, p = gmm_params(x, 3)
    p = np.array([1 - 5e-7, 5e-7, 0], np.float32)
    q, _ = va.lib().gmm_posteriors(x, m, c, p)
    assert np.all(q[:, 1:] == 0)
    q1, ll1 = va.lib().gmm_posteriors(x, m[:1], c[:1], np.ones(1, np.float32))
    assert np.all(q1 == 1)
    _, llo = oracle().posteriors(x, m[:1], c[:1], np.ones(1))
    assert abs(ll1 - llo) / abs(llo) < 1e-6


# ---------------------------------------------------------------- EM

def test_em_custom_init_float_vs_oracle_and_reference(cfg1):
    x, m, c, p = cfg1
    out = va.vl_gmm(x, 64, init="custom", means=m, covs=c, priors=p, max_iter=15)
    mo, co, po, llo, _ = oracle().em(x, m, c, p, 15)
    assert rel(out[0], mo) < 1e-4 and rel(out[1], co) < 1e-4 and rel(out[2], po) < 1e-4
    assert abs(out[3] - llo) / abs(llo) < 1e-5
    R = reference()
    if R is not None:
        R.set_num_threads(1)
        mr, cr, pr, llr, qr = R.gmm(x, 64, init="custom", means=m, covs=c, priors=p, max_iter=15)
        assert rel(out[0], mr) < 1e-4 and rel(out[1], cr) < 1e-4 and rel(out[2], pr) < 1e-4
        assert rel(out[4], qr) < 1e-3


def test_em_double_with_restart_vs_oracle():
    rng = np.random.default_rng(8)
    N, D, K = 3000, 12, 8
    x = np.concatenate([rng.standard_normal((N // 2, D)), rng.standard_normal((N // 2, D)) + 4])
    m = rng.standard_normal((K, D))
    m[5] = 60.0  # far from all data: the restart path must split a mode
    c = np.ones((K, D))
    p = np.ones(K) / K
    out = va.vl_gmm(x, K, init="custom", means=m, covs=c, priors=p, max_iter=10)
    mo, co, po, llo, qo = oracle().em(x, m, c, p, 10)
    assert rel(out[0], mo) < 1e-9 and rel(out[1], co) < 1e-9 and rel(out[2], po) < 1e-9
    assert abs(out[3] - llo) / abs(llo) < 1e-10
    assert rel(out[4], qo) < 1e-9


@needs_ref
@pytest.mark.parametrize("init", ["kmeans", "rand"])
def test_gmm_cluster_seeded_init_matches_reference(init):
    """Full vl_gmm_cluster with the reference's own initialisers (k-means via
    ANN kd-forest, or k-means++) from the same generator seed."""
    x = synth.mixture(20000, 32, 32, seed=5)
    R = reference()
    R.set_num_threads(1)
    res = []
    for lib in (va.lib(), R):
        lib.seed(49000)
        res.append(lib.gmm(x, 32, init=init, max_iter=20, cov_bound=1e-4))
    (m, c, p, ll, q), (mr, cr, pr, llr, qr) = res
    assert rel(m, mr) < 1e-4 and rel(c, cr) < 1e-4 and rel(p, pr) < 1e-4
    assert abs(ll - llr) / abs(llr) < 1e-5


# ---------------------------------------------------------------- Fisher

FLAGS = {"plain": 0, "sqrt": 1, "norm": 2, "improved": 3, "fast": 7}


@pytest.mark.parametrize("case", ["cfg1", "cfg2"])
@pytest.mark.parametrize("flags", list(FLAGS.values()), ids=list(FLAGS))
def test_fisher_float(case, flags, request):
    x, m, c, p = request.getfixturevalue(case)
    xs = x[:2000]
    enc, n = va.vl_fisher(xs, m, c, p, flags)
    eo, no = oracle().fisher(xs, m, c, p, flags)
    assert rel(enc, eo) < 1e-4
    R = reference()
    if R is not None:
        er, nr = R.fisher(xs, m, c, p, flags)
        assert rel(enc, er) < 1e-4
        assert abs(int(n) - int(nr)) <= max(2, nr // 10000)


@pytest.mark.parametrize("flags", [0, 3, 7])
def test_fisher_double(cfg1, flags):
    x, m, c, p = [a.astype(np.float64) for a in cfg1]
    enc, n = va.vl_fisher(x[:1500], m, c, p, flags)
    eo, no = oracle().fisher(x[:1500], m, c, p, flags)
    assert rel(enc, eo) < 1e-10
    assert n == no


def test_fisher_batch_uneven_equals_per_image(cfg2):
    x, m, c, p = cfg2
    counts = [0, 1, 700, 64, 65, 1, 1300, 3, 0, 257, 129]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    xs = x[: int(off[-1])]
    enc, terms = va.vl_fisher_batch(xs, off, m, c, p, 3)
    for i in range(len(counts)):
        e1, n1 = va.vl_fisher(xs[off[i]:off[i + 1]], m, c, p, 3)
        assert np.array_equal(enc[i], e1), i
        assert terms[i] == n1
    assert np.all(enc[0] == 0) and np.all(enc[8] == 0)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_fisher_segments_in_parallel_equal_sequential(cfg2, dt):
    # a batch large enough to fill the device sums each image's segments in
    # one workgroup, one image alone sums them in parallel workgroups: the
    # encodings must agree bit for bit (and stay within the oracle's bar)
    x, m, c, p = (a.astype(dt) for a in cfg2)
    K = m.shape[0]
    rng = np.random.default_rng(3)
    counts = rng.integers(1, 1300, size=300)
    counts[[0, 7]] = [1299, 257]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    xs = np.concatenate([x] * int(np.ceil(off[-1] / x.shape[0])))[: int(off[-1])]
    enc, terms = va.vl_fisher_batch(xs, off, m, c, p, 3)
    for i in (0, 7, 42):
        e1, n1 = va.vl_fisher(xs[off[i]:off[i + 1]], m, c, p, 3)
        assert np.array_equal(enc[i], e1), i
        assert terms[i] == n1
    eo, _ = oracle().fisher(xs[off[0]:off[1]], m, c, p, 3)
    assert rel(enc[0], eo) < (1e-4 if dt == np.float32 else 1e-10)


def test_fisher_skips_tiny_prior_component(cfg1):
    x, m, c, p = cfg1
    p = p.copy()
    p[3] = 1e-7
    p /= p.sum()
    enc, _ = va.vl_fisher(x[:500], m, c, p, 0)
    K, D = m.shape
    assert np.all(enc[3 * D:4 * D] == 0) and np.all(enc[K * D + 3 * D:K * D + 4 * D] == 0)
    eo, _ = oracle().fisher(x[:500], m, c, p, 0)
    assert rel(enc, eo) < 1e-4


# ---------------------------------------------------------------- VLAD

@pytest.mark.parametrize("dt", [np.float32, np.float64])
@pytest.mark.parametrize("flags", [0, 1, 2, 8, 15])
def test_vlad_dense_onehot_equals_hard(dt, flags):
    # a dense matrix of one-hot rows (and empty rows) is encoded from its
    # cluster indices: the same numbers as the dense computation
    rng = np.random.default_rng(5)
    x = rng.standard_normal((700, 96)).astype(dt)
    mu = rng.standard_normal((40, 96)).astype(dt)
    a = rng.integers(0, 40, 700).astype(np.uint32)
    a[::97] = 0xFFFFFFFF
    dense = np.zeros((700, 40), dt)
    rows = np.nonzero(a != 0xFFFFFFFF)[0]
    dense[rows, a[rows]] = 1
    enc = va.vl_vlad(x, mu, dense, flags)
    assert np.array_equal(enc, va.vl_vlad_hard(x, mu, a, flags))
    assert rel(enc, oracle().vlad(x, mu, dense, flags)) < (1e-5 if dt == np.float32 else 1e-12)
    R = reference()
    if R is not None:
        assert np.array_equal(enc, R.vlad(x, mu, dense, flags))
    # one weight other than 1 keeps the dense computation
    dense2 = dense.copy()
    dense2[rows[0], a[rows[0]]] = 0.5
    enc2 = va.vl_vlad(x, mu, dense2, flags)
    assert rel(enc2, oracle().vlad(x, mu, dense2, flags)) < (1e-5 if dt == np.float32 else 1e-12)
    if R is not None:
        assert np.array_equal(enc2, R.vlad(x, mu, dense2, flags))


@pytest.mark.parametrize("flags", [0, 1, 2, 4, 8, 15, 4 | 8])
def test_vlad_soft_float(flags):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((256, 128)).astype(np.float32)
    mu = rng.standard_normal((16, 128)).astype(np.float32)
    a = rng.uniform(size=(256, 16)).astype(np.float32)
    a /= a.sum(1, keepdims=True)
    enc = va.vl_vlad(x, mu, a, flags)
    assert rel(enc, oracle().vlad(x, mu, a, flags)) < 1e-5
    R = reference()
    if R is not None:
        assert np.array_equal(enc, R.vlad(x, mu, a, flags))  # same arithmetic, same order


def test_vlad_known_answers():
    x = np.array([[1.0], [2.0], [3.0]], np.float32)
    eye = np.eye(3, dtype=np.float32)
    assert np.array_equal(va.vl_vlad(x, np.zeros((3, 1), np.float32), eye, 4), [1, 2, 3])
    mu = np.array([[0.0], [1.0], [2.0]], np.float32)
    assert np.array_equal(va.vl_vlad(x, mu, eye, 4), [1, 1, 1])
    assert np.array_equal(va.vl_vlad(np.concatenate([x, x]), mu, np.concatenate([eye, eye]), 4), [2, 2, 2])


@pytest.mark.parametrize("n,D,K,dt", [(5000, 80, 256, np.float32), (4500, 13, 37, np.float32),
                                      (2049, 128, 300, np.float64), (6200, 96, 1024, np.float32),
                                      (1, 7, 3, np.float32)])
def test_vlad_hard_chunks_blocks_and_norm_buffers(n, D, K, dt):
    # rows beyond one staged chunk of cluster indices (2048), cluster blocks
    # of 4..32 (chosen from the number of workgroups), encodings longer than
    # one LDS buffer of the global norm with a ragged tail (K*D not a multiple
    # of the 16-byte reads): bitwise the reference's dense loop
    rng = np.random.default_rng(n + D + K)
    x = rng.standard_normal((n, D)).astype(dt)
    mu = rng.standard_normal((K, D)).astype(dt)
    a = rng.integers(0, K, n).astype(np.uint32)
    a[::53] = 0xFFFFFFFF
    dense = np.zeros((n, K), dt)
    rows = np.nonzero(a != 0xFFFFFFFF)[0]
    dense[rows, a[rows]] = 1
    R = reference()
    for flags in (0, 1, 3, 8, 15):
        enc = va.vl_vlad_hard(x, mu, a, flags)
        if R is not None:
            assert np.array_equal(enc, R.vlad(x, mu, dense, flags)), flags
        # float sums of a few rows minus mass * mean cancel: 5e-5 from FP64
        assert rel(enc, oracle().vlad(x, mu, dense, flags)) < (5e-5 if dt == np.float32 else 1e-12)


@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_vlad_hard_wide_descriptors(dt):
    # cluster indices with D > 128 take the dense-assignment kernel (one
    # (cluster, dimension) chain per thread, 16-dimension slices, the last
    # one partial) and the separate per-cluster normalisation
    rng = np.random.default_rng(200)
    n, D, K = 2500, 200, 37
    x = rng.standard_normal((n, D)).astype(dt)
    mu = rng.standard_normal((K, D)).astype(dt)
    a = rng.integers(0, K, n).astype(np.uint32)
    a[::41] = 0xFFFFFFFF
    dense = np.zeros((n, K), dt)
    rows = np.nonzero(a != 0xFFFFFFFF)[0]
    dense[rows, a[rows]] = 1
    R = reference()
    for flags in (0, 1, 3, 15):
        enc = va.vl_vlad_hard(x, mu, a, flags)
        if R is not None:
            assert np.array_equal(enc, R.vlad(x, mu, dense, flags)), flags
        assert rel(enc, oracle().vlad(x, mu, dense, flags)) < (5e-5 if dt == np.float32 else 1e-12)


@needs_ref
@pytest.mark.parametrize("dt", [np.float32, np.float64])
def test_vlad_dimension_1024_dense_and_hard(dt):
    # D above the per-cluster norm's LDS stage (512 values): the norm is summed
    # and divided from global memory in the same order
    rng = np.random.default_rng(1024)
    n, D, K = 700, 1024, 9
    x = rng.standard_normal((n, D)).astype(dt)
    mu = rng.standard_normal((K, D)).astype(dt)
    a = rng.integers(0, K, n).astype(np.uint32)
    dense = np.zeros((n, K), dt)
    dense[np.arange(n), a] = 1
    soft = rng.uniform(0, 1, (n, K)).astype(dt)
    soft[soft < 0.6] = 0
    R = reference()
    for flags in (1, 15):
        assert np.array_equal(va.vl_vlad_hard(x, mu, a, flags), R.vlad(x, mu, dense, flags)), flags
        assert np.array_equal(va.vl_vlad(x, mu, dense, flags), R.vlad(x, mu, dense, flags)), flags
        assert np.array_equal(va.vl_vlad(x, mu, soft, flags), R.vlad(x, mu, soft, flags)), flags


def _vlad_batch_case(seed, n_img, N_max, D, K):
    rng = np.random.default_rng(seed)
    counts = rng.integers(1, N_max, n_img)
    counts[1], counts[2] = 0, 1
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    x = rng.standard_normal((int(off[-1]), D)).astype(np.float32)
    mu = rng.standard_normal((K, D)).astype(np.float32)
    return rng, off, x, mu


@pytest.mark.parametrize("K,n_img,D", [(512, 300, 80), (256, 80, 64), (37, 20, 200)])
def test_vlad_batch_hard_blocks_equal_single_images(K, n_img, D):
    # batches large enough for 32- and 16-cluster blocks (LDS accumulators; a
    # single image uses 4-cluster blocks and register chains), and cluster
    # indices with D > 128 (the dense-assignment kernel), with an empty and a
    # one-row image: each image's encoding bit for bit as when encoded alone
    from vlfeat_amd.gate import DeviceArray
    rng, off, x, mu = _vlad_batch_case(K + n_img, n_img, 400, D, K)
    a = rng.integers(0, K, x.shape[0]).astype(np.uint32)
    a[::29] = 0xFFFFFFFF
    L = va.lib()
    dx, dm, doff, da = DeviceArray(x), DeviceArray(mu), DeviceArray(off), DeviceArray(a)
    denc = DeviceArray(nbytes=4 * n_img * K * D)
    for flags in (0, 3):
        L.vlhip_vlad_encode_batch(1, denc.ptr, dm.ptr, D, K, dx.ptr, doff.ptr, n_img, None, da.ptr, flags)
        enc = denc.to_host(np.float32, (n_img, K * D))
        for i in sorted(set(list(range(0, n_img, max(1, n_img // 12))) + [1, 2, n_img - 1])):
            e1 = va.vl_vlad_hard(x[off[i]:off[i + 1]], mu, a[off[i]:off[i + 1]], flags)
            assert np.array_equal(enc[i], e1), (i, flags)


def test_vlad_batch_soft_equals_single_images():
    # dense (soft) assignments over 40 images: the image-major workgroup order
    # of many-image launches gives the same encodings as single-image calls
    from vlfeat_amd.gate import DeviceArray
    K, D, n_img = 48, 80, 40
    rng, off, x, mu = _vlad_batch_case(9, n_img, 300, D, K)
    q = rng.uniform(size=(x.shape[0], K)).astype(np.float32) ** 4
    q /= q.sum(1, keepdims=True)
    L = va.lib()
    dx, dm, doff, dq = DeviceArray(x), DeviceArray(mu), DeviceArray(off), DeviceArray(q)
    denc = DeviceArray(nbytes=4 * n_img * K * D)
    for flags in (0, 1, 15):
        L.vlhip_vlad_encode_batch(1, denc.ptr, dm.ptr, D, K, dx.ptr, doff.ptr, n_img, dq.ptr, None, flags)
        enc = denc.to_host(np.float32, (n_img, K * D))
        for i in (0, 1, 2, 17, n_img - 1):
            e1 = va.vl_vlad(x[off[i]:off[i + 1]], mu, q[off[i]:off[i + 1]], flags)
            assert np.array_equal(enc[i], e1), (i, flags)


def _f32_global_norm(z):
    """vlad.c:230-241 in float32, term by term: n = n + z*z, sqrt, max 1e-12."""
    n = np.float32(0)
    for v in z.astype(np.float32):
        n = np.float32(n + np.float32(v * v))
    n = np.float32(np.sqrt(n))
    n = n if n > 1e-12 else np.float32(1e-12)
    return (z.astype(np.float32) / n).astype(np.float32)


@pytest.mark.parametrize("case", ["ties", "subnormal", "overflow", "nan", "zeros", "mixed"])
def test_vlad_global_norm_exact_sum(case):
    # one descriptor per cluster, zero means, flags 0: the encoding before the
    # global normalisation is the data itself, so the squared norm's float
    # sum sees chosen terms: exact ties (2^-24 on top of 1), subnormal squares
    # and a start below the normal range, squares that overflow to inf, a
    # NaN, long runs of zeros, and sums over many powers of two
    K, D = 96, 64
    rng = np.random.default_rng(len(case))
    z = rng.standard_normal(K * D).astype(np.float32)
    if case == "ties":
        z[:] = np.float32(2.0 ** -12)
        z[::7] = 1.0
        z[::13] = np.float32(3 * 2.0 ** -13)
        z[0] = 1.0
    elif case == "subnormal":
        z[: K * D // 2] = (rng.uniform(0.5, 2.0, K * D // 2) * 1e-21).astype(np.float32)
        z[K * D // 2::5] = np.float32(1e-19)
    elif case == "overflow":
        z[100] = np.float32(1.5e19)
        z[200] = np.float32(1.5e19)
    elif case == "nan":
        z[3000] = np.nan
    elif case == "zeros":
        z[:2000] = 0
        z[5000:] = 0
    else:
        z *= np.exp2(rng.integers(-30, 30, K * D)).astype(np.float32)
    x = z.reshape(K, D)
    mu = np.zeros((K, D), np.float32)
    a = np.arange(K, dtype=np.uint32)
    enc = va.vl_vlad_hard(x, mu, a, 0)
    np.testing.assert_array_equal(enc, _f32_global_norm(z))
    R = reference()
    if R is not None:
        np.testing.assert_array_equal(enc, R.vlad(x, mu, np.eye(K, dtype=np.float32), 0))


def test_vlad_hard_equals_one_hot_and_double(cfg2):
    x, m, c, p = cfg2
    a, _ = va.vl_kmeans_quantize(x, m)
    onehot = np.eye(m.shape[0], dtype=np.float32)[a]
    for flags in (0, 3, 8):
        assert np.array_equal(va.vl_vlad_hard(x, m, a, flags), va.vl_vlad(x, m, onehot, flags))
    xd, md = x.astype(np.float64), m.astype(np.float64)
    assert rel(va.vl_vlad(xd, md, onehot.astype(np.float64), 1), oracle().vlad(xd, md, onehot, 1)) < 1e-12


# ---------------------------------------------------------------- k-means

@pytest.mark.parametrize("dtype,D", [(np.float32, 80), (np.float32, 13), (np.float64, 64)])
def test_quantize_bit_exact(dtype, D):
    rng = np.random.default_rng(D)
    x = rng.standard_normal((20000, D)).astype(dtype)
    cen = rng.standard_normal((300, D)).astype(dtype)
    a, d = va.vl_kmeans_quantize(x, cen)
    ao, do = oracle().quantize(x, cen)
    assert np.mean(a == ao) > 0.999
    R = reference()
    if R is not None:
        km = R.kmeans_new(dtype)
        R.vl_kmeans_set_centers(km, cen.ctypes.data, D, 300)
        ar, dr = R.kmeans_quantize(km, x)
        R.vl_kmeans_delete(km)
        assert np.array_equal(a, ar)
        assert np.array_equal(d, dr)


@needs_ref
@pytest.mark.parametrize("D,K,scale", [(128, 1024, 1.0), (37, 100, 1.0), (80, 512, 1e3), (64, 256, 1e-3),
                                       (16, 64, 1.0)])
def test_quantize_screened_ties_and_scales(D, K, scale):
    # float data with D <= 128 and K >= 64 takes the matrix-core screen; the
    # assignments and distances must still be the reference's bit for bit,
    # including exact ties (duplicate centres, rows equal to a centre) where
    # the first index wins
    rng = np.random.default_rng(D + K)
    x = (synth.sift_like(6000, D, seed=D, project=D != 128) if D in (64, 80, 128)
         else rng.standard_normal((6000, D))).astype(np.float32) * np.float32(scale)
    cen = x[rng.choice(x.shape[0], K, replace=False)].copy()
    cen[7] = cen[3]
    cen[K - 1] = cen[3]
    x[:50] = cen[3]                       # distance 0 to centres 3, 7 and K-1
    x[50:60] = cen[5] + np.float32(scale * 1e-4)
    a, d = va.vl_kmeans_quantize(x, cen)
    assert np.all(a[:50] == 3)
    R = reference()
    km = R.kmeans_new(np.float32)
    R.vl_kmeans_set_centers(km, cen.ctypes.data, D, K)
    ar, dr = R.kmeans_quantize(km, x)
    R.vl_kmeans_delete(km)
    assert np.array_equal(a, ar)
    assert np.array_equal(d, dr)


@needs_ref
@pytest.mark.parametrize("K", [512, 1030])
def test_quantize_screened_crowded_windows(K):
    # the screen keeps at most two centres per lane within 2 eps of the lane's
    # running minimum and sends the row to the full scan beyond that; here
    # rows sit at the centre of 2D centres x0 +- 0.5 e_d, all at exactly the
    # same distance 0.25 (every lane overflows; the first index must win), or
    # just off it (near ties inside the window), or sweep past a decreasing
    # run of centres (the window restarts when the minimum drops by more than
    # 2 eps)
    D = 80
    rng = np.random.default_rng(K)
    cen = (rng.standard_normal((K, D)) * 4).astype(np.float32)
    x0 = np.round(rng.standard_normal(D)).astype(np.float32)
    ring = np.concatenate([x0 + 0.5 * np.eye(D, dtype=np.float32), x0 - 0.5 * np.eye(D, dtype=np.float32)])
    cen[100:100 + 2 * D] = ring[rng.permutation(2 * D)]
    # a run of centres approaching x1, one per lane position, farther first
    x1 = (rng.standard_normal(D) * 2).astype(np.float32)
    u = rng.standard_normal(D).astype(np.float32)
    u /= np.linalg.norm(u)
    for j, k in enumerate(range(300, 300 + 16 * 12, 16)):
        cen[k] = x1 + np.float32(3.0 * 0.5 ** j) * u
    x = (rng.standard_normal((4000, D)) * 4).astype(np.float32)
    x[:40] = x0
    x[40:80] = x0 + (rng.standard_normal((40, D)) * 1e-4).astype(np.float32)
    x[80:120] = x1
    x[120:160] = x1 + (rng.standard_normal((40, D)) * 1e-3).astype(np.float32)
    # rows without a finite distance take the full scan: distance +inf, the
    # first centre (the reference leaves their assignment undefined)
    x[160:163] = np.nan
    x[163, 7] = np.nan
    a, d = va.vl_kmeans_quantize(x, cen)
    assert np.all(a[160:164] == 0) and np.all(d[160:164] == np.inf)
    R = reference()
    km = R.kmeans_new(np.float32)
    R.vl_kmeans_set_centers(km, cen.ctypes.data, D, K)
    ar, dr = R.kmeans_quantize(km, x)
    R.vl_kmeans_delete(km)
    # all 2D ring centres tie at exactly 0.25: the first of them, 100, wins
    assert np.all(dr[:40] == np.float32(0.25)) and np.all(ar[:40] == 100)
    fin = np.ones(len(x), bool)
    fin[160:164] = False
    assert np.array_equal(a[fin], ar[fin])
    assert np.array_equal(d, dr)


@needs_ref
@pytest.mark.parametrize("algo,comparisons", [(va.VlKMeansLloyd, 100), (va.VlKMeansANN, 20),
                                              (va.VlKMeansANN, 5000)])
def test_kmeans_cluster_matches_reference(algo, comparisons):
    x = synth.mixture(5000, 16, 20, seed=9)
    R = reference()
    R.set_num_threads(1)
    res = []
    for lib in (va.lib(), R):
        lib.seed(1)
        km = lib.kmeans_new(np.float32, algorithm=algo, max_iter=10, num_trees=2, max_comparisons=comparisons)
        e = lib.vl_kmeans_cluster(km, x.ctypes.data, 16, 5000, 40)
        res.append((lib.kmeans_centers(km, np.float32), e))
        lib.vl_kmeans_delete(km)
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1]


@needs_ref
@pytest.mark.parametrize("dtype,D,algo", [(np.float32, 13, va.VlKMeansLloyd), (np.float32, 24, va.VlKMeansANN),
                                          (np.float64, 13, va.VlKMeansLloyd), (np.float64, 16, va.VlKMeansANN)])
def test_kmeans_refine_with_empty_clusters_matches_reference(dtype, D, algo):
    """Refinement from given centers, a quarter of them far from every datum:
    those clusters are empty after the first assignment and restart from data
    drawn by the seeded generator (kmeans.c:919-925 / 1095-1101); centers and
    energy must equal the reference's bit for bit (Lloyd and an exhaustive ANN
    search, which also applies the keep-previous rule of kmeans.c:750-763)."""
    N, K = 30000, 48
    x = synth.mixture(N, D, 12, seed=21, dtype=dtype)
    rng = np.random.default_rng(3)
    cen0 = x[rng.choice(N, K, replace=False)].copy()
    cen0[::4] += dtype(50.0)
    R = reference()
    R.set_num_threads(1)
    res = []
    for lib in (va.lib(), R):
        lib.seed(11)
        km = lib.kmeans_new(dtype, algorithm=algo, max_iter=15, num_trees=2, max_comparisons=0)
        lib.vl_kmeans_set_centers(km, cen0.ctypes.data, D, K)
        e = lib.vl_kmeans_refine_centers(km, x.ctypes.data, N)
        res.append((lib.kmeans_centers(km, dtype), e))
        lib.vl_kmeans_delete(km)
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1]


@needs_ref
def test_kmeans_cluster_large_clusters_matches_reference():
    """Lloyd on 200,000 descriptors and 16 clusters (12,500 members per cluster
    on average, the unrolled part of the serial per-cluster sums): bit-exact
    centers and energy."""
    N, D, K = 200_000, 80, 16
    x = synth.sift_like(N, D, seed=5)
    R = reference()
    R.set_num_threads(0)
    res = []
    for lib in (va.lib(), R):
        lib.seed(2)
        km = lib.kmeans_new(np.float32, algorithm=va.VlKMeansLloyd, max_iter=8)
        e = lib.vl_kmeans_cluster(km, x.ctypes.data, D, N, K)
        res.append((lib.kmeans_centers(km, np.float32), e))
        lib.vl_kmeans_delete(km)
    assert np.array_equal(res[0][0], res[1][0])
    assert res[0][1] == res[1][1]
