"""Pins of the FP64 oracle (CPU only).

Each oracle routine is checked against something other than its own formula:
the closed forms and known answers of the reference's MATLAB tests
(toolbox/xtest/vl_test_fisher.m, vl_test_vlad.m), an independent scipy
restatement of the mixture log-likelihood, brute force, invariants, and --
where it was compiled here -- the reference library's own double path.
"""
import numpy as np
import pytest
from scipy.special import logsumexp

from oracle import oracle, reference


def simple_fisher(x, mu, sigma2, prior, fast=False):
    """vl_test_fisher.m:57-77 in numpy (columns = data, as in MATLAB)."""
    sigma = np.sqrt(sigma2)
    K = mu.shape[1]
    delta = [(x - mu[:, [k]]) / sigma[:, [k]] for k in range(K)]
    q = np.stack([np.log(prior[k]) - 0.5 * np.sum(np.log(sigma2[:, k])) - 0.5 * np.sum(delta[k] ** 2, 0)
                  for k in range(K)])
    q = np.exp(q - q.max(0))
    q = q / q.sum(0)
    n = x.shape[1]
    if fast:
        i = q.argmax(0)
        q = np.zeros_like(q)
        q[i, np.arange(n)] = 1
    u = [delta[k] @ q[k] / n / np.sqrt(prior[k]) for k in range(K)]
    v = [(delta[k] ** 2 - 1) @ q[k] / n / np.sqrt(2 * prior[k]) for k in range(K)]
    return np.concatenate(u + v)


@pytest.fixture(scope="module")
def fisher_case():
    rng = np.random.default_rng(0)
    D, N, K = 5, 21, 3
    x = rng.standard_normal((D, N))
    mu = rng.standard_normal((D, K))
    return x, mu, np.ones((D, K)), np.ones(K) / K


@pytest.mark.parametrize("variant", ["basic", "norm", "sqrt", "improved", "fast"])
def test_oracle_fisher_closed_form(fisher_case, variant):
    x, mu, s2, pri = fisher_case
    phi = simple_fisher(x, mu, s2, pri, fast=(variant == "fast"))
    if variant in ("sqrt", "improved", "fast"):
        phi = np.sign(phi) * np.sqrt(np.abs(phi))
    if variant in ("norm", "improved", "fast"):
        phi = phi / np.linalg.norm(phi)
    flags = {"basic": 0, "norm": 2, "sqrt": 1, "improved": 3, "fast": 7}[variant]
    enc, _ = oracle().fisher(x.T, mu.T, s2.T, pri, flags)
    np.testing.assert_allclose(enc, phi, atol=1e-10, rtol=0)


def test_oracle_vlad_known_answers():
    o = oracle()
    x = np.array([[1.0], [2.0], [3.0]])
    eye = np.eye(3)
    assert np.array_equal(o.vlad(x, np.zeros((3, 1)), eye, 4), [1, 2, 3])
    assert np.array_equal(o.vlad(x, np.array([[0.0], [1.0], [2.0]]), eye, 4), [1, 1, 1])
    xx = np.concatenate([x, x])
    aa = np.concatenate([eye, eye])
    assert np.array_equal(o.vlad(xx, np.array([[0.0], [1.0], [2.0]]), aa, 4), [2, 2, 2])


@pytest.mark.parametrize("variant", ["rand", "norm", "sqrt", "individual", "mass"])
def test_oracle_vlad_closed_form(variant):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((128, 256))
    mu = rng.standard_normal((128, 16))
    a = rng.uniform(size=(16, 256))
    a = a / a.sum(0)
    phi = np.concatenate([x @ a[i] - a[i].sum() * mu[:, i] for i in range(16)])  # vl_test_vlad.m:68-73
    flags = 4
    if variant == "norm":
        phi, flags = phi / np.linalg.norm(phi), 0
    elif variant == "sqrt":
        phi = np.sign(phi) * np.sqrt(np.abs(phi))
        phi, flags = phi / np.linalg.norm(phi), 2
    elif variant == "individual":
        p = phi.reshape(16, 128)
        phi, flags = (p / np.sqrt((p ** 2).sum(1, keepdims=True))).ravel(), 4 | 1
    elif variant == "mass":
        p = phi.reshape(16, 128)
        phi, flags = (p / a.sum(1)[:, None]).ravel(), 4 | 8
    enc = oracle().vlad(x.T, mu.T, a.T, flags)
    np.testing.assert_allclose(enc, phi, atol=1e-10, rtol=0)


def test_oracle_posteriors_match_mixture_density():
    rng = np.random.default_rng(3)
    N, D, K = 200, 7, 5
    x = rng.standard_normal((N, D))
    mu = rng.standard_normal((K, D))
    cov = rng.uniform(0.3, 2.0, (K, D))
    pri = rng.dirichlet(np.ones(K))
    # independent restatement: log N(x; mu, diag cov) summed with log priors
    lp = (np.log(pri)[None] - 0.5 * np.log(2 * np.pi * cov).sum(1)[None]
          - 0.5 * (((x[:, None, :] - mu[None]) ** 2) / cov[None]).sum(2))
    q, ll = oracle().posteriors(x, mu, cov, pri)
    np.testing.assert_allclose(ll, logsumexp(lp, axis=1).sum(), rtol=1e-12)
    np.testing.assert_allclose(q, np.exp(lp - logsumexp(lp, axis=1, keepdims=True)), atol=1e-13)
    np.testing.assert_allclose(q.sum(1), 1, atol=1e-13)


def test_oracle_posteriors_tiny_prior_is_excluded():
    x = np.zeros((3, 2))
    mu = np.zeros((2, 2))
    cov = np.ones((2, 2))
    q, _ = oracle().posteriors(x, mu, cov, np.array([1 - 1e-7, 1e-7]))
    assert np.all(q[:, 1] == 0)  # gmm.c:746-747: prior < 1e-6 -> log weight -inf


def test_oracle_mstep_one_hot_gives_cluster_moments():
    rng = np.random.default_rng(5)
    N, D, K = 300, 4, 3
    x = rng.standard_normal((N, D)) * 2 + 1
    a = rng.integers(0, K, N)
    q = np.eye(K)[a]
    m, c, p = oracle().maximization(q, x, np.zeros((K, D)), np.zeros((K, D)), np.zeros(K), 1e-4)
    for k in range(K):
        xs = x[a == k]
        np.testing.assert_allclose(m[k], xs.mean(0), rtol=1e-12)
        np.testing.assert_allclose(c[k], xs.var(0), rtol=1e-10)
        assert abs(p[k] - len(xs) / N) < 1e-14


def test_oracle_quantize_brute_force():
    rng = np.random.default_rng(6)
    x = rng.standard_normal((500, 9))
    c = rng.standard_normal((17, 9))
    a, d = oracle().quantize(x, c)
    dd = ((x[:, None] - c[None]) ** 2).sum(2)
    assert np.array_equal(a, dd.argmin(1))
    np.testing.assert_allclose(d, dd.min(1), rtol=1e-12)


# ---- against the reference library (double path) ---------------------------

needs_ref = pytest.mark.skipif(reference() is None, reason="reference library not built here")


@needs_ref
def test_oracle_matches_reference_double_posteriors_and_fisher():
    ref = reference()
    rng = np.random.default_rng(7)
    N, D, K = 400, 16, 8
    x = rng.standard_normal((N, D))
    mu = rng.standard_normal((K, D))
    cov = rng.uniform(0.5, 2, (K, D))
    pri = rng.dirichlet(np.ones(K))
    q0, ll0 = ref.gmm_posteriors(x, mu, cov, pri)
    q1, ll1 = oracle().posteriors(x, mu, cov, pri)
    np.testing.assert_allclose(q1, q0, atol=1e-12)
    np.testing.assert_allclose(ll1, ll0, rtol=1e-12)
    for flags in (0, 1, 2, 3, 7):
        e0, n0 = ref.fisher(x, mu, cov, pri, flags)
        e1, n1 = oracle().fisher(x, mu, cov, pri, flags)
        np.testing.assert_allclose(e1, e0, atol=1e-11)
        assert n0 == n1


@needs_ref
def test_oracle_matches_reference_double_em_with_restart():
    ref = reference()
    rng = np.random.default_rng(8)
    N, D, K = 600, 6, 6
    x = np.concatenate([rng.standard_normal((N // 2, D)), rng.standard_normal((N // 2, D)) + 4])
    mu = rng.standard_normal((K, D))
    mu[5] = 50.0  # far from all data: forces the empty-mode restart path
    cov = np.ones((K, D))
    pri = np.ones(K) / K
    m0, c0, p0, ll0, q0 = ref.gmm(x, K, init="custom", means=mu, covs=cov, priors=pri, max_iter=12)
    m1, c1, p1, ll1, q1 = oracle().em(x, mu, cov, pri, 12)
    np.testing.assert_allclose(m1, m0, atol=1e-8)
    np.testing.assert_allclose(c1, c0, atol=1e-8)
    np.testing.assert_allclose(p1, p0, atol=1e-10)
    np.testing.assert_allclose(ll1, ll0, rtol=1e-10)


@needs_ref
def test_oracle_matches_reference_vlad_and_quantize():
    ref = reference()
    rng = np.random.default_rng(9)
    x = rng.standard_normal((300, 12))
    mu = rng.standard_normal((10, 12))
    a = rng.uniform(size=(300, 10))
    for flags in (0, 1, 2, 4, 8, 15):
        np.testing.assert_allclose(oracle().vlad(x, mu, a, flags), ref.vlad(x, mu, a, flags), atol=1e-12)
    km = ref.kmeans_new(np.float64)
    try:
        ref.vl_kmeans_set_centers(km, mu.ctypes.data, 12, 10)
        a0, d0 = ref.kmeans_quantize(km, x)
    finally:
        ref.vl_kmeans_delete(km)
    a1, d1 = oracle().quantize(x, mu)
    assert np.array_equal(a0, a1)
    np.testing.assert_allclose(d1, d0, rtol=1e-13)


def test_planted_sign_error_is_caught(fisher_case):
    """A plausible mistake (v without the -1 term) must fail the closed form."""
    x, mu, s2, pri = fisher_case
    enc, _ = oracle().fisher(x.T, mu.T, s2.T, pri, 0)
    K, D = mu.shape[1], mu.shape[0]
    bad = enc.copy()
    q, _ = oracle().posteriors(x.T, mu.T, s2.T, pri)
    bad[K * D:] += (q.sum(0) / x.shape[1] / np.sqrt(2 * pri)).repeat(D)
    assert not np.allclose(bad, simple_fisher(x, mu, s2, pri), atol=1e-6)


# ---- learned-gate perceptron (or_mlp_loss_grads) -------------------------------

def _mlp_case(D=5, H1=7, H2=6, K=9, rows=4, seed=0):
    rng = np.random.default_rng(seed)
    n = D * H1 + H1 + H1 * H2 + H2 + H2 * K + K
    params = rng.normal(0, 0.5, n)
    xin = rng.normal(0, 1, (rows, D))
    q = rng.dirichlet(np.ones(K), rows)
    return params, xin, q, (D, H1, H2, K)


def test_mlp_oracle_matches_finite_differences():
    params, xin, q, (D, H1, H2, K) = _mlp_case()
    O = oracle()
    loss, g = O.mlp_loss_grads(params, D, H1, H2, K, xin, q)
    eps = 1e-6
    fd = np.empty_like(params)
    for i in range(params.size):
        p1 = params.copy(); p1[i] += eps
        p2 = params.copy(); p2[i] -= eps
        fd[i] = (O.mlp_loss_grads(p1, D, H1, H2, K, xin, q, False)[0]
                 - O.mlp_loss_grads(p2, D, H1, H2, K, xin, q, False)[0]) / (2 * eps)
    assert np.linalg.norm(g - fd) / np.linalg.norm(fd) < 1e-7
    # a plausible slip (GELU derivative without the a*phi(a) term) would not pass
    assert np.abs(g - fd).max() < 1e-7


def test_mlp_oracle_softmax_identities():
    D, H1, H2, K, rows = 4, 3, 5, 8, 6
    n = D * H1 + H1 + H1 * H2 + H2 + H2 * K + K
    rng = np.random.default_rng(1)
    xin = rng.normal(0, 1, (rows, D))
    q = rng.dirichlet(np.ones(K), rows)
    O = oracle()
    # zero network: uniform prediction, loss = log K, dL/db3 = 1/K - mean q
    loss, g = O.mlp_loss_grads(np.zeros(n), D, H1, H2, K, xin, q)
    assert abs(loss - np.log(K)) < 1e-12
    np.testing.assert_allclose(g[-K:], 1.0 / K - q.mean(0), atol=1e-14)
    # a large output bias on the target class of one-hot rows drives the loss to ~0
    p = np.zeros(n)
    p[-K:] = 0.0
    p[-K + 3] = 40.0
    oh = np.zeros((rows, K)); oh[:, 3] = 1
    loss, _ = O.mlp_loss_grads(p, D, H1, H2, K, xin, oh, False)
    assert loss < 1e-15
