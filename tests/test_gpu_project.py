"""Pre-encoding projection (apps/recognition/encodeImage.m:91-97): y = P (x - c),
optionally renormalised to unit L2 norm, as a device kernel and as the load
stage of an encode plan.

Checks: against the closed form in float64 (per-row relative L2 <= 2e-6: FP32
products and sums over at most 200 terms; shapes for both kernels — 16-byte
aligned data with 32 <= Din <= 128 a multiple of 16 streamed from global
memory, the rest through LDS chunks — with ragged row counts), unit norms after renormalisation,
the zero row (norm floor 1e-12), a PCA-whitening projection learned as
trainEncoder.m:150-166 learns it whitening its own training data (identity
covariance), and a plan with the load stage giving bit for bit the encodings of
the same plan fed with the projected descriptors."""
import numpy as np
import pytest

import vlfeat_amd as va
from oracle import oracle
from vlfeat_amd import synth
from vlfeat_amd.gate import DeviceArray

pytestmark = pytest.mark.gpu


def rel_rows(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return np.linalg.norm(a - b, axis=1) / np.maximum(np.linalg.norm(b, axis=1), 1e-300)


def project(x, P, c, renorm):
    L = va.lib()
    N, Din = x.shape
    D = P.shape[0]
    dx, dP, dc = DeviceArray(x), DeviceArray(np.ascontiguousarray(P, np.float32)), DeviceArray(c.astype(np.float32))
    dout = DeviceArray(nbytes=4 * max(N, 1) * D)
    L.vlhip_project(dout.ptr, dx.ptr, N, Din, D, dP.ptr, dc.ptr, int(renorm))
    return dout.to_host(np.float32, (N, D))


def closed_form(x, P, c, renorm):
    xc = (x - c).astype(np.float32).astype(np.float64)  # the centring is a float subtraction
    y = xc @ P.astype(np.float64).T
    if renorm:
        y = y / np.maximum(np.linalg.norm(y, axis=1, keepdims=True), 1e-12)
    return y


@pytest.mark.parametrize("N,Din,D", [(1000, 128, 80), (777, 37, 13), (130, 200, 128), (5, 8, 1), (64, 32, 16),
                                     (31, 128, 80), (4097, 64, 128), (3001, 48, 7), (20000, 128, 80), (33, 16, 5)])
@pytest.mark.parametrize("renorm", [0, 1])
def test_project_matches_closed_form(N, Din, D, renorm):
    rng = np.random.default_rng(N + Din + D)
    x = rng.standard_normal((N, Din)).astype(np.float32) * np.float32(0.3) + np.float32(0.1)
    P = rng.standard_normal((D, Din)).astype(np.float32) / np.float32(np.sqrt(Din))
    c = rng.standard_normal(Din).astype(np.float32) * np.float32(0.1)
    y = project(x, P, c, renorm)
    ref = closed_form(x, P, c, renorm)
    err = rel_rows(y, ref)
    assert err.max() < 2e-6, err.max()
    if renorm:
        np.testing.assert_allclose(np.linalg.norm(y.astype(np.float64), axis=1), 1.0, atol=2e-6)


def test_project_zero_row_and_empty():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((70, 24)).astype(np.float32)
    c = x[17].copy()
    P = rng.standard_normal((10, 24)).astype(np.float32)
    y = project(x, P, c, 1)
    assert np.all(y[17] == 0)           # 0 * (1 / max(0, 1e-12))
    assert np.all(np.isfinite(y))
    assert project(x[:0], P, c, 1).shape == (0, 10)


def test_learned_whitening_whitens_training_data():
    """PCA + whitening learned as trainEncoder.m:150-166 (mean, eigenvectors of
    the covariance, rows scaled by 1 / sqrt(eigenvalue)): the projected training
    data have zero mean and identity covariance."""
    x = synth.sift_like(20000, 128, seed=6, project=False)
    mu = x.mean(0, dtype=np.float64)
    xc = x.astype(np.float64) - mu
    d, V = np.linalg.eigh(xc.T @ xc / x.shape[0])
    order = np.argsort(d)[::-1][:64]
    P = (V[:, order] / np.sqrt(d[order])).T.astype(np.float32)
    y = project(x, P, mu.astype(np.float32), 0).astype(np.float64)
    assert np.abs(y.mean(0)).max() < 1e-4
    cov = y.T @ y / y.shape[0] - np.outer(y.mean(0), y.mean(0))
    assert np.abs(cov - np.eye(64)).max() < 1e-3


def test_plan_load_stage_equals_projected_input():
    L = va.lib()
    rng = np.random.default_rng(11)
    counts = np.array([700, 1, 0, 2500, 333, 1200], np.int64)
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    n, N, D, K, Din = len(counts), int(off[-1]), 80, 96, 128
    raw = synth.sift_like(N, 128, seed=12, project=False)
    P = synth.projection(D, Din, seed=12).astype(np.float32)
    c = raw.mean(0).astype(np.float32)
    y = project(raw, P, c, 1)
    m = y[rng.choice(N, K, replace=False)].copy()
    cov = (np.var(y, 0)[None, :] * rng.uniform(0.1, 0.5, (K, D))).astype(np.float32)
    p = rng.dirichlet(np.ones(K)).astype(np.float32)

    draw, dy, dm, dc, dp = (DeviceArray(a) for a in (raw, y, m, cov, p))
    dP, dcen = DeviceArray(P), DeviceArray(c)
    outs = []
    for with_stage in (True, False):
        denc = DeviceArray(nbytes=4 * n * 2 * K * D)
        dvlad = DeviceArray(nbytes=4 * n * K * D)
        dterms = DeviceArray(nbytes=8 * n)
        plan = L.vlhip_encode_plan_new(D, K, off.ctypes.data, n, 3, 1)
        try:
            if with_stage:
                L.vlhip_encode_plan_set_projection(plan, Din, dP.ptr, dcen.ptr, 1)
            for _ in range(2):  # first run: tuning and capture; second: graph replay
                L.vlhip_encode_plan_run(plan, (draw if with_stage else dy).ptr, dm.ptr, dc.ptr, dp.ptr, denc.ptr,
                                        dvlad.ptr, dterms.ptr)
            outs.append((denc.to_host(np.float32, (n, 2 * K * D)), dvlad.to_host(np.float32, (n, K * D)),
                         dterms.to_host(np.uint64, (n,))))
        finally:
            L.vlhip_encode_plan_delete(plan)
    assert np.array_equal(outs[0][0], outs[1][0])
    assert np.array_equal(outs[0][1], outs[1][1])
    assert np.array_equal(outs[0][2], outs[1][2])
    yo = closed_form(raw, P, c, 1).astype(np.float32)
    for i in (0, 3, 4):
        eo, _ = oracle().fisher(yo[off[i]:off[i + 1]], m, cov, p, 3)
        assert rel_rows(outs[0][0][i:i + 1], eo[None]).max() < 2e-4
