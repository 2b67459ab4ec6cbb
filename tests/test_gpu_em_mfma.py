"""EM on the matrix cores (vl_gmm_set_precision with BF16X3 / F16X3): one EM
step's log-likelihood and M-step moments against the FP64 oracle's posteriors,
whole EM runs against the oracle's EM, the restart path, and the posteriors
handed out after a matrix-core EM (the FP32 E-step of the final parameters).

Tolerances (relative L2 unless said otherwise), F16X3 / BF16X3:
  log-likelihood of one step         1e-6 / 1e-5 (relative)
  moments S0, mass, S1, S2 of a step 2e-5 / 2e-4
  means / covariances / priors after 10 EM iterations  2e-4 / 2e-3
The FP32 parity path is printed beside them for scale."""
import numpy as np
import pytest

import vlfeat_amd as va
from oracle import oracle
from vlfeat_amd import synth
from vlfeat_amd.gate import DeviceArray

pytestmark = pytest.mark.gpu

F16, BF16 = 3, 1  # VL_GMM_PRECISION_F16X3, VL_GMM_PRECISION_BF16X3
STEP_LL = {F16: 1e-6, BF16: 1e-5}
STEP_MOM = {F16: 2e-5, BF16: 2e-4}
EM_PARAM = {F16: 2e-4, BF16: 2e-3}


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def start(x, K, seed):
    rng = np.random.default_rng(seed)
    m = x[rng.choice(x.shape[0], K, replace=False)].astype(np.float32)
    c = (np.var(x, 0)[None, :] * rng.uniform(0.2, 0.6, (K, x.shape[1]))).astype(np.float32)
    p = rng.dirichlet(np.ones(K) * 4).astype(np.float32)
    return m, c, p


CASES = {
    "sift128_k64": (lambda: synth.sift_like(20000, 128, seed=6, project=False), 64),
    "sift80_k256": (lambda: synth.sift_like(30000, 80, seed=1), 256),
    "mix64_k64": (lambda: synth.mixture(30000, 64, 64, seed=0), 64),
    "mix24_k37": (lambda: synth.mixture(9000, 24, 10, seed=3), 37),
    "mix48_k160": (lambda: synth.mixture(20000, 48, 32, seed=4), 160),
    "mix112_k90": (lambda: synth.mixture(9000, 112, 12, seed=8), 90),
}


def fp64_moments(q, x, m, K):
    """M-step moments of gmm.c:1093-1149 from FP64 posteriors (threshold 0.01 / K)."""
    x = x.astype(np.float64)
    m = m.astype(np.float64)
    qt = np.where(q < 0.01 / K, 0.0, q)
    S0 = qt.sum(0)
    S1 = qt.T @ x
    S2 = qt.T @ (x * x) - 2 * m * S1 + m * m * S0[:, None]
    return S0, q.sum(0), S1, S2


@pytest.mark.parametrize("mode", [F16, BF16])
@pytest.mark.parametrize("case", list(CASES))
def test_em_step_against_oracle(case, mode):
    x, K = CASES[case][0](), CASES[case][1]
    N, D = x.shape
    m, c, p = start(x, K, seed=11)
    L = va.lib()
    dx, dm, dc, dp = (DeviceArray(a) for a in (x, m, c, p))
    out = np.zeros(2 * K + 2 * K * D)
    ll = L.vlhip_gmm_em_step(mode, out.ctypes.data, dp.ptr, dm.ptr, dc.ptr, D, K, dx.ptr, N)
    ll_only = L.vlhip_gmm_em_step(mode, None, dp.ptr, dm.ptr, dc.ptr, D, K, dx.ptr, N)
    assert ll_only == ll
    q, llo = oracle().posteriors(x, m, c, p)
    S0, mass, S1, S2 = fp64_moments(q, x, m, K)
    got = (out[:K], out[K:2 * K], out[2 * K:2 * K + K * D].reshape(K, D), out[2 * K + K * D:].reshape(K, D))
    errs = [rel(g, e) for g, e in zip(got, (S0, mass, S1, S2))]
    # the FP32 parity path for scale
    qf, llf = L.gmm_posteriors(x, m, c, p)
    print(f"{case} mode {mode}: LL {abs(ll - llo) / abs(llo):.2e} (FP32 {abs(llf - llo) / abs(llo):.2e}), "
          f"S0 {errs[0]:.2e} mass {errs[1]:.2e} S1 {errs[2]:.2e} S2 {errs[3]:.2e}")
    assert abs(ll - llo) <= STEP_LL[mode] * abs(llo)
    assert max(errs) <= STEP_MOM[mode], errs


@pytest.mark.parametrize("mode", [F16, BF16])
@pytest.mark.parametrize("case", ["sift128_k64", "sift80_k256"])
def test_em_run_against_oracle(case, mode):
    x, K = CASES[case][0](), CASES[case][1]
    m0, c0, p0 = start(x, K, seed=12)
    mo, co, po, llo, _ = oracle().em(x, m0, c0, p0, 10)
    m, c, pr, ll, q = va.vl_gmm(x, K, init="custom", means=m0, covs=c0, priors=p0, max_iter=10, precision=mode)
    mf, cf, pf, llf, _ = va.vl_gmm(x, K, init="custom", means=m0, covs=c0, priors=p0, max_iter=10,
                                   want_post=False)
    e = [rel(m, mo), rel(c, co), rel(pr, po), abs(ll - llo) / abs(llo)]
    ef = [rel(mf, mo), rel(cf, co), rel(pf, po), abs(llf - llo) / abs(llo)]
    print(f"{case} mode {mode}: means {e[0]:.2e} covs {e[1]:.2e} priors {e[2]:.2e} LL {e[3]:.2e} | "
          f"FP32 {ef[0]:.2e} {ef[1]:.2e} {ef[2]:.2e} {ef[3]:.2e}")
    assert max(e[:3]) <= EM_PARAM[mode], e
    assert e[3] <= 10 * STEP_LL[mode]
    # posteriors handed out after the matrix-core EM: the FP32 E-step of the
    # final parameters, bit for bit
    qr, _ = va.lib().gmm_posteriors(x, m, c, pr)
    assert np.array_equal(q, qr)


@pytest.mark.parametrize("mode", [F16, BF16])
def test_em_restart_path(mode):
    # one starting component far from all data with a small prior: its mass
    # vanishes and it is restarted from iteration 2 (gmm.c:815-999)
    x = synth.mixture(12000, 32, 8, seed=7)
    K = 12
    m0, c0, p0 = start(x, K, seed=13)
    m0[5] = 40.0
    p0[5] = 1e-3
    p0 /= p0.sum()
    mo, co, po, llo, _ = oracle().em(x, m0, c0, p0, 8)
    m, c, pr, ll, _ = va.vl_gmm(x, K, init="custom", means=m0, covs=c0, priors=p0, max_iter=8, precision=mode,
                                want_post=False)
    mf, cf, pf, llf, _ = va.vl_gmm(x, K, init="custom", means=m0, covs=c0, priors=p0, max_iter=8,
                                   want_post=False)
    print(f"restart mode {mode}: means {rel(m, mo):.2e} covs {rel(c, co):.2e} priors {rel(pr, po):.2e} | "
          f"FP32 {rel(mf, mo):.2e} {rel(cf, co):.2e} {rel(pf, po):.2e}")
    # without the restart the far component keeps a zero prior for good
    assert pr[5] > 1e-3 and pf[5] > 1e-3 and po[5] > 1e-3
    assert max(rel(m, mo), rel(c, co), rel(pr, po)) <= EM_PARAM[mode]


def test_precision_setting_roundtrip():
    L = va.lib()
    g = L.vl_gmm_new(1, 16, 4)
    try:
        assert L.vl_gmm_get_precision(g) == 0
        L.vl_gmm_set_precision(g, F16)
        assert L.vl_gmm_get_precision(g) == F16
    finally:
        L.vl_gmm_delete(g)
