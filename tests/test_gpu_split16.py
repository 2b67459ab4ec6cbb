"""Matrix-core Fisher encoder with compensated split-16 operands (BF16x3,
F16x3) against the FP64 oracle.

Tolerance: relative L2 error of the Fisher vector <= 1e-3 (the precision
target of the fast path); the FP16 split is checked to a tighter 2e-4."""
import numpy as np
import pytest

import vlfeat_amd as va
from oracle import oracle
from vlfeat_amd import synth

pytestmark = pytest.mark.gpu

TOL = {va.MODE_BF16X3: 1e-3, va.MODE_F16X3: 2e-4}


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def params(x, K, seed):
    rng = np.random.default_rng(seed)
    m = x[rng.choice(x.shape[0], K, replace=False)].astype(np.float32)
    c = (np.var(x, 0)[None, :] * rng.uniform(0.1, 0.5, (K, x.shape[1]))).astype(np.float32)
    p = rng.dirichlet(np.ones(K)).astype(np.float32)
    return m, c, p


CASES = {
    "sift80_k128": (lambda: synth.sift_like(3000, 80, seed=1), 128),
    "mix64_k64": (lambda: synth.mixture(3000, 64, 64, seed=0), 64),
    "sift40_k100": (lambda: synth.sift_like(2500, 40, seed=2), 100),
    "mix24_k37": (lambda: synth.mixture(1500, 24, 10, seed=3), 37),
    "sift96_k160": (lambda: synth.sift_like(2000, 96, seed=5), 160),
    "sift128_k200": (lambda: synth.sift_like(2000, 128, seed=6, project=False), 200),
    # odd numbers of 16-feature tiles on the paired second pass (K <= 128 or D > 80)
    "mix112_k90": (lambda: synth.mixture(1500, 112, 12, seed=8), 90),
    "mix16_k20": (lambda: synth.mixture(1500, 16, 6, seed=9), 20),
}


@pytest.mark.parametrize("mode", [va.MODE_BF16X3, va.MODE_F16X3], ids=["bf16x3", "f16x3"])
@pytest.mark.parametrize("case", list(CASES))
@pytest.mark.parametrize("flags", [0, 3, 7])
def test_split16_fisher_vs_fp64(mode, case, flags):
    gen, K = CASES[case]
    x = gen()
    m, c, p = params(x, K, seed=K)
    eo, no = oracle().fisher(x, m, c, p, flags)
    with va.fisher_precision(mode):
        enc, n = va.vl_fisher(x, m, c, p, flags)
    err = rel(enc, eo)
    print(f"{case} mode={mode} flags={flags}: rel L2 err {err:.3e}, terms {n} vs {no}")
    assert err <= TOL[mode]
    assert abs(int(n) - int(no)) <= max(3, no // 2000)


@pytest.mark.parametrize("mode", [va.MODE_BF16X3, va.MODE_F16X3], ids=["bf16x3", "f16x3"])
def test_split16_batch_matches_single_images(mode):
    x = synth.sift_like(4000, 80, seed=4)
    m, c, p = params(x, 192, seed=5)
    counts = [0, 1, 700, 64, 65, 1, 1300, 3, 0, 257, 129, 500]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    xs = x[: int(off[-1])]
    with va.fisher_precision(mode):
        enc, terms = va.vl_fisher_batch(xs, off, m, c, p, 3)
        for i in range(len(counts)):
            e1, n1 = va.vl_fisher(xs[off[i]:off[i + 1]], m, c, p, 3)
            assert np.array_equal(enc[i], e1), i
            assert terms[i] == n1
    for i in range(len(counts)):
        if counts[i]:
            eo, _ = oracle().fisher(xs[off[i]:off[i + 1]], m, c, p, 3)
            assert rel(enc[i], eo) <= TOL[mode]
    assert np.all(enc[0] == 0)
