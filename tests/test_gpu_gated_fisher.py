"""Gated Fisher encoder (MODE_GATED) against the dense F16X3 encoder and the
FP64 oracle.

The exact first pass judges the gate: a group is served sparsely only when,
in every row, the gate's 16 candidates carry an exact posterior mass (their
logits normalised by the first pass's log-sum-exp over all components) of at
least `need`; the network's own predicted mass only pre-screens rows.

* With a gate that is never confident every group takes the dense path, and
  the encodings equal the F16X3 encoder's bit for bit.
* On a mixture with peaked posteriors a trained gate serves groups at need
  0.999 with every image (the maximum over the images, not a median) within
  1e-3 relative L2 of the FP64 oracle; the served fraction and the worst
  image are printed for 0.99, 0.999 and 0.9999.
* Over 100 images of 3,000 descriptors (more than the gate's probe) the worst
  image stays within 1e-3 of the FP64 oracle at need 0.999.
"""
import concurrent.futures as cf

import numpy as np
import pytest

import vlfeat_amd as va
from oracle import oracle
from vlfeat_amd import synth
from vlfeat_amd.gate import Gate

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


@pytest.fixture(scope="module")
def peaked():
    K, D = 256, 32
    x = synth.mixture(60000, D, 96, seed=11)
    m, c, p = gmm_params(x, K, seed=12)
    g = Gate(D, K, 128, 128, seed=3)
    g.standardize_from(x)
    g.train(x, m, c, p, steps=500, batch=2048, lr=2e-3)
    return x, m, c, p, g


@pytest.mark.parametrize("flags", [3, 7])
def test_gated_never_confident_equals_f16x3(peaked, flags):
    x, m, c, p, g = peaked
    counts = [300, 1, 64, 700, 33]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    xs = x[: int(off[-1])]
    with va.fisher_precision(va.MODE_F16X3):
        e16, t16 = va.vl_fisher_batch(xs, off, m, c, p, flags)
    va.set_fisher_gate(g, 16, 1.5)  # unreachable mass: every group dense
    try:
        with va.fisher_precision(va.MODE_GATED):
            eg, tg = va.vl_fisher_batch(xs, off, m, c, p, flags)
    finally:
        va.set_fisher_gate(None)
    assert np.array_equal(eg, e16)
    assert np.array_equal(tg, t16)


@pytest.mark.parametrize("flags", [0, 3, 7])
def test_gated_peaked_mixture_vs_fp64(peaked, flags):
    x, m, c, p, g = peaked
    counts = [3000, 1500, 2048, 257, 31]
    off = np.concatenate([[0], np.cumsum(counts)]).astype(np.uint64)
    xs = x[10000: 10000 + int(off[-1])]
    with va.fisher_precision(va.MODE_F16X3):
        ed, td = va.vl_fisher_batch(xs, off, m, c, p, flags)
    ref = [oracle().fisher(xs[off[i]:off[i + 1]], m, c, p, flags) for i in range(len(counts))]
    worst = {}
    for need in (0.99, 0.999, 0.9999):
        va.set_fisher_gate(g, 16, need)
        va.fisher_gate_stats(reset=True)
        try:
            with va.fisher_precision(va.MODE_GATED):
                eg, tg = va.vl_fisher_batch(xs, off, m, c, p, flags)
        finally:
            va.set_fisher_gate(None)
        groups, served = va.fisher_gate_stats(reset=True)
        errs = [rel(eg[i], ref[i][0]) for i in range(len(counts))]
        dterms = [abs(int(tg[i]) - int(ref[i][1])) / max(1, ref[i][1]) for i in range(len(counts))]
        worst[need] = (max(errs), max(dterms), served / max(groups, 1))
        print(f"flags={flags} need={need}: served {served}/{groups} groups, max rel err {max(errs):.2e} "
              f"(dense {max(rel(ed[i], ref[i][0]) for i in range(len(counts))):.2e}), "
              f"max term deviation {max(dterms):.2e}")
    err, dt, frac = worst[0.999]
    assert err <= 1e-3
    assert dt <= 0.01
    assert frac >= 0.1
    # a stricter bar never serves more groups
    assert worst[0.9999][2] <= worst[0.999][2] <= worst[0.99][2]


def test_gated_100_images_worst_image(peaked):
    # a batch larger than the gate's probe: a gate that serves nothing on the
    # probe is not run on the rest (dense path, bit for bit F16X3); a useful
    # gate keeps serving the whole batch, and the worst of 100 images stays
    # within 1e-3 of the FP64 oracle
    _, m, c, p, g = peaked
    x = synth.mixture(300000, 32, 96, seed=11)
    off = (np.arange(101) * 3000).astype(np.uint64)
    with va.fisher_precision(va.MODE_F16X3):
        e16, t16 = va.vl_fisher_batch(x, off, m, c, p, 3)
    va.set_fisher_gate(g, 16, 1.5)
    try:
        with va.fisher_precision(va.MODE_GATED):
            eg, tg = va.vl_fisher_batch(x, off, m, c, p, 3)
    finally:
        va.set_fisher_gate(None)
    assert np.array_equal(eg, e16) and np.array_equal(tg, t16)
    O = oracle()
    with cf.ThreadPoolExecutor(8) as ex:
        ref = list(ex.map(lambda i: O.fisher(x[off[i]:off[i + 1]], m, c, p, 3)[0], range(100)))
    dense = np.array([rel(e16[i], ref[i]) for i in range(100)])
    stats = {}
    for need in (0.999, 0.9999):
        va.set_fisher_gate(g, 16, need)
        va.fisher_gate_stats(reset=True)
        try:
            with va.fisher_precision(va.MODE_GATED):
                eg, _ = va.vl_fisher_batch(x, off, m, c, p, 3)
        finally:
            va.set_fisher_gate(None)
        groups, served = va.fisher_gate_stats(reset=True)
        errs = np.array([rel(eg[i], ref[i]) for i in range(100)])
        stats[need] = (served / groups, float(errs.max()))
        print(f"100 images need={need}: served {served}/{groups} groups, rel err to FP64 median "
              f"{np.median(errs):.2e}, 90th pct {np.percentile(errs, 90):.2e}, max {errs.max():.2e} "
              f"(dense F16X3 max {dense.max():.2e})")
    assert stats[0.999][0] >= 0.1          # the gate stays on for the whole batch
    assert stats[0.999][1] <= 1e-3         # every image within the bound
    assert stats[0.9999][1] <= 1e-3
