"""Learned gate: GPU forward/backward against the FP64 oracle, training, the
candidate selection, and what the trained network gains over the same
selection without training and over a nearest-means heuristic.

Tolerances: the GPU multiplies BF16 operands (inputs, weights, activations
and back-propagated errors are rounded to BF16 before each product) with FP32
accumulation; the oracle evaluates the same BF16-rounded weights and inputs in
FP64.  Per-tensor gradient relative L2 error <= 1e-2, loss relative error <= 2e-3.
"""
import numpy as np
import pytest

import vlfeat_amd as va
from oracle import oracle
from vlfeat_amd import synth
from vlfeat_amd.gate import Gate, bf16_round

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


@pytest.mark.parametrize("D,H1,H2,K,rows", [(80, 256, 256, 1024, 64), (24, 40, 24, 37, 50)])
def test_gate_gradients_match_fp64_oracle(D, H1, H2, K, rows):
    x = synth.sift_like(4000, D, seed=3) if D == 80 else synth.mixture(4000, D, 8, seed=3)
    m, c, p = gmm_params(x, K, seed=4)
    g = Gate(D, K, H1, H2, seed=7)
    g.standardize_from(x)
    xb = x[:rows]
    q, _ = oracle().posteriors(xb, m, c, p)
    loss, grads = g.gradients(xb, q.astype(np.float32))
    P, mean, istd = g.params()
    # the oracle sees what the GPU multiplies: BF16 weights and inputs
    Pr = P.copy()
    for i, t in enumerate(g.split(Pr)):
        if i % 2 == 0:
            t[...] = bf16_round(t)
    xin = bf16_round((xb.astype(np.float32) - mean) * istd)
    lo, go = oracle().mlp_loss_grads(Pr, D, H1, H2, K, xin, q.astype(np.float32).astype(np.float64))
    print(f"loss gpu {loss:.6f} oracle {lo:.6f}")
    assert abs(loss - lo) / abs(lo) < 2e-3
    names = ["W1", "b1", "W2", "b2", "W3", "b3"]
    for name, a, b in zip(names, g.split(grads), g.split(go)):
        e = rel(a, b)
        print(f"{name}: rel L2 {e:.3e}")
        assert e < 1e-2, name


def test_gate_training_reduces_loss():
    x = synth.sift_like(60000, 80, seed=5)
    m, c, p = gmm_params(x, 64, seed=6)
    g = Gate(80, 64, 128, 128, seed=1)
    g.standardize_from(x)
    trace = g.train(x, m, c, p, steps=300, batch=2048, lr=2e-3)
    first, last = float(trace[:10].mean()), float(trace[-10:].mean())
    print(f"loss {first:.4f} -> {last:.4f}")
    assert np.all(np.isfinite(trace))
    assert last < 0.5 * first


def test_gate_select_matches_numpy_topk():
    x = synth.mixture(3000, 16, 32, seed=8)
    m, c, p = gmm_params(x, 48, seed=9)
    g = Gate(16, 48, 64, 64, seed=2)
    g.standardize_from(x)
    g.train(x, m, c, p, steps=150, batch=1024, lr=3e-3)
    M, need = 8, 0.9
    idx, mass, conf = g.select(x, M, need)
    z = g.logits(x).astype(np.float64)
    pz = np.exp(z - z.max(1, keepdims=True))
    pz /= pz.sum(1, keepdims=True)
    order = np.argsort(-pz, axis=1, kind="stable")[:, :M]
    topmass = np.take_along_axis(pz, order, 1).sum(1)
    sel = idx[:, 0] >= 0
    # rows that were selected: same top-M set and mass; screened rows: true mass below need
    for i in np.nonzero(sel)[0]:
        assert set(idx[i]) == set(order[i]), i
    np.testing.assert_allclose(mass[sel], topmass[sel], rtol=1e-4, atol=1e-5)
    assert np.all(topmass[~sel] < need + 1e-4)
    np.testing.assert_array_equal(conf, sel & (topmass >= need - 1e-5) | (conf & sel))
    assert np.all(conf <= sel)
    print(f"selected {sel.mean():.3f}, confident {conf.mean():.3f}")


def test_gate_gain_over_untrained_and_nearest_means():
    """What the trained network gains: on a mixture with peaked posteriors the
    trained gate's 16 candidates carry >= 0.99 of the exact posterior mass for
    most descriptors (they can skip the dense E-step); the same network
    untrained, and the 16 nearest means by Euclidean distance, do worse."""
    K, D, M = 256, 32, 16
    x = synth.mixture(40000, D, 96, seed=11)
    m, c, p = gmm_params(x, K, seed=12)
    q, _ = oracle().posteriors(x[:3000], m, c, p)

    def captured(cands):
        return np.take_along_axis(q, cands, 1).sum(1)

    g0 = Gate(D, K, 128, 128, seed=3)
    g0.standardize_from(x)
    g1 = Gate(D, K, 128, 128, seed=3)
    g1.standardize_from(x)
    g1.train(x, m, c, p, steps=400, batch=2048, lr=2e-3)
    top = lambda z: np.argsort(-z, axis=1, kind="stable")[:, :M]
    cap_untrained = captured(top(g0.logits(x[:3000])))
    cap_trained = captured(top(g1.logits(x[:3000])))
    d2 = ((x[:3000, None, :].astype(np.float64) - m[None].astype(np.float64)) ** 2).sum(-1)
    cap_nearest = captured(np.argsort(d2, axis=1, kind="stable")[:, :M])
    exact_top = np.sort(q, 1)[:, ::-1][:, :M].sum(1)
    frac = lambda a: float((a >= 0.99).mean())
    print(f"rows with >=0.99 of the mass in the 16 candidates: exact top-16 {frac(exact_top):.3f}, "
          f"trained {frac(cap_trained):.3f}, untrained {frac(cap_untrained):.3f}, nearest means {frac(cap_nearest):.3f}; "
          f"mean captured mass trained {cap_trained.mean():.4f} untrained {cap_untrained.mean():.4f} "
          f"nearest {cap_nearest.mean():.4f}")
    assert frac(cap_trained) > frac(cap_untrained) + 0.3
    assert cap_trained.mean() > cap_nearest.mean()
    assert frac(cap_trained) >= 0.8 * frac(exact_top)
