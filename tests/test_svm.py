"""SVM solvers (vl/svm.c) on the host: the solver loop that runs when the
caller supplies its own data functions (vl_svm_new_with_abstract_data), and
the predefined losses, against the reference compiled from source.

The host loop performs the reference's double operations in its order, so
models, biases, statistics, scores and the diagnostic trajectory are equal bit
for bit for both solvers and all five losses, with weights, bias multipliers,
float and double data and a homogeneous kernel map.  Independently of the
reference: weak duality (primal objective >= dual objective at every SDCA
diagnostic), the duality gap below epsilon when SDCA reports convergence, and
a separable problem classified without error."""
import numpy as np
import pytest

import vlfeat_amd as va
from oracle import reference

needs_ref = pytest.mark.skipif(reference() is None, reason="reference library not built")

SDCA, SGD = 2, 1
LOSSES = {"hinge": 0, "hinge2": 1, "l1": 2, "l2": 3, "logistic": 4}


def problem(n, d, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(d)
    x = rng.standard_normal((n, d))
    y = np.sign(x @ w + 0.3 * rng.standard_normal(n))
    y[y == 0] = 1
    return x.astype(dtype), y


def same(a, b):
    for key in ("model", "scores"):
        assert np.array_equal(a[key], b[key]), key
    assert a["bias"] == b["bias"]
    for f, v in a["statistics"].items():
        w = b["statistics"][f]
        assert v == w or (np.isnan(v) and np.isnan(w)), (f, v, w)
    assert len(a["diagnostics"]) == len(b["diagnostics"])
    for da, db in zip(a["diagnostics"], b["diagnostics"]):
        assert da == db


@needs_ref
@pytest.mark.parametrize("solver", [SDCA, SGD], ids=["sdca", "sgd"])
@pytest.mark.parametrize("loss", list(LOSSES))
def test_host_loop_equals_reference(solver, loss):
    x, y = problem(300, 24, seed=LOSSES[loss] + 10 * solver)
    kw = dict(solver=solver, loss=LOSSES[loss], seed=7, diag_freq=97, diagnostics=True, epsilon=1e-3)
    a = va.lib().svm_train(x, y, 0.01, abstract=True, **kw)
    b = reference().svm_train(x, y, 0.01, abstract=True, **kw)
    same(a, b)
    assert a["statistics"]["iteration"] > 0


@needs_ref
@pytest.mark.parametrize("solver", [SDCA, SGD], ids=["sdca", "sgd"])
def test_host_loop_options_equal_reference(solver):
    x, y = problem(200, 17, seed=3, dtype=np.float64)
    w = np.random.default_rng(1).uniform(0.0, 2.0, len(y))
    w[::11] = 0.0
    for extra in (dict(bias_multiplier=0.5, weights=w), dict(bias_multiplier=2.0, bias_lr=0.3, max_iter=450),
                  dict(hom=(1, 1.0, 1, -1.0, 1))):
        kw = dict(solver=solver, loss=0, seed=11, diag_freq=64, diagnostics=True, **extra)
        xx = np.abs(x) if "hom" in extra else x
        a = va.lib().svm_train(xx, y, 0.02, abstract=True, **kw)
        b = reference().svm_train(xx, y, 0.02, abstract=True, **kw)
        same(a, b)


@needs_ref
def test_predefined_losses_equal_reference():
    L, R = va.lib(), reference()
    rng = np.random.default_rng(0)
    for _ in range(200):
        inner, label, alpha, n2 = rng.normal(0, 2), rng.choice([-1.0, 1.0]), rng.uniform(-1, 1), rng.uniform(0.01, 3)
        assert L.vl_svm_hinge_loss(inner, label) == R.vl_svm_hinge_loss(inner, label)
        beta = abs(alpha)
        a = label * beta
        assert L.vl_svm_logistic_dca_update(a, inner, n2, label) == R.vl_svm_logistic_dca_update(a, inner, n2, label)


def test_sdca_duality_and_separable_problem():
    """Weak duality and convergence, without the reference."""
    x, y = problem(400, 10, seed=5, dtype=np.float64)
    out = va.lib().svm_train(x, y, 0.001, solver=SDCA, loss=0, seed=1, abstract=True, diagnostics=True,
                             epsilon=1e-4)
    for d in out["diagnostics"]:
        assert d["objective"] >= d["dualObjective"] - 1e-12
    st = out["statistics"]
    if st["status"] == 2:  # converged
        assert st["dualityGap"] < 1e-4
    # separable: labels from a noiseless linear rule
    rng = np.random.default_rng(2)
    xs = rng.standard_normal((300, 5))
    ys = np.sign(xs @ np.array([1.0, -2.0, 0.5, 0.0, 1.5]) + 0.1)
    out = va.lib().svm_train(xs, ys, 1e-4, solver=SDCA, loss=0, seed=3, abstract=True, max_iter=30000)
    pred = np.sign(xs @ out["model"] + out["bias"])
    assert np.mean(pred == ys) >= 0.99
