"""SVM solvers on the GPU (persistent cooperative launch per run of
iterations) against the reference compiled from source.

The GPU forms each inner product in a fixed parallel order instead of the
reference's one-by-one double sum, so results agree to rounding, not bit for
bit: model and bias within 1e-9 (relative), every statistic of every
diagnostic within 1e-9 (relative, or 1e-12 absolute near zero), the same
number of diagnostics, the same final iteration and status.  Shapes cover one
workgroup (D <= 4096), a few, and the Fisher-vector dimension of the
recognition application's encodings (2KD = 40,960 at K = 256, D = 80), float
and double points, weights with zeros, bias multipliers, a homogeneous kernel
map applied on the device, both solvers and all five losses."""
import numpy as np
import pytest

import vlfeat_amd as va
from oracle import reference

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(reference() is None, reason="reference library not built")]

SDCA, SGD, NONE = 2, 1, 0
LOSSES = {"hinge": 0, "hinge2": 1, "l1": 2, "l2": 3, "logistic": 4}


def problem(n, d, seed, dtype=np.float32):
    rng = np.random.default_rng(seed)
    w = rng.standard_normal(d) / np.sqrt(d)
    x = rng.standard_normal((n, d)) / np.sqrt(d)
    y = np.sign(x @ w + 0.1 * rng.standard_normal(n))
    y[y == 0] = 1
    return x.astype(dtype), y


def close(a, b, tol=1e-9):
    assert np.linalg.norm(a["model"] - b["model"]) <= tol * max(np.linalg.norm(b["model"]), 1e-300)
    assert abs(a["bias"] - b["bias"]) <= tol * max(abs(b["bias"]), 1e-3)
    assert len(a["diagnostics"]) == len(b["diagnostics"])
    for da, db in zip(a["diagnostics"] + [a["statistics"]], b["diagnostics"] + [b["statistics"]]):
        assert da["status"] == db["status"] and da["iteration"] == db["iteration"]
        for f in ("objective", "regularizer", "loss", "dualObjective", "dualLoss", "dualityGap", "scoresVariation"):
            u, v = da[f], db[f]
            if np.isinf(v) or np.isnan(v):
                assert u == v or (np.isnan(u) and np.isnan(v)), (f, u, v)
            else:
                assert abs(u - v) <= max(tol * abs(v), 1e-12), (f, u, v)
    nz = b["scores"] != 0
    assert np.allclose(a["scores"][nz], b["scores"][nz], rtol=tol, atol=1e-12)


def run_both(x, y, lam, **kw):
    a = va.lib().svm_train(x, y, lam, diagnostics=True, **kw)
    b = reference().svm_train(x, y, lam, diagnostics=True, **kw)
    return a, b


@pytest.mark.parametrize("solver", [SDCA, SGD], ids=["sdca", "sgd"])
@pytest.mark.parametrize("loss", list(LOSSES))
def test_gpu_svm_losses(solver, loss):
    x, y = problem(400, 3000, seed=LOSSES[loss] + 7 * solver)
    a, b = run_both(x, y, 0.01, solver=solver, loss=LOSSES[loss], seed=5, diag_freq=150, epsilon=1e-3)
    print(f"{loss} solver {solver}: iterations {a['statistics']['iteration']}, objective "
          f"{a['statistics']['objective']:.6g} vs {b['statistics']['objective']:.6g}")
    close(a, b)


@pytest.mark.parametrize("solver", [SDCA, SGD], ids=["sdca", "sgd"])
@pytest.mark.parametrize("dims", [(300, 24), (250, 9000), (120, 40960)], ids=["d24", "d9000", "d40960"])
@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_gpu_svm_shapes(solver, dims, dtype):
    n, d = dims
    x, y = problem(n, d, seed=d, dtype=dtype)
    a, b = run_both(x, y, 0.005, solver=solver, loss=0, seed=9, epsilon=1e-4)
    close(a, b)


@pytest.mark.parametrize("solver", [SDCA, SGD], ids=["sdca", "sgd"])
def test_gpu_svm_weights_bias_and_map(solver):
    x, y = problem(350, 5000, seed=2)
    w = np.random.default_rng(3).uniform(0.0, 2.0, len(y))
    w[::9] = 0.0
    a, b = run_both(x, y, 0.01, solver=solver, loss=1, seed=4, weights=w, bias_multiplier=0.5, diag_freq=100,
                    max_iter=1200)
    close(a, b)
    xa = np.abs(x)
    a, b = run_both(xa, y, 0.01, solver=solver, loss=0, seed=6, hom=(1, 1.0, 1, -1.0, 1), bias_multiplier=2.0)
    close(a, b)


def test_gpu_svm_warm_start_and_evaluate():
    """Solver None only evaluates the model and bias it is given (the
    statistics, scores and a converged status)."""
    x, y = problem(300, 6000, seed=8)
    L, R = va.lib(), reference()
    a = L.svm_train(x, y, 0.01, solver=SDCA, loss=0, seed=1, max_iter=200)
    b = R.svm_train(x, y, 0.01, solver=SDCA, loss=0, seed=1, max_iter=200)
    close(a, b)
    kw = dict(solver=NONE, model=a["model"], bias=a["bias"], diagnostics=True)
    ea = L.svm_train(x, y, 0.01, **kw)
    eb = R.svm_train(x, y, 0.01, **kw)
    close(ea, eb)
    assert ea["statistics"]["status"] == 2

