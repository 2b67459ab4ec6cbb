"""Golden vectors written by the reference itself (tests/golden/make_golden.py
runs VLFeat 0.9.21 compiled from source): the FP64 oracle against the
reference's double path, and the GPU library against the reference's float and
double paths, without needing the reference at test time.

Tolerances: oracle vs reference double 1e-9 relative (different summation
order only); GPU float vs reference float as in the parity tests (Fisher / VLAD
1e-4 relative L2, posteriors 1e-4, GMM means / variances 1e-4, hard
assignments bit-exact); GPU double vs reference double 1e-9.
"""
import hashlib
import json
from pathlib import Path

import numpy as np
import pytest

from oracle import oracle

G = Path(__file__).resolve().parent / "golden"


@pytest.fixture(scope="module")
def gold():
    with np.load(G / "golden.npz", allow_pickle=False) as z:
        return {k: z[k] for k in z.files}


def rel(a, b):
    a = np.asarray(a, np.float64).ravel()
    b = np.asarray(b, np.float64).ravel()
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))


def test_golden_manifest(gold):
    manifest = json.loads((G / "golden_manifest.json").read_text())
    assert sorted(manifest) == sorted(gold)
    for k, v in gold.items():
        assert hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() == manifest[k], k


@pytest.mark.parametrize("tag", ["mix", "sift"])
def test_oracle_matches_golden_double(gold, tag):
    O = oracle()
    x, m, c, p = (gold[f"{tag}_{s}"].astype(np.float64) for s in ("x", "means", "covs", "priors"))
    q, ll = O.posteriors(x, m, c, p)
    assert rel(q, gold[f"{tag}_post_d"]) < 1e-9
    assert abs(ll - float(gold[f"{tag}_ll_d"])) <= 1e-9 * abs(ll)
    for flags in (0, 3, 7):
        e, n = O.fisher(x, m, c, p, flags)
        assert rel(e, gold[f"{tag}_fisher{flags}_d"]) < 1e-9, flags
        assert n == int(gold[f"{tag}_fisher{flags}_terms_d"])
    a, d = O.quantize(x, m)
    np.testing.assert_array_equal(a, gold[f"{tag}_assign_d"])
    assert rel(d, gold[f"{tag}_dist_d"]) < 1e-12
    onehot = np.zeros((x.shape[0], m.shape[0]))
    onehot[np.arange(x.shape[0]), a] = 1
    for flags in (0, 1, 2, 4):
        assert rel(O.vlad(x, m, onehot, flags), gold[f"{tag}_vlad{flags}_d"]) < 1e-9, flags


def test_oracle_em_matches_golden_double(gold):
    x = gold["em_x"].astype(np.float64)
    m, c, p, ll, _ = oracle().em(x, gold["em_means0"], gold["em_covs0"], gold["em_priors0"], 20)
    assert rel(m, gold["em_custom_means_d"]) < 1e-9
    assert rel(c, gold["em_custom_covs_d"]) < 1e-9
    assert rel(p, gold["em_custom_priors_d"]) < 1e-9
    assert abs(ll - float(gold["em_custom_ll_d"])) <= 1e-9 * abs(ll)


# ---------------------------------------------------------------- GPU library

@pytest.mark.gpu
@pytest.mark.parametrize("tag", ["mix", "sift"])
@pytest.mark.parametrize("suf", ["f", "d"])
def test_gpu_matches_golden_encoders(gold, tag, suf):
    import vlfeat_amd as va
    dt = np.float32 if suf == "f" else np.float64
    tol = 1e-4 if suf == "f" else 1e-9
    x, m, c, p = (gold[f"{tag}_{s}"].astype(dt) for s in ("x", "means", "covs", "priors"))
    q, ll = va.lib().gmm_posteriors(x, m, c, p)
    assert rel(q, gold[f"{tag}_post_{suf}"]) < tol
    for flags in (0, 3, 7):
        e, n = va.vl_fisher(x, m, c, p, flags)
        assert rel(e, gold[f"{tag}_fisher{flags}_{suf}"]) < tol, flags
        assert abs(int(n) - int(gold[f"{tag}_fisher{flags}_terms_{suf}"])) <= (2 if suf == "f" else 0)
    a, d = va.vl_kmeans_quantize(x, m)
    np.testing.assert_array_equal(a, gold[f"{tag}_assign_{suf}"])
    assert rel(d, gold[f"{tag}_dist_{suf}"]) < (1e-6 if suf == "f" else 1e-12)
    onehot = np.zeros((x.shape[0], m.shape[0]), dt)
    onehot[np.arange(x.shape[0]), gold[f"{tag}_assign_{suf}"]] = 1
    for flags in (0, 1, 2, 4):
        assert rel(va.vl_vlad(x, m, onehot, flags), gold[f"{tag}_vlad{flags}_{suf}"]) < tol, flags


@pytest.mark.gpu
@pytest.mark.parametrize("suf", ["f", "d"])
def test_gpu_matches_golden_em(gold, suf):
    import vlfeat_amd as va
    dt = np.float32 if suf == "f" else np.float64
    tol = 1e-4 if suf == "f" else 1e-9
    x = gold["em_x"].astype(dt)
    m, c, p, ll, _ = va.vl_gmm(x, 6, init="custom", means=gold["em_means0"].astype(dt),
                               covs=gold["em_covs0"].astype(dt), priors=gold["em_priors0"].astype(dt),
                               max_iter=20, want_post=False)
    assert rel(m, gold[f"em_custom_means_{suf}"]) < tol
    assert rel(c, gold[f"em_custom_covs_{suf}"]) < tol
    assert rel(p, gold[f"em_custom_priors_{suf}"]) < tol
    # default k-means initialisation: the seeded RNG trajectory through the ANN
    # k-means, the kd-forest and the EM restarts
    va.lib().seed(5489)
    m, c, p, ll, _ = va.vl_gmm(x, 6, init="kmeans", max_iter=20, want_post=False)
    assert rel(m, gold[f"em_kmeans_means_{suf}"]) < tol
    assert rel(c, gold[f"em_kmeans_covs_{suf}"]) < tol
    assert rel(p, gold[f"em_kmeans_priors_{suf}"]) < tol


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["lloyd", "ann"])
def test_gpu_matches_golden_kmeans(gold, name):
    import vlfeat_amd as va
    from vlfeat_amd.capi import VlKMeansANN, VlKMeansLloyd, VlKMeansRandomSelection
    x = gold["km_x"]
    va.lib().seed(5489)
    algo = VlKMeansLloyd if name == "lloyd" else VlKMeansANN
    centers, e = va.vl_kmeans(x, 8, algorithm=algo, init=VlKMeansRandomSelection, max_iter=30, num_trees=2,
                              max_comparisons=40)
    assert rel(centers, gold[f"km_{name}_centers"]) < 1e-5
    assert abs(e - float(gold[f"km_{name}_energy"])) <= 1e-5 * abs(e)
