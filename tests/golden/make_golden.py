"""Write the golden vectors of this path from the reference itself.

Runs VLFeat 0.9.21's own C implementation, compiled from the reference sources
by oracle/Makefile into oracle/_ref/libvlref.so, on small seeded inputs and
stores inputs and outputs in golden.npz (plain arrays, no pickles) with a
SHA-256 manifest.  Only needed when the fixtures are regenerated; the tests
read the committed files.

    python tests/golden/make_golden.py
"""
from __future__ import annotations

import hashlib
import json
import sys
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
ROOT = HERE.parent.parent
sys.path.insert(0, str(ROOT))

from oracle import build, reference  # noqa: E402
from vlfeat_amd import synth  # noqa: E402
from vlfeat_amd.capi import VlKMeansANN, VlKMeansLloyd, VlKMeansRandomSelection  # noqa: E402


def gmm_params(x, K, seed):
    rng = np.random.default_rng(seed)
    m = x[rng.choice(x.shape[0], K, replace=False)].astype(np.float32)
    c = (np.var(x, 0)[None, :] * rng.uniform(0.1, 0.5, (K, x.shape[1]))).astype(np.float32)
    p = rng.dirichlet(np.ones(K)).astype(np.float32)
    return m, c, p


def main() -> None:
    build(with_reference=True)
    R = reference()
    assert R is not None, "reference library not built"
    R.set_num_threads(1)
    out: dict[str, np.ndarray] = {}

    # E-step and encoders: config-1 style mixture (D=16, K=12) and SIFT-like (D=80, K=24)
    for tag, x, K in [("mix", synth.mixture(1200, 16, 10, seed=21), 12),
                      ("sift", synth.sift_like(900, 80, seed=22), 24)]:
        m, c, p = gmm_params(x, K, seed=23)
        out[f"{tag}_x"], out[f"{tag}_means"], out[f"{tag}_covs"], out[f"{tag}_priors"] = x, m, c, p
        for dt, suf in [(np.float32, "f"), (np.float64, "d")]:
            xd, md, cd, pd = (a.astype(dt) for a in (x, m, c, p))
            q, ll = R.gmm_posteriors(xd, md, cd, pd)
            out[f"{tag}_post_{suf}"] = q
            out[f"{tag}_ll_{suf}"] = np.array(ll)
            for flags in (0, 3, 7):
                e, n = R.fisher(xd, md, cd, pd, flags)
                out[f"{tag}_fisher{flags}_{suf}"] = e
                out[f"{tag}_fisher{flags}_terms_{suf}"] = np.array(n, dtype=np.uint64)
            # VLAD with hard assignments from the exact quantizer of the means
            km = R.kmeans_new(dt)
            try:
                R.vl_kmeans_set_centers(km, md.ctypes.data, md.shape[1], md.shape[0])
                a, dist = R.kmeans_quantize(km, xd)
            finally:
                R.vl_kmeans_delete(km)
            out[f"{tag}_assign_{suf}"] = a
            out[f"{tag}_dist_{suf}"] = dist
            onehot = np.zeros((x.shape[0], K), dt)
            onehot[np.arange(x.shape[0]), a] = 1
            for flags in (0, 1, 2, 4):
                out[f"{tag}_vlad{flags}_{suf}"] = R.vlad(xd, md, onehot, flags)

    # GMM-EM trajectories: custom init (20 iterations) and the default k-means
    # init (seeded RNG through the ANN k-means, kd-forest and restarts)
    x = synth.mixture(3000, 8, 6, seed=31)
    m, c, p = gmm_params(x, 6, seed=32)
    out["em_x"], out["em_means0"], out["em_covs0"], out["em_priors0"] = x, m, c, p
    for dt, suf in [(np.float32, "f"), (np.float64, "d")]:
        xd = x.astype(dt)
        mm, cc, pp, ll, _ = R.gmm(xd, 6, init="custom", means=m.astype(dt), covs=c.astype(dt),
                                  priors=p.astype(dt), max_iter=20, want_post=False)
        out[f"em_custom_means_{suf}"], out[f"em_custom_covs_{suf}"] = mm, cc
        out[f"em_custom_priors_{suf}"], out[f"em_custom_ll_{suf}"] = pp, np.array(ll)
        R.seed(5489)
        mm, cc, pp, ll, _ = R.gmm(xd, 6, init="kmeans", max_iter=20, want_post=False)
        out[f"em_kmeans_means_{suf}"], out[f"em_kmeans_covs_{suf}"] = mm, cc
        out[f"em_kmeans_priors_{suf}"], out[f"em_kmeans_ll_{suf}"] = pp, np.array(ll)

    # k-means: Lloyd and ANN with random-selection seeding
    x = synth.mixture(2500, 12, 8, seed=41)
    out["km_x"] = x
    for algo, name in [(VlKMeansLloyd, "lloyd"), (VlKMeansANN, "ann")]:
        R.seed(5489)
        km = R.kmeans_new(np.float32, algo, VlKMeansRandomSelection, 30, 2, 40, 1)
        try:
            e = R.vl_kmeans_cluster(km, x.ctypes.data, x.shape[1], x.shape[0], 8)
            out[f"km_{name}_centers"] = R.kmeans_centers(km, np.float32)
            out[f"km_{name}_energy"] = np.array(e)
        finally:
            R.vl_kmeans_delete(km)

    np.savez_compressed(HERE / "golden.npz", **out)
    manifest = {k: hashlib.sha256(np.ascontiguousarray(v).tobytes()).hexdigest() for k, v in sorted(out.items())}
    (HERE / "golden_manifest.json").write_text(json.dumps(manifest, indent=1) + "\n")
    print(f"wrote {len(out)} arrays to {HERE / 'golden.npz'}")


if __name__ == "__main__":
    main()
