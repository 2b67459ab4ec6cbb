"""A C program compiled against include/vl/*.h and linked with -lvl
(tests/c_client/vl_client.c) runs the reference's application sequence --
GMM by EM, an improved Fisher vector, k-means quantisation, VLAD, kd-forest
words -- and its outputs equal the same calls made through ctypes, bit for
bit (the FP32 paths are deterministic)."""
import shutil
import subprocess
from pathlib import Path

import numpy as np
import pytest

import vlfeat_amd as va
from vlfeat_amd import synth

pytestmark = pytest.mark.gpu
ROOT = Path(__file__).resolve().parent.parent


def test_c_client_matches_ctypes(tmp_path):
    cc = shutil.which("cc") or shutil.which("gcc")
    if cc is None:
        pytest.skip("no C compiler")
    exe = tmp_path / "vl_client"
    lib = ROOT / "vlfeat_amd" / "lib"
    subprocess.run([cc, "-std=c99", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(ROOT / "tests" / "c_client" / "vl_client.c"),
                    f"-L{lib}", "-lvl", f"-Wl,-rpath,{lib}", "-o", str(exe)], check=True)
    N, D, K, n = 20000, 80, 64, 3000
    x = synth.sift_like(N, D, seed=12)
    rng = np.random.default_rng(3)
    m0 = x[rng.choice(N, K, replace=False)]
    c0 = np.tile(np.var(x, 0), (K, 1)).astype(np.float32)
    p0 = np.full(K, 1.0 / K, np.float32)
    x.tofile(tmp_path / "data.f32")
    np.concatenate([m0.ravel(), c0.ravel(), p0]).astype(np.float32).tofile(tmp_path / "init.f32")
    r = subprocess.run([str(exe), str(tmp_path), str(N), str(D), str(K), str(n)], capture_output=True, text=True,
                       timeout=300)
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    rd = lambda name, dt: np.fromfile(tmp_path / name, dtype=dt)
    m, c, p, ll, _ = va.vl_gmm(x, K, init="custom", means=m0, covs=c0, priors=p0, max_iter=10, want_post=False)
    assert np.array_equal(rd("gmm.f32", np.float32), m.ravel())
    assert np.array_equal(rd("gmm_covs.f32", np.float32), c.ravel())
    assert np.array_equal(rd("gmm_priors.f32", np.float32), p)
    assert rd("ll.f64", np.float64)[0] == ll
    fv, _ = va.vl_fisher(x[:n], m, c, p, va.FISHER_IMPROVED)
    assert np.array_equal(rd("fisher.f32", np.float32), fv)
    a, _ = va.vl_kmeans_quantize(x[:n], m)
    assert np.array_equal(rd("assign.u32", np.uint32), a)
    oh = np.zeros((n, K), np.float32)
    oh[np.arange(n), a] = 1
    assert np.array_equal(rd("vlad.f32", np.float32), va.vl_vlad(x[:n], m, oh, va.VLAD_SQUARE_ROOT))
    # an exhaustive kd-forest query finds the exact nearest means (the forest
    # orders distances in double; ties between equal distances aside)
    assert (rd("words.u32", np.uint32) == a).mean() > 0.999
