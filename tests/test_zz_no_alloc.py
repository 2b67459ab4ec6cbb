"""vlhip_kmeans_quantize_noalloc (no device allocation, static centre table)
against vlhip_kmeans_quantize on the same device inputs: assignments and
distances bit for bit, on consecutive calls of decreasing size (the second
call reuses the table the first one filled), for the matrix-core screened
shapes and for the exhaustive kernel's (double data, wide descriptors, few
centres).  The size check runs before any device work, so its test needs no
GPU: a child process asks for more centres than the static table holds and
must abort with the message."""
import subprocess
import sys
from pathlib import Path

import numpy as np
import pytest

import vlfeat_amd as va
from vlfeat_amd import synth
from vlfeat_amd.gate import DeviceArray

ROOT = Path(__file__).resolve().parent.parent


def both(x, c, lanes=8):
    L = va.lib()
    N, D = x.shape
    K = c.shape[0]
    dt = 1 if x.dtype == np.float32 else 2
    dx, dc = DeviceArray(x), DeviceArray(c)
    out = []
    for fn in (L.vlhip_kmeans_quantize, L.vlhip_kmeans_quantize_noalloc):
        da = DeviceArray(nbytes=4 * N)
        dd = DeviceArray(nbytes=x.itemsize * N)
        fn(dt, da.ptr, dd.ptr, dx.ptr, dc.ptr, N, D, K, lanes)
        L.vlhip_synchronize()
        out.append((da.to_host(np.uint32, (N,)), dd.to_host(x.dtype, (N,))))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("shapes", [
    [(200000, 80, 4096), (3000, 80, 512)],      # the table's maximum, then a smaller call
    [(50000, 128, 1024), (777, 37, 64)],        # widest screened dimension, then the smallest K
])
def test_noalloc_quantize_equals_allocating_entry_point(shapes):
    for N, D, K in shapes:
        x = synth.sift_like(N, D, seed=N + D, project=D != 128)
        c = synth.sift_like(K, D, seed=K + 7, project=D != 128)
        c[K // 2] = c[K // 3]  # duplicate centres: the first index wins
        (a1, d1), (a2, d2) = both(x, c)
        print(f"N={N} D={D} K={K}: {int((a1 == a2).sum())}/{N} assignments equal")
        assert np.array_equal(a1, a2), (N, D, K)
        assert np.array_equal(d1, d2), (N, D, K)


@pytest.mark.gpu
@pytest.mark.parametrize("N,D,K,dt", [(4000, 200, 300, np.float32), (3000, 64, 256, np.float64),
                                       (5000, 80, 40, np.float32)])
def test_noalloc_quantize_exhaustive_shapes(N, D, K, dt):
    rng = np.random.default_rng(N)
    x = rng.standard_normal((N, D)).astype(dt)
    c = rng.standard_normal((K, D)).astype(dt)
    for lanes in (8, 4):
        (a1, d1), (a2, d2) = both(x, c, lanes)
        assert np.array_equal(a1, a2) and np.array_equal(d1, d2), lanes


def test_noalloc_quantize_rejects_more_centers_than_its_table():
    code = ("import vlfeat_amd as va; "
            "va.lib().vlhip_kmeans_quantize_noalloc(1, None, None, None, None, 10, 80, 4097, 8)")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       cwd=str(ROOT))
    assert r.returncode != 0
    assert "at most 4096 centers" in r.stderr
