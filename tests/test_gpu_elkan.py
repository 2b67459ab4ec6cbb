"""Elkan's k-means refinement (kmeans.c:1162-1745) against the reference's
own Elkan, compiled from source.

The reference's Elkan keeps float bounds on the distances and stops only when
no datum moves (Lloyd stops on a small relative energy change), so its result
generally differs from its Lloyd's; this library runs the same bound
arithmetic on the GPU, so centers, energy and the kept center-distance table
must equal the reference's Elkan bit for bit: float and double data,
dimensions that are not a multiple of the SIMD lane count, empty clusters
restarted from seeded draws, and large clusters."""
import ctypes as C

import numpy as np
import pytest

import vlfeat_amd as va
from oracle import reference
from vlfeat_amd import synth
from vlfeat_amd.capi import VlKMeans

pytestmark = [pytest.mark.gpu,
              pytest.mark.skipif(reference() is None, reason="reference library not built")]


def center_distances(lib, km, dtype, K):
    s = C.cast(km, C.POINTER(VlKMeans)).contents
    ct = C.c_float if dtype == np.float32 else C.c_double
    return np.ctypeslib.as_array(C.cast(s.centerDistances, C.POINTER(ct)), (K * K,)).copy()


def run(lib, x, K, algo, seed, max_iter, dtype, centers=None, reps=1, verbosity=0):
    N, D = x.shape
    lib.seed(seed)
    km = lib.kmeans_new(dtype, algorithm=algo, max_iter=max_iter, reps=reps, verbosity=verbosity)
    if centers is None:
        e = lib.vl_kmeans_cluster(km, x.ctypes.data, D, N, K)
    else:
        lib.vl_kmeans_set_centers(km, centers.ctypes.data, D, K)
        e = lib.vl_kmeans_refine_centers(km, x.ctypes.data, N)
    c = lib.kmeans_centers(km, dtype)
    cd = center_distances(lib, km, dtype, K) if algo == va.VlKMeansElkan else None
    lib.vl_kmeans_delete(km)
    return c, e, cd


@pytest.mark.parametrize("dtype,D,K,N", [(np.float32, 16, 40, 5000), (np.float32, 13, 24, 8000),
                                         (np.float64, 13, 24, 6000), (np.float64, 32, 50, 4000),
                                         (np.float32, 80, 256, 60000)],
                         ids=["f32-d16", "f32-d13", "f64-d13", "f64-d32", "f32-sift80-k256"])
def test_elkan_cluster_matches_reference_elkan(dtype, D, K, N):
    x = synth.sift_like(N, D, seed=D) if D == 80 else synth.mixture(N, D, K // 2, seed=D, dtype=dtype)
    x = x.astype(dtype)
    R = reference()
    R.set_num_threads(0)
    a = run(va.lib(), x, K, va.VlKMeansElkan, 3, 12, dtype)
    b = run(R, x, K, va.VlKMeansElkan, 3, 12, dtype)
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1]
    assert np.array_equal(a[2], b[2])


def test_elkan_takes_the_reference_elkan_trajectory_not_lloyd():
    """A case where the reference's Elkan and Lloyd disagree (Lloyd stops when
    the relative energy change falls below 1e-4, Elkan only when no datum
    moves): this library's Elkan follows Elkan's."""
    x = synth.sift_like(20000, 32, seed=4)
    R = reference()
    R.set_num_threads(0)
    found = False
    for seed in range(1, 6):
        el = run(va.lib(), x, 100, va.VlKMeansElkan, seed, 200, np.float32)
        er = run(R, x, 100, va.VlKMeansElkan, seed, 200, np.float32)
        lr = run(R, x, 100, va.VlKMeansLloyd, seed, 200, np.float32)
        assert np.array_equal(el[0], er[0]) and el[1] == er[1]
        if not np.array_equal(er[0], lr[0]):
            found = True
            print(f"seed {seed}: reference Elkan and Lloyd differ by {np.abs(er[0] - lr[0]).max():.3g}; "
                  f"energies {er[1]:.8g} / {lr[1]:.8g}")
            break
    assert found


@pytest.mark.parametrize("dtype", [np.float32, np.float64], ids=["f32", "f64"])
def test_elkan_restarts_empty_clusters_as_reference(dtype):
    """Given centers, a quarter of them far from every datum: their clusters
    are empty after the first assignment and restart from data drawn by the
    seeded generator, as in the reference."""
    N, D, K = 30000, 13, 48
    x = synth.mixture(N, D, 12, seed=21, dtype=dtype)
    rng = np.random.default_rng(3)
    cen0 = x[rng.choice(N, K, replace=False)].copy()
    cen0[::4] += dtype(50.0)
    R = reference()
    R.set_num_threads(0)
    a = run(va.lib(), x, K, va.VlKMeansElkan, 11, 15, dtype, centers=cen0)
    b = run(R, x, K, va.VlKMeansElkan, 11, 15, dtype, centers=cen0)
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1]
    assert np.array_equal(a[2], b[2])


def test_elkan_large_clusters_repetitions_and_verbose():
    N, D, K = 200_000, 80, 16
    x = synth.sift_like(N, D, seed=5)
    R = reference()
    R.set_num_threads(0)
    a = run(va.lib(), x, K, va.VlKMeansElkan, 2, 8, np.float32, reps=2, verbosity=1)
    b = run(R, x, K, va.VlKMeansElkan, 2, 8, np.float32, reps=2)
    assert np.array_equal(a[0], b[0])
    assert a[1] == b[1]
