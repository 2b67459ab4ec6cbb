"""Host-side parts of the library that run without a GPU: the MT19937
generator, the randomised kd-forest and random-selection seeding.  Compared
with published known answers and, where it was compiled here, with the
reference library driven through the same C calls."""
import ctypes as C

import numpy as np
import pytest

import vlfeat_amd
from oracle import reference

needs_ref = pytest.mark.skipif(reference() is None, reason="reference library not built here")


def draws(lib, n):
    r = lib.vl_get_rand()
    return [lib.vl_rand_uint32(r) for _ in range(n)]


def test_mt19937_known_answers():
    L = vlfeat_amd.lib()
    r = L.vl_get_rand()
    key = (C.c_uint32 * 4)(0x123, 0x234, 0x345, 0x456)
    L.vl_rand_seed_by_array(r, key, 4)
    # first outputs of the published mt19937ar test (init_by_array of this key)
    assert draws(L, 5) == [1067595299, 955945823, 477289528, 4107218783, 4228976476]
    L.vl_rand_seed(r, 5489)
    assert draws(L, 1) == [3499211612]  # the standard default-seed first output


@needs_ref
def test_mt19937_matches_reference():
    L, R = vlfeat_amd.lib(), reference()
    for s in (0, 1, 49000, 2 ** 31 + 7):
        L.seed(s)
        R.seed(s)
        assert draws(L, 1500) == draws(R, 1500)


def _kdforest(lib, data, trees, comparisons):
    kd = lib.lib
    kd.vl_kdforest_new.restype = C.c_void_p
    kd.vl_kdforest_new.argtypes = [C.c_uint32, C.c_uint64, C.c_uint64, C.c_int]
    kd.vl_kdforest_build.argtypes = [C.c_void_p, C.c_uint64, C.c_void_p]
    kd.vl_kdforest_set_max_num_comparisons.argtypes = [C.c_void_p, C.c_uint64]
    kd.vl_kdforest_query.restype = C.c_uint64
    kd.vl_kdforest_query.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_void_p]
    kd.vl_kdforest_delete.argtypes = [C.c_void_p]
    t = 1 if data.dtype == np.float32 else 2
    f = kd.vl_kdforest_new(t, data.shape[1], trees, 1)
    kd.vl_kdforest_set_max_num_comparisons(f, comparisons)
    kd.vl_kdforest_build(f, data.shape[0], data.ctypes.data)
    return kd, f


class Neighbor(C.Structure):
    _fields_ = [("distance", C.c_double), ("index", C.c_uint64)]


@needs_ref
@pytest.mark.parametrize("dtype,trees,comparisons,n", [
    (np.float32, 1, 20, 300), (np.float32, 3, 15, 500), (np.float64, 4, 40, 400), (np.float32, 2, 0, 200),
    (np.float32, 1, 50, 3000)])
def test_kdforest_build_and_query_match_reference(dtype, trees, comparisons, n):
    rng = np.random.default_rng(11)
    pts = rng.standard_normal((n, 10)).astype(dtype)
    pts[7] = pts[3]  # duplicates exercise the zero-variance leaves
    queries = rng.standard_normal((200, 10)).astype(dtype)
    out = []
    for lib in (vlfeat_amd.lib(), reference()):
        lib.seed(123)
        kd, f = _kdforest(lib, pts, trees, comparisons)
        nb = (Neighbor * 3)()
        res = []
        for qv in queries:
            ncomp = kd.vl_kdforest_query(f, nb, 3, qv.ctypes.data)
            res.append((ncomp, [(nb[i].index, nb[i].distance) for i in range(3)]))
        kd.vl_kdforest_delete(f)
        out.append((res, draws(lib, 3)))  # generator state after the build
    assert out[0] == out[1]


@needs_ref
@pytest.mark.parametrize("dtype", [np.float32, np.float64])
def test_random_selection_seeding_matches_reference(dtype):
    rng = np.random.default_rng(12)
    x = rng.standard_normal((2000, 8)).astype(dtype)
    x[100:140] = x[0]  # duplicates must be skipped identically
    res = []
    for lib in (vlfeat_amd.lib(), reference()):
        lib.seed(77)
        km = lib.kmeans_new(dtype)
        lib.vl_kmeans_init_centers_with_rand_data(km, x.ctypes.data, 8, 2000, 50)
        res.append(lib.kmeans_centers(km, dtype))
        lib.vl_kmeans_delete(km)
    assert np.array_equal(res[0], res[1])


@needs_ref
@pytest.mark.parametrize("dtype,trees,comparisons", [(np.float32, 2, 30), (np.float64, 3, 0)])
def test_kdforest_query_with_array_and_getters_match_reference(dtype, trees, comparisons):
    # the array query of kdtree.c:870 over several host threads, as
    # toolbox/misc/vl_kdtreequery.c and apps/recognition/encodeImage.m use it
    rng = np.random.default_rng(31)
    pts = rng.standard_normal((1500, 12)).astype(dtype)
    queries = rng.standard_normal((400, 12)).astype(dtype)
    nn = 4
    out = []
    for lib in (vlfeat_amd.lib(), reference()):
        lib.seed(5)
        kd, f = _kdforest(lib, pts, trees, comparisons)
        kd.vl_kdforest_query_with_array.restype = C.c_uint64
        kd.vl_kdforest_query_with_array.argtypes = [C.c_void_p, C.c_void_p, C.c_uint64, C.c_uint64, C.c_void_p,
                                                     C.c_void_p]
        for name, res in (("vl_kdforest_get_num_trees", C.c_uint64), ("vl_kdforest_get_data_dimension", C.c_uint64),
                          ("vl_kdforest_get_data_type", C.c_uint32),
                          ("vl_kdforest_get_thresholding_method", C.c_int)):
            getattr(kd, name).restype = res
            getattr(kd, name).argtypes = [C.c_void_p]
        lib.set_num_threads(4)
        idx = np.empty((400, nn), np.uint32)
        dist = np.empty((400, nn), dtype)
        ncomp = kd.vl_kdforest_query_with_array(f, idx.ctypes.data, nn, 400, dist.ctypes.data, queries.ctypes.data)
        lib.set_num_threads(0)
        getters = (kd.vl_kdforest_get_num_trees(f), kd.vl_kdforest_get_data_dimension(f),
                   kd.vl_kdforest_get_data_type(f), kd.vl_kdforest_get_thresholding_method(f))
        single = []
        nb = (Neighbor * nn)()
        for qv in queries[:20]:
            kd.vl_kdforest_query(f, nb, nn, qv.ctypes.data)
            single.append([nb[i].index for i in range(nn)])
        kd.vl_kdforest_delete(f)
        out.append((idx, dist, ncomp, getters, single))
    (i0, d0, c0, g0, s0), (i1, d1, c1, g1, s1) = out
    assert np.array_equal(i0, i1) and np.array_equal(d0, d1)
    assert c0 == c1 and g0 == g1 and s0 == s1
    assert g0 == (trees, 12, 1 if dtype == np.float32 else 2, 0)
    assert [list(r) for r in i0[:20]] == s0  # the array form equals one query at a time


def test_error_state_is_per_thread():
    import threading
    L = vlfeat_amd.lib().lib
    L.vl_set_last_error.restype = C.c_int
    L.vl_set_last_error.argtypes = [C.c_int, C.c_char_p, C.c_int]
    L.vl_get_last_error.restype = C.c_int
    L.vl_get_last_error_message.restype = C.c_char_p
    assert L.vl_set_last_error(3, b"bad argument %d", 42) == 3
    assert L.vl_get_last_error() == 3 and L.vl_get_last_error_message() == b"bad argument 42"
    seen = []

    def other():
        seen.append((L.vl_get_last_error(), L.vl_get_last_error_message()))
        L.vl_set_last_error(4, None, 0)
        seen.append((L.vl_get_last_error(), L.vl_get_last_error_message()))
    t = threading.Thread(target=other)
    t.start()
    t.join()
    assert seen == [(0, b""), (4, b"")]
    assert L.vl_get_last_error() == 3 and L.vl_get_last_error_message() == b"bad argument 42"
    L.vl_set_last_error(0, None, 0)


def test_kmeans_rejects_l1_distance_with_a_message():
    # only the L2 distance is provided: an L1 k-means aborts with a message
    # before any device work (no GPU needed)
    import subprocess
    import sys
    from pathlib import Path
    code = ("import ctypes as C, numpy as np, vlfeat_amd as va; L = va.lib(); "
            "km = L.vl_kmeans_new(1, 0); x = np.zeros((100, 4), np.float32); "
            "L.vl_kmeans_cluster(km, x.ctypes.data, 4, 100, 3)")
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120,
                       cwd=str(Path(__file__).resolve().parent.parent))
    assert r.returncode != 0
    assert "only the L2 distance" in r.stdout + r.stderr
