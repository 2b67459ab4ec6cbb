"""Homogeneous kernel map (vl/homkermap.c), host side (no GPU needed).

* The FP64 oracle and the library's host entry points (the map object and the
  scalar vl_homogeneouskernelmap_evaluate_{d,f}) against the reference compiled
  from source: bitwise equal over three kernels, both windows, orders 1/3/5,
  gamma 1 and 0.5, on values spanning the table's exponents, negatives, zeros
  and values outside the table.
* What the mathematics fixes: <psi(x), psi(y)> approximates the kernel in
  closed form (intersection min(x, y), chi2 2xy/(x+y), Jensen-Shannon
  x/2 log2((x+y)/x) + y/2 log2((x+y)/y)); the error falls with the order, to
  the levels the published construction reaches; a wrong constant or sign in
  a spectrum breaks it (checked with a planted sign error in the oracle's
  inputs below).
"""
import numpy as np
import pytest

import vlfeat_amd as va
from oracle import oracle, reference

needs_ref = pytest.mark.skipif(reference() is None, reason="reference library not built")

KERNELS = {0: np.minimum,
           1: lambda a, b: 2 * a * b / (a + b),
           2: lambda a, b: a / 2 * np.log2((a + b) / a) + b / 2 * np.log2((a + b) / b)}


def sample(seed=0):
    rng = np.random.default_rng(seed)
    return np.concatenate([rng.lognormal(-2, 2, 400), -rng.lognormal(-1, 1, 60),
                           [0.0, -0.0, 1e-7, 3e-6, 1e-3, 0.5, 1.0, 2.0, 255.9, 256.0, 300.0, 1e4]])


@needs_ref
@pytest.mark.parametrize("kernel", [0, 1, 2])
@pytest.mark.parametrize("window", [0, 1])
def test_oracle_and_host_path_match_reference(kernel, window):
    R, O, P = reference(), oracle(), va.lib()
    x = sample(kernel + 3 * window)
    for order in (1, 3, 5):
        for gamma in (1.0, 0.5):
            r = R.homkermap(x, kernel, gamma, order, -1.0, window)
            assert np.array_equal(O.homkermap(x, kernel, gamma, order, -1.0, window), r)
            assert np.array_equal(P.homkermap(x, kernel, gamma, order, -1.0, window), r)
            x32 = x.astype(np.float32)
            r32 = R.homkermap(x32, kernel, gamma, order, -1.0, window)
            assert np.array_equal(P.homkermap(x32, kernel, gamma, order, -1.0, window), r32)
            assert np.array_equal(O.homkermap(x32, kernel, gamma, order, -1.0, window).astype(np.float32), r32)


@needs_ref
def test_getters_match_reference():
    R, P = reference(), va.lib()
    for lib in (R, P):
        h = lib.vl_homogeneouskernelmap_new(2, 0.5, 4, -1.0, 1)
        assert lib.vl_homogeneouskernelmap_get_order(h) == 4
        assert lib.vl_homogeneouskernelmap_get_dimension(h) == 9
        assert lib.vl_homogeneouskernelmap_get_kernel_type(h) == 2
        assert lib.vl_homogeneouskernelmap_get_window_type(h) == 1
        lib.vl_homogeneouskernelmap_delete(h)


@pytest.mark.parametrize("kernel,bound3,bound5", [(1, 1e-2, 4e-3), (2, 5e-3, 2e-3), (0, 2.0, 1.0)])
def test_map_approximates_the_kernel(kernel, bound3, bound5):
    # rectangular window, gamma 1: <psi(x), psi(y)> against the kernel itself
    # (the intersection kernel is not smooth at x = y, hence its looser bound)
    O = oracle()
    rng = np.random.default_rng(7)
    x, y = rng.uniform(0.01, 1.0, 500), rng.uniform(0.01, 1.0, 500)
    exact = KERNELS[kernel](x, y)
    errs = []
    for order in (1, 3, 5):
        px, py = O.homkermap(x, kernel, 1.0, order, -1.0, 1), O.homkermap(y, kernel, 1.0, order, -1.0, 1)
        errs.append(np.max(np.abs((px * py).sum(1) - exact) / exact))
    assert errs[0] > errs[1] > errs[2]
    assert errs[1] < bound3 and errs[2] < bound5
    # the map of a negative value is the negated map of its magnitude, zero maps to zero
    v = np.array([0.3, -0.3, 0.0])
    m = O.homkermap(v, kernel, 1.0, 3, -1.0, 1)
    assert np.array_equal(m[1], -m[0]) and not m[2].any()
    # a planted error (the map of chi2 used for JS) is caught by the same check
    if kernel == 2:
        px, py = O.homkermap(x, 1, 1.0, 5, -1.0, 1), O.homkermap(y, 1, 1.0, 5, -1.0, 1)
        assert np.max(np.abs((px * py).sum(1) - exact) / exact) > 10 * bound5
