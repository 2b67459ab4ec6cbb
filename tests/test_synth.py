"""The seeded C generator of the benchmark workloads (libvlsynth,
include/vlsynth.h).

Pinned independently of its own code: the SplitMix64 core against the
published first output of the generator from state 0, the per-row stream
start against a Python restatement of the header's definition, the
logarithm and exponential against libm, the polar-method normals by their
moments, the projection by orthonormality and by the projected rows
equalling the projection applied to the unprojected rows; and the
documented properties: bits depend only on (seed, row) (chunks equal the
whole), SIFT rows rectified, unit norm, clipped before renormalisation,
mixture parameters in their ranges, counts clamped with the median near 3000.
The numpy generator (other draws, same distributions) is compared by
distribution only."""
import math

import numpy as np
import pytest

from vlfeat_amd import synth

M64 = (1 << 64) - 1


def mix64(z):
    z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M64
    z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M64
    return z ^ (z >> 31)


def stream_uniforms(seed, stream, row, k):
    s = mix64(seed ^ mix64(((stream * 0x9E3779B97F4A7C15) & M64) ^ mix64(row + 1)))
    out = []
    for _ in range(k):
        s = (s + 0x9E3779B97F4A7C15) & M64
        out.append((mix64(s) >> 11) * 2.0 ** -53)
    return out


def test_splitmix64_core_and_streams():
    # SplitMix64 from state 0: first output 0xe220a8397b1dcdaf (the published reference value)
    assert mix64(0x9E3779B97F4A7C15) == 0xE220A8397B1DCDAF
    L = synth._synth()
    for seed, stream, row in [(0, 0, 0), (1, 1, 12345), (7, 5, 2**40 + 3)]:
        want = stream_uniforms(seed, stream, row, 4)
        got = [L.vlsynth_uniform(seed, stream, row, k) for k in range(4)]
        assert got == want


def test_log_exp_against_libm():
    L = synth._synth()
    rng = np.random.default_rng(0)
    xs = np.concatenate([10.0 ** rng.uniform(-300, 300, 2000), rng.uniform(0.5, 2.0, 2000), [1.0, 2.0, 0.5]])
    for x in xs:
        a, b = L.vlsynth_log(float(x)), math.log(x)
        assert abs(a - b) <= 4 * math.ulp(b) + (0.0 if b else 1e-300), (x, a, b)
    for x in np.concatenate([rng.uniform(-700, 700, 2000), rng.uniform(-1, 1, 2000), [0.0]]):
        a, b = L.vlsynth_exp(float(x)), math.exp(x)
        assert abs(a - b) <= 4 * math.ulp(b), (x, a, b)
    assert L.vlsynth_log(0.0) == -math.inf and math.isnan(L.vlsynth_log(-1.0))
    assert L.vlsynth_exp(800.0) == math.inf and L.vlsynth_exp(-800.0) == 0.0


def test_bits_depend_only_on_seed_and_row():
    a = synth.c_sift_like(5000, 80, seed=1)
    b = synth.c_sift_like(5000, 80, seed=1)
    assert np.array_equal(a, b)
    parts = [synth.c_sift_like(n, 80, seed=1, row_offset=o) for o, n in [(0, 1234), (1234, 3000), (4234, 766)]]
    assert np.array_equal(np.concatenate(parts), a)
    assert not np.array_equal(synth.c_sift_like(5000, 80, seed=2), a)
    x1, *_ = synth.c_mixture(3000, 24, 10, seed=3)
    x2, *_ = synth.c_mixture(1000, 24, 10, seed=3, row_offset=2000)
    assert np.array_equal(x1[2000:], x2)


def test_sift_like_rows_and_projection():
    z = synth.c_sift_like(20000, 128, seed=5).astype(np.float64)
    assert z.min() >= 0.0
    assert np.abs(np.linalg.norm(z, axis=1) - 1.0).max() < 1e-6
    assert 0.45 < np.mean(z == 0) < 0.55                      # rectified Gaussian
    P = synth.c_projection(80, 128, 5)
    assert np.abs(P @ P.T - np.eye(80)).max() < 1e-13
    y = synth.c_sift_like(20000, 80, seed=5).astype(np.float64)
    assert np.abs(y - z @ P.T).max() < 1e-6                   # the projection of the same rows
    # the numpy generator: same distribution of the unprojected rows
    zn = synth.sift_like(20000, 128, seed=5, project=False).astype(np.float64)
    for f in (lambda v: v.mean(), lambda v: np.mean(v == 0), lambda v: v.max(axis=1).mean()):
        assert abs(f(z) - f(zn)) < 0.02 * abs(f(zn)) + 1e-3


def test_mixture_parameters_and_normals():
    x, mu, sg, w = synth.c_mixture(400000, 8, 1, seed=11)
    assert w[0] == 1.0 and np.all(np.abs(mu) <= 1) and np.all((sg >= 0.1) & (sg <= 0.5))
    z = (x.astype(np.float64) - mu[0]) / sg[0]
    assert abs(z.mean()) < 5e-3 and abs(z.var() - 1.0) < 5e-3
    assert abs(np.mean(z ** 4) - 3.0) < 0.05                   # Gaussian kurtosis
    assert abs(np.mean(z > 1.0) - 0.158655) < 3e-3
    _, mu, sg, w = synth.c_mixture(0, 64, 64, seed=0)
    assert abs(w.sum() - 1.0) < 1e-12 and w.min() > 0


def test_lognormal_counts():
    c = synth.c_lognormal_counts(4096)
    assert c.min() >= 200 and c.max() <= 20000
    assert abs(np.median(c) - 3000) < 150
    z = (np.log(np.clip(c, 201, 19999)) - math.log(3000)) / 0.5
    assert abs(z.std() - 1.0) < 0.05
    assert np.array_equal(c, synth.c_lognormal_counts(4096))
