"""The oracle's time-lapse merge (State4D::merge / mergeCov,
src/state4d.cpp:90-236) pinned by what the merge means: the current state
is x = x_static + x_dynamic, so its mean is the sum of the means and its
covariance is [I I] Sigma [I I]^H of the joint 6x6 prior Sigma, computed here
as a float64 matrix product of random Hermitian positive definite joints."""
import numpy as np

import oracle

PAIRS = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]


def _joint(rng, n):
    X = rng.standard_normal((n, 6, 6)) + 1j * rng.standard_normal((n, 6, 6))
    return X @ np.conj(np.swapaxes(X, 1, 2)) + 0.5 * np.eye(6)


def test_merge_cov_is_covariance_of_the_sum():
    rng = np.random.default_rng(3)
    n = 500
    S = _joint(rng, n)
    ss = [S[:, a, b] for a, b in PAIRS]
    dd = [S[:, 3 + a, 3 + b] for a, b in PAIRS]
    sd = [S[:, a, 3 + b] for a in range(3) for b in range(3)]
    got = oracle.merge_cov_4d(ss, dd, sd)
    T = np.concatenate([np.eye(3), np.eye(3)], axis=1)
    want = T @ S @ T.T
    for q, (a, b) in enumerate(PAIRS):
        assert np.allclose(got[q], want[:, a, b], rtol=1e-6, atol=1e-6 * np.abs(want[:, a, b]).max()), (a, b)


def test_merge_mean_is_the_sum():
    rng = np.random.default_rng(4)
    ms = [(rng.standard_normal(300) + 1j * rng.standard_normal(300)).astype(np.complex64) for _ in range(3)]
    md = [(rng.standard_normal(300) + 1j * rng.standard_normal(300)).astype(np.complex64) for _ in range(3)]
    got = oracle.merge_mean_4d(ms, md)
    for l in range(3):
        assert np.array_equal(got[l], ms[l] + md[l])
