"""Pins the oracle's grid fills, element-wise operators, data variance and
facies geomodel update (oracle/crava_oracle.c) with closed forms worked out
from the reference's definitions, not by restating its loops:

* fillInParamCorr without gradients is circCorrT[k] * corrXY[j,i] * scale
  exactly (the interpolation weight is 0); an integer gradient is an exact
  cyclic shift of circCorrT by the lateral lag; a fractional gradient
  interpolates a linear circCorrT exactly up to float rounding;
* fillInErrCorr puts corrXY at zero lag only; with a lateral shift s the
  weight at lag 0 is 1 - s and the rest goes to lag -1 (s > 0), or 1 + |s| and
  |s| at lag +1 (s < 0), as the reference's truncation toward zero gives;
* computeDataVariance on values whose squares and sums are exact in double;
* operators and geomodel against numpy float32 arithmetic (IEEE, no FMA).
"""
import numpy as np
import pytest

import oracle

RM = np.float32(-99999.0)


def _inputs(nxp, nyp, nzp, seed=0):
    rng = np.random.default_rng(seed)
    cxy = rng.uniform(0.1, 1.0, (nyp, nxp)).astype(np.float32)
    ct = rng.uniform(0.1, 1.0, nzp).astype(np.float32)
    return cxy, ct


def _lag(i, n):
    return i if i < n - i else i - n


@pytest.mark.parametrize("n", [(9, 6, 8), (16, 5, 12)])
def test_param_corr_no_gradient_is_separable_product(n):
    nxp, nyp, nzp = n
    cxy, ct = _inputs(*n)
    s = np.float32(0.37)
    g = oracle.fill_param_corr(nxp, nyp, nzp, cxy, ct, 0.0, 0.0, s)
    exp = (ct[:, None, None] * cxy[None, :, :]) * s
    assert np.array_equal(g[:, :, :nxp], exp.astype(np.float32))
    assert np.all(g[:, :, nxp:] == RM * s)


def test_param_corr_integer_gradient_is_cyclic_shift():
    nxp, nyp, nzp = 10, 7, 11
    cxy, ct = _inputs(nxp, nyp, nzp, 1)
    g = oracle.fill_param_corr(nxp, nyp, nzp, cxy, ct, 1.0, -2.0, 1.0)
    for k in range(nzp):
        for j in range(nyp):
            for i in range(nxp):
                kk = (k + _lag(i, nxp) * 1 + _lag(j, nyp) * -2) % nzp
                assert g[k, j, i] == np.float32(ct[kk] * cxy[j, i])


def test_param_corr_fractional_gradient_interpolates_linear_trend():
    nxp, nyp, nzp = 8, 6, 64
    cxy, _ = _inputs(nxp, nyp, nzp, 2)
    a, b = 0.5, 0.01
    ct = (a + b * np.arange(nzp)).astype(np.float32)
    gi, gj = 0.3, -0.45
    g = oracle.fill_param_corr(nxp, nyp, nzp, cxy, ct, gi, gj, 1.0)
    for k in range(8, nzp - 8):       # away from the cyclic wrap of the linear trend
        for j in range(nyp):
            for i in range(nxp):
                t = k + _lag(i, nxp) * gi + _lag(j, nyp) * gj
                assert g[k, j, i] == pytest.approx((a + b * t) * cxy[j, i], rel=2e-6)


def test_err_corr_zero_lag_only():
    nxp, nyp, nzp = 9, 5, 7
    cxy, _ = _inputs(nxp, nyp, nzp, 3)
    g = oracle.fill_err_corr(nxp, nyp, nzp, cxy, 0.0, 0.0)
    assert np.array_equal(g[0, :, :nxp], cxy)
    assert np.all(g[1:, :, :nxp] == 0.0)
    assert np.all(g[:, :, nxp:] == RM)


def test_err_corr_shift_weights():
    nxp, nyp, nzp = 6, 1, 9
    cxy = np.ones((nyp, nxp), dtype=np.float32)
    gi = 0.25
    g = oracle.fill_err_corr(nxp, nyp, nzp, cxy, gi, 0.0)
    for i in range(nxp):
        s = np.float32(_lag(i, nxp) * gi)
        col = g[:, 0, i]
        if s == 0:
            assert col[0] == 1.0 and np.all(col[1:] == 0)
        elif 0 < s < 1:
            assert col[0] == np.float32(1 - s) and col[nzp - 1] == np.float32(s)
            assert np.all(col[1:nzp - 1] == 0)
        elif -1 < s < 0:
            assert col[0] == np.float32(1 - s) and col[1] == np.float32(-s)
            assert np.all(col[2:] == 0)
        else:
            assert np.all(col == 0)


def test_data_variance_exact_sums():
    nx, ny, nz, nxp, nyp, nzp = 5, 4, 6, 8, 5, 9
    rnxp = 2 * (nxp // 2 + 1)
    rng = np.random.default_rng(4)
    g = (rng.integers(-8, 9, (nzp, nyp, rnxp)) * 0.25).astype(np.float32)   # squares exact, sums exact
    box = g[:nz, :ny, :nx].astype(np.float64)
    nz_mask = box != 0
    exp = np.float32((box[nz_mask] ** 2).sum() / nz_mask.sum())
    v, cnt = oracle.data_variance(g, nx, ny, nz, nxp, nyp, nzp)
    assert cnt == int(nz_mask.sum())
    assert np.float32(v) == exp
    v0, c0 = oracle.data_variance(np.zeros_like(g), nx, ny, nz, nxp, nyp, nzp)
    assert v0 == 0.0 and c0 == 0


@pytest.mark.parametrize("transformed", [False, True])
def test_grid_ops_match_float32_arithmetic(transformed):
    rng = np.random.default_rng(5)
    a0 = rng.standard_normal(2 * 60).astype(np.float32)
    b = rng.standard_normal(2 * 60).astype(np.float32)
    s = np.float32(1.7)
    ops = [0, 1, 2, 3, 4, 8] if transformed else [0, 1, 2, 3, 5, 6, 7, 8]
    for op in ops:
        a = a0.copy()
        oracle.grid_op(op, a, b, transformed, float(s))
        if transformed:
            x, y = a0.view(np.complex64), b.view(np.complex64)
            # textbook product with each float32 product and sum rounded once
            # (numpy's own complex64 product may fuse operations)
            xr, xi, yr, yi = x.real, x.imag, y.real, y.imag
            prod = (xr * yr - xi * yi) + 1j * (xi * yr + xr * yi).astype(np.float32)
            exp = {0: x + y, 1: x - y, 2: prod, 3: -x, 4: np.conj(x), 8: y}[op]
            assert np.array_equal(a.view(np.complex64), exp.astype(np.complex64)), op
        else:
            exp = {0: a0 + b, 1: a0 - b, 2: a0 * b, 3: -a0, 5: a0 * s, 6: a0 + s,
                   7: np.full_like(a0, s), 8: b}[op]
            assert np.array_equal(a, exp.astype(np.float32)), op


def test_facies_geomodel():
    nx, ny, nz, nfac = 5, 3, 4, 3
    rnx = 2 * (nx // 2 + 1)
    rng = np.random.default_rng(6)
    prob = [rng.uniform(0, 1, (nz, ny, rnx)).astype(np.float32) for _ in range(nfac + 1)]
    prior = np.array([0.2, 0.5, 0.3], dtype=np.float32)
    before = [p.copy() for p in prob]
    oracle.facies_geomodel(nx, ny, nz, prior, prob)
    for l in range(nfac):
        exp = prior[l] * before[nfac] + before[l]
        assert np.array_equal(prob[l][:, :, :nx], exp[:, :, :nx])
        assert np.all(prob[l][:, :, nx:] == RM)
    assert np.array_equal(prob[nfac], before[nfac])
