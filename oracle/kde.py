"""CPU restatement of the facies density tables: the PosteriorElasticPDF3D
constructor (src/posteriorelasticpdf3d.cpp:12-138, SetupSmoothingGaussian3D
:360-413) and the per-density-index setup of
FaciesProb::MakePosteriorElasticPDF3D (src/faciesprob.cpp:926-1135,
CalculateVariances :2332-2377).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker, never by the
product.  The convolution is computed in float64 with numpy's FFT (the
reference runs FFTW 2.1.2 in float), the histogram, the smoother and the
variances with the reference's float roundings.
"""
from __future__ import annotations

import math

import numpy as np

IMISSING = -99999


def histogram(d1, d2, d3, n1, n2, n3, lim):
    """Counts of the points per lattice cell (float32), times float(1/npts);
    [n3][n2][n1].  Points outside the lattice are dropped
    (FFTGrid::setRealValue refuses them)."""
    npts = len(d1)
    h = np.zeros((n3, n2, n1), np.float32)
    dx, dy, dz = (lim[1] - lim[0]) / n1, (lim[3] - lim[2]) / n2, (lim[5] - lim[4]) / n3
    for p in range(npts):
        i = math.floor((d1[p] - lim[0]) / dx)
        j = math.floor((d2[p] - lim[2]) / dy)
        k = math.floor((d3[p] - lim[4]) / dz)
        if 0 <= i < n1 and 0 <= j < n2 and 0 <= k < n3:
            h[k, j, i] += np.float32(1.0)
    return h * np.float32(np.float32(1.0) / np.float32(npts)) if npts else h * np.float32(np.inf)


def smoother(sigma, n1, n2, n3, lim):
    """exp(-x^T sigma^-1 x / 2) at the wrapped lattice lags (float32), divided
    by its sequential float32 sum; [n3][n2][n1]."""
    si = np.linalg.inv(np.asarray(sigma, np.float64))
    dx, dy, dz = (lim[1] - lim[0]) / n1, (lim[3] - lim[2]) / n2, (lim[5] - lim[4]) / n3

    def lag(n):
        j = np.arange(n)
        return np.where(j <= n // 2, j, j - n).astype(np.float64)

    jj = lag(n1)[None, None, :] * dx
    kk = lag(n2)[None, :, None] * dy
    ll = lag(n3)[:, None, None] * dz
    q = (jj * jj * si[0, 0] + kk * kk * si[1, 1] + ll * ll * si[2, 2] + 2 * jj * kk * si[1, 0] +
         2 * jj * ll * si[2, 0] + 2 * kk * ll * si[2, 1])
    s = np.exp(-0.5 * q).astype(np.float32)
    total = np.cumsum(s.ravel(), dtype=np.float32)[-1]      # the reference's running float sum
    return (s / total).astype(np.float32)


def kde_table(d1, d2, d3, sigma, n1, n2, n3, lim):
    """The density table [n3][n2][n1] in float64: the circular convolution of
    the histogram with the smoother."""
    h = histogram(d1, d2, d3, n1, n2, n3, lim).astype(np.float64)
    s = smoother(sigma, n1, n2, n3, lim).astype(np.float64)
    return np.fft.irfftn(np.fft.rfftn(h) * np.fft.rfftn(s), s=h.shape, axes=(0, 1, 2))


def sample_variances(a, b, r, facies):
    """CalculateVariances: float means, float accumulators of double squares."""
    a, b, r = (np.asarray(x, np.float32) for x in (a, b, r))
    keep = [i for i in range(len(a)) if facies[i] != IMISSING]
    n = len(keep)
    m = [np.float32(0.0)] * 3
    for i in keep:
        m = [m[0] + a[i], m[1] + b[i], m[2] + r[i]]
    m = [np.float32(x / np.float32(n)) for x in m]
    v = [np.float32(0.0)] * 3
    for i in keep:
        for c, x in enumerate((a[i], b[i], r[i])):
            d = float(np.float32(x - m[c]))
            v[c] = np.float32(float(v[c]) + d * d)
    return [np.float32(x / np.float32(n - 1)) for x in v]


def density_setup(a, b, r, facies, sigma_e=None, no_vs=False):
    """One density index of MakePosteriorElasticPDF3D: (smoothing covariance
    3x3, limits [6])."""
    nobs = sum(1 for f in facies if f != IMISSING)
    var = sample_variances(a, b, r, facies)
    if no_vs:
        var[1] = np.float32(5 * var[0])
    hopt = np.float32((4.0 / 7) ** (1.0 / 7) * float(nobs) ** (-1.0 / 7))
    kstd = [np.float32(hopt * np.float32(math.sqrt(float(v)))) for v in var]
    se = np.zeros((3, 3))
    if sigma_e is not None:
        s = np.asarray(sigma_e, np.float64).reshape(3, 3)
        if not no_vs:
            for j in range(3):
                for k in range(j + 1):
                    se[k, j] = se[j, k] = s[k, j]
        else:
            se[0, 0] = s[0, 0]
            se[0, 2] = se[2, 0] = s[0, 1]
            se[2, 2] = s[1, 1]
    for d in range(3):
        k2 = float(np.float32(kstd[d] * kstd[d]))
        if se[d, d] < k2:
            se[d, d] = k2
    lim = []
    for d, x in enumerate((a, b, r)):
        x = np.asarray(x, np.float32)
        lim.append(float(x.min()) - 5.0 * math.sqrt(se[d, d]))
        lim.append(float(x.max()) + 5.0 * math.sqrt(se[d, d]))
    return se, lim
