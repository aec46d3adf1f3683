"""Oracle of the local-noise correction: Crava::computeG,
Crava::newPosteriorCovPointwise, Crava::correctAlphaBetaRho and the
realisation transform of Crava::simulate (src/crava.cpp:2130-2470,
:1547-1566), restated in numpy, trace by trace.

TEST INFRASTRUCTURE ONLY: the checker.  The product never imports it.

Eigen-decompositions.  NRLib::ComputeEigenVectors
(libs/nrlib/flens/nrlib_flens.cpp:114-121) calls LAPACK's dgeev through
FLENS (right eigenvectors, real parts of the eigenvalues).  LAPACK is a
third-party dependency that the reference tree does not hold; here it is
netlib's dgeev as scipy ships it (scipy 1.15.3, OpenBLAS 0.3.28 with the
reference LAPACK), called directly, so the oracle's eigenvectors carry the
signs that dgeev gives.  lib_matr_eigen (libs/lib/lib_matr.c:921-1000,
Householder tridiagonalisation + QL) only enters through a matrix function
(V f(L) V^T), which does not depend on the eigenvectors' signs; it is
restated with numpy's symmetric solver.

The products are FLENS (BLAS dgemm) products in the reference, whose
summation order is the BLAS library's; numpy's @ stands for them.
"""
from __future__ import annotations

import numpy as np
from scipy.linalg import lapack

RMISSING = -99999.000  # src/definitions.h:35


def eigen_vectors(m: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """NRLib::ComputeEigenVectors: (real parts of the eigenvalues, right
    eigenvectors as columns), LAPACK dgeev order."""
    a = np.array(m, dtype=np.float64, order="F")
    wr, wi, vl, vr, info = lapack.dgeev(a, compute_vl=0, compute_vr=1)
    if info != 0:
        raise RuntimeError(f"dgeev info {info}")
    return np.asarray(wr, dtype=np.float64), np.asarray(vr, dtype=np.float64)


def sort3x3(ev: np.ndarray, vec: np.ndarray) -> tuple[np.ndarray, np.ndarray]:
    """NRLib::Sort3x3 (libs/nrlib/flens/nrlib_flens.cpp:211-246): eigenvalues
    in descending order (first maximum wins a tie), vectors permuted along."""
    index = [-1, -1, -1]
    h1 = np.zeros(3)
    h2 = np.zeros((3, 3))
    for j in range(3):
        mx = -100000.0
        for i in range(3):
            if j == 0 or (j == 1 and i != index[0]) or (j == 2 and i != index[0] and i != index[1]):
                if ev[i] > mx:
                    mx = ev[i]
                    index[j] = i
        h1[j] = mx
        h2[:, j] = vec[:, index[j]]
    return h1, h2


def _fn(m: np.ndarray, f, keep) -> np.ndarray:
    """V diag(f(l) where keep(l), else 0) V^T with dgeev's V (the pattern of
    every ComputeEigenVectors call in src/crava.cpp:2150-2310)."""
    ev, v = eigen_vectors(m)
    d = np.zeros((len(ev), len(ev)))
    for i, l in enumerate(ev):
        if keep(l):
            d[i, i] = f(l)
    return (v @ d) @ v.T


def compute_g(err_theta_cov: np.ndarray, prior_var0: np.ndarray, post_var0: np.ndarray,
              resign: bool = False) -> np.ndarray:
    """Crava::computeG, src/crava.cpp:2130-2211.  Returns G [ntheta][3].

    resign=True re-signs dgeev's eigenvectors of A so that each one's
    component of largest magnitude is positive (the host helper's rule), to
    compare that helper's algebra with this restatement."""
    nt = err_theta_cov.shape[0]
    sm = np.array(prior_var0, dtype=np.float64)
    sdelta = sm - np.array(post_var0, dtype=np.float64)
    sdelta = _fn(sdelta, lambda l: l, lambda l: l > 0.0)
    sinv = _fn(sm, lambda l: 1.0 / np.sqrt(l), lambda l: l > 0.0000001)
    a = (sinv @ sdelta) @ sinv
    ev, vec = eigen_vectors(a)
    ev, vec = sort3x3(ev, vec)
    if resign:
        for k in range(3):
            if vec[np.argmax(np.abs(vec[:, k])), k] < 0.0:
                vec[:, k] = -vec[:, k]
    lg = np.zeros((nt, 3))
    with np.errstate(invalid="ignore"):
        for i in range(min(nt, 3)):
            lg[i, i] = np.sqrt(ev[i] / (1.0 - ev[i]))
    ece = _fn(err_theta_cov, np.sqrt, lambda l: l > 0.0)
    return ((ece @ lg) @ vec.T) @ sinv


def robust_inv_sqrt_delta(prior_var0, post_var0) -> np.ndarray:
    """The trace-independent factor of newPosteriorCovPointwise
    (src/crava.cpp:2289-2307): (Sm - Smd)^(-1/2), eigenvalues <= 1e-7
    dropped."""
    sd = np.array(prior_var0, dtype=np.float64) - np.array(post_var0, dtype=np.float64)
    return _fn(sd, lambda l: 1.0 / np.sqrt(l), lambda l: l > 0.0000001)


def new_posterior_cov_pointwise(g, scales, err_theta_cov, prior_var0, post_var0):
    """Crava::newPosteriorCovPointwise, src/crava.cpp:2213-2312.

    Returns (sigmanew, sigmamdnew): sqrt(Sm - Smd_new) (Sm - Smd)^(-1/2) and
    sqrt(Smd_new), Smd_new = Sm - Sm G^T (G Sm G^T + D E D)^+ G Sm,
    D = diag(sqrt(scales))."""
    nt = err_theta_cov.shape[0]
    d = np.diag(np.sqrt(np.asarray(scales, dtype=np.float64)))
    sigmaenew = (d @ err_theta_cov) @ d
    sm = np.array(prior_var0, dtype=np.float64)
    help_ = ((g @ sm) @ g.T) + sigmaenew
    pinv = _fn(help_, lambda l: 1.0 / l, lambda l: l > 0.00000001)
    deltanew = (((sm @ g.T) @ pinv) @ g) @ sm
    smdnew = _fn(sm - deltanew, np.sqrt, lambda l: l > 0.0)
    sqdelta = _fn(deltanew, np.sqrt, lambda l: l > 0.0)
    sigmanew = sqdelta @ robust_inv_sqrt_delta(prior_var0, post_var0)
    assert sigmaenew.shape == (nt, nt)
    return sigmanew, smdnew


def inv_sqrt_post(post_var0) -> np.ndarray:
    """sigmamdold of correctAlphaBetaRho (src/crava.cpp:2376-2385):
    lib_matr_eigen of Smd, eigenvalues > 0 mapped to 1/sqrt."""
    w, v = np.linalg.eigh(np.array(post_var0, dtype=np.float64))
    d = np.diag([1.0 / np.sqrt(x) if x > 0.0 else 0.0 for x in w])
    return (v @ d) @ v.T


def min_scale(table: np.ndarray) -> float:
    """Grid2D::FindMin(RMISSING) (libs/nrlib/grid/grid2d.hpp:182-191)."""
    t = np.asarray(table, dtype=np.float64).reshape(-1)
    ok = t[t != RMISSING]
    return float(ok.min()) if ok.size else RMISSING


def trace_matrices(g, scale_tables, err_theta_cov, prior_var0, post_var0):
    """Per-trace sigmanew and (*sigmamdnew_)(i,j) of correctAlphaBetaRho
    (src/crava.cpp:2411-2438).  scale_tables: ntheta arrays [ny][nx] of
    getLocalNoiseScale(angle)(i, j).  Returns two arrays [ny][nx][3][3]."""
    nt = len(scale_tables)
    ny, nx = scale_tables[0].shape
    mins = [min_scale(t) for t in scale_tables]
    old = inv_sqrt_post(post_var0)
    sn = np.zeros((ny, nx, 3, 3))
    tc = np.zeros((ny, nx, 3, 3))
    for j in range(ny):
        for i in range(nx):
            sc = [float(scale_tables[a][j, i]) / mins[a] for a in range(nt)]
            s, smd = new_posterior_cov_pointwise(g, sc, err_theta_cov, prior_var0, post_var0)
            sn[j, i] = s
            # lib_matr_prod, libs/lib/lib_matr.c:1190-1225 (left to right)
            for r in range(3):
                for c in range(3):
                    x = 0.0
                    for k in range(3):
                        x = x + smd[r, k] * old[k, c]
                    tc[j, i, r, c] = x
    return sn, tc


def apply_correction(post, mean2, sigmanew, nx, ny, nz):
    """The trace loop of correctAlphaBetaRho (src/crava.cpp:2447-2454) on
    real padded grids [nzp][nyp][rnxp] (float32), in place: float
    differences, then a double sum from the left, rounded to float."""
    for j in range(ny):
        for i in range(nx):
            s = sigmanew[j, i]
            d = [(post[c][:nz, j, i] - mean2[c][:nz, j, i]).astype(np.float32).astype(np.float64) for c in range(3)]
            m = [mean2[c][:nz, j, i].astype(np.float64) for c in range(3)]
            for r in range(3):
                v = m[r] + s[r, 0] * d[0]
                v = v + s[r, 1] * d[1]
                v = v + s[r, 2] * d[2]
                post[r][:nz, j, i] = v.astype(np.float32)


def transform_realisation(seed, trace_cov, nx, ny, nz):
    """The local-noise block of Crava::simulate (src/crava.cpp:1547-1566) on
    real padded grids, in place: double sum from the left of the float
    draws, rounded to float."""
    for j in range(ny):
        for i in range(nx):
            m = trace_cov[j, i]
            x = [seed[c][:nz, j, i].astype(np.float64) for c in range(3)]
            out = []
            for r in range(3):
                v = m[r, 0] * x[0]
                v = v + m[r, 1] * x[1]
                v = v + m[r, 2] * x[2]
                out.append(v.astype(np.float32))
            for r in range(3):
                seed[r][:nz, j, i] = out[r]
