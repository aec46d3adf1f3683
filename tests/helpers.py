"""Shared helpers for the parity tests (CPU side only)."""
from __future__ import annotations

import numpy as np


def rel_l2(x, y) -> float:
    x = np.asarray(x)
    y = np.asarray(y)
    den = np.linalg.norm(y.ravel())
    num = np.linalg.norm((x - y).ravel())
    return float(num / den) if den > 0 else float(num)


def real_part(g: np.ndarray, nxp: int) -> np.ndarray:
    """Cells [.., :nxp] of a padded real grid."""
    return g[:, :, :nxp]


def seeded_real_grid(nxp, nyp, nzp, seed, dtype=np.float32) -> np.ndarray:
    rng = np.random.default_rng(seed)
    rnxp = 2 * (nxp // 2 + 1)
    g = np.zeros((nzp, nyp, rnxp), dtype=dtype)
    g[:, :, :nxp] = rng.standard_normal((nzp, nyp, nxp)).astype(dtype)
    return g


def to_complex(g: np.ndarray) -> np.ndarray:
    return g.view(np.complex64 if g.dtype == np.float32 else np.complex128)


def fourier_inputs(prob, which="oracle"):
    """Forward-transform the time-domain inputs of a synth.Problem on the CPU
    (oracle or reference FFT).  Returns float32 arrays in place-transformed form."""
    import oracle
    cfg = prob.cfg
    n = (cfg.nxp, cfg.nyp, cfg.nzp)
    mu = [g.cpu().numpy().copy() for g in prob.mu]
    d = [g.cpu().numpy().copy() for g in prob.data]
    cv = [g.cpu().numpy().copy() for g in prob.cov]
    ec = prob.errcorr.cpu().numpy().copy()
    for g in mu + d:
        oracle.fft_r2c(g, *n, False, which)
    for g in cv + [ec]:
        oracle.fft_r2c(g, *n, True, which)
    return mu, d, cv, ec


def as_f64(grids):
    return [np.ascontiguousarray(g.view(np.complex64).astype(np.complex128)).view(np.float64) for g in grids]
