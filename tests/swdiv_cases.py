"""Inputs of the wavelet-division and interpolation tests (shared by the
oracle and GPU tests)."""
import numpy as np


def swdiv_inputs(nx, ny, nz, nxp, nyp, nzp, nt, seed, local=True, rel_range=(0.6, 1.0), wavelet="ricker"):
    """Returns (positional args after the sizes, keyword tables) of
    oracle.divide_by_scale_wavelet / core.divide_by_scale_wavelet."""
    rng = np.random.default_rng(seed)
    rnxp = 2 * (nxp // 2 + 1)
    grids = [rng.standard_normal((nzp, nyp, rnxp)).astype(np.float32) for _ in range(nt)]
    t = np.arange(nzp)
    t = np.minimum(t, nzp - t).astype(np.float64)
    if wavelet == "ricker":
        wav = np.stack([np.exp(-(t / 2.5) ** 2) * np.cos((0.5 + 0.1 * l) * t) for l in range(nt)])
    else:   # delta + 0.3 delta(t - 2): |W|^2 in [0.49, 1.69]
        wav = np.zeros((nt, nzp))
        wav[:, 0] = 1.0
        wav[:, 2] = 0.3
    wav = wav.astype(np.float32)
    wnorm = np.sqrt((wav.astype(np.float64) ** 2).sum(1)).astype(np.float32)
    A = rng.standard_normal((nt, 3)).astype(np.float32)
    L0 = np.array([[0.08, 0, 0], [0.06, 0.09, 0], [0.02, 0.01, 0.03]])
    var0 = L0 @ L0.T
    ct = np.where(t < nz, np.exp(-(t / 4.0) ** 2), 0.0).astype(np.float32)
    ev = (rng.uniform(0.5, 2.0, nt) * 1e-3).astype(np.float32)
    rel = rng.uniform(*rel_range, (nyp, nxp)) if rel_range else np.ones((nyp, nxp))
    kw = {}
    if local:
        kw["shift"] = [rng.uniform(-6, 6, (ny, nx)).astype(np.float32) if l % 2 == 0 else None for l in range(nt)]
        kw["gain"] = [rng.uniform(0.7, 1.3, (ny, nx)).astype(np.float32) if l != 1 else None for l in range(nt)]
    args = [grids, wav, wnorm, A, var0, ct, ev, rel, 4.0, 100.0, 5.0, 55.0, 0.1]
    return args, kw


def isq_grid(nx, ny, nz, nxp, nyp, nzp, seed, n_bad, n_weak=0, cluster=False):
    """A real padded grid whose volume has n_bad zero traces and n_weak traces
    scaled by 1e-3 (scattered, or in one block when cluster)."""
    rng = np.random.default_rng(seed)
    rnxp = 2 * (nxp // 2 + 1)
    g = rng.standard_normal((nzp, nyp, rnxp)).astype(np.float32)
    idx = rng.permutation(nx * ny)
    if cluster:
        i0, j0 = nx // 3, ny // 3
        w = max(1, int(np.sqrt(n_bad)))
        bad = [(j0 + a) * nx + i0 + b for a in range(w) for b in range(w) if j0 + a < ny and i0 + b < nx][:n_bad]
        idx = np.array(bad + [t for t in idx if t not in bad])
    for t in idx[:n_bad]:
        g[:, t // nx, t % nx] = 0.0
    for t in idx[n_bad:n_bad + n_weak]:
        g[:, t // nx, t % nx] *= np.float32(1e-3)
    return g
