"""Seeded synthetic inversion inputs (no field data ships with the reference).

Builds, from one seed, everything the posterior path consumes, following the
formulas of the reference where it defines them:

* prior covariance Sigma0 from standard deviations and correlations;
* lateral correlation corrXY on the padded grid (ModelGeneral::findCorrXYGrid,
  src/modelgeneral.cpp:4073-4108) with a spherical variogram
  (SphericalVario::corr, src/vario.cpp:66-73);
* circular temporal correlation (SeismicParametersHolder::computeCircCorrT /
  makeCircCorrTPosDef, src/seismicparametersholder.cpp:431-499) with the
  low-cut integer of src/modelgeneral.cpp:3961;
* the six prior covariance grids (FFTGrid::fillInParamCorr with zero
  gradients, src/fftgrid.cpp:1119-1164, times Sigma0_ab,
  src/seismicparametersholder.cpp:214-219) and the error-correlation grid
  (FFTGrid::fillInErrCorr, src/fftgrid.cpp:1065-1116);
* Ricker wavelets in FFT order (Wavelet::Ricker, src/wavelet.cpp:1354-1358;
  Wavelet::resample ordering and norm, :416-463; Wavelet::scale, :477-486);
* the PP reflection matrix (ModelAVODynamic::setupDefaultReflectionMatrix,
  src/modelavodynamic.cpp:876-928);
* angular noise correlation Sigma_e,lm = s_l s_m exp(-3 |dtheta| / 10 deg);
* background trends, a Gaussian perturbation drawn from the prior and seismic
  d = K m + white noise at the requested S/N.

Arrays are torch tensors on the requested device (CPU for tests, the GPU for
large benchmarks).  Grids use the FFTGrid real layout [nzp][nyp][rnxp].
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np
import torch

RMISSING = -99999.0


@dataclass
class Config:
    name: str
    nx: int
    ny: int
    nz: int
    nxp: int
    nyp: int
    nzp: int
    thetas_deg: list
    peak_hz: list
    wavelet_scale: list
    snr: list
    dx: float = 25.0
    dy: float = 25.0
    dt: float = 4.0
    lateral_range: float = 800.0
    vertical_range_ms: float = 40.0
    std: tuple = (0.08, 0.12, 0.04)
    corr: tuple = (0.7, 0.5, 0.4)   # (ab, ar, br)
    wnc: float = 0.1
    low_cut: float = 5.0
    high_cut: float = 55.0
    ang_range_deg: float = 10.0
    seed: int = 1


def find_closest_factorable(n: int) -> int:
    """Smallest 2^a 3^b 5^c 7^d >= n (FFTGrid::findClosestFactorableNumber,
    src/fftgrid.cpp:2837-2866)."""
    m = max(1, n)
    while True:
        k = m
        for p in (2, 3, 5, 7):
            while k % p == 0:
                k //= p
        if k == 1:
            return m
        m += 1


def config(name: str, seed: int = 1) -> Config:
    rng = np.random.default_rng(seed + 1000)
    if name == "plain_cpu":      # 64x64x128 padded 80x80x160, 3 angles
        return Config(name, 64, 64, 128, 80, 80, 160, [10.0, 20.0, 30.0], [30.0] * 3, [1.0] * 3,
                      [5.0] * 3, seed=seed)
    if name == "plain_gpu":      # 256^3 padded 288^3, 5 angles
        th = list(np.linspace(5.0, 35.0, 5))
        return Config(name, 256, 256, 256, 288, 288, 288, th, [25.0] * 5,
                      list(rng.uniform(0.8, 1.2, 5)), list(rng.uniform(3.0, 10.0, 5)), seed=seed)
    if name == "precision":      # 512x512x320 FFT grid (interior 448x448x256), 6 angles
        # per-angle S/N from 1e4 down to 20: low noise makes M ill-conditioned
        th = list(np.linspace(4.0, 40.0, 6))
        return Config(name, 448, 448, 256, 512, 512, 320, th, [25.0] * 6,
                      list(rng.uniform(0.8, 1.2, 6)), [1e4, 3e3, 1e3, 300.0, 100.0, 20.0], seed=seed)
    if name == "differentiation":  # 256x256x192, 5 angles
        th = list(np.linspace(5.0, 35.0, 5))
        nxp, nzp = find_closest_factorable(int(math.ceil(256 * 1.1))), find_closest_factorable(int(math.ceil(192 * 1.1)))
        return Config(name, 256, 256, 192, nxp, nxp, nzp, th, [25.0] * 5,
                      list(rng.uniform(0.8, 1.2, 5)), list(rng.uniform(3.0, 10.0, 5)), seed=seed)
    if name == "tiny":           # brute-force sized
        return Config(name, 4, 4, 8, 4, 4, 8, [10.0, 30.0], [60.0] * 2, [1.0] * 2, [4.0] * 2,
                      lateral_range=60.0, vertical_range_ms=12.0, low_cut=0.0, high_cut=1e9, seed=seed)
    if name == "small":          # quick parity size, odd nxp on purpose
        return Config(name, 12, 10, 20, 15, 12, 24, [8.0, 18.0, 28.0, 38.0], [35.0] * 4, [1.0, 0.9, 1.1, 1.05],
                      [6.0, 5.0, 4.0, 3.0], low_cut=5.0, high_cut=100.0, seed=seed)
    raise KeyError(name)


def random_config(rng: np.random.Generator, idx: int, seed: int = 1, nxy=(32, 192), nzr=(64, 256)) -> Config:
    """One volume of the batching workload: nx, ny in [32,192], nz in
    [64,256] (or the given inclusive ranges), ntheta in [3,8], Ricker
    20-40 Hz, ranges 300-1500 m / 20-60 ms, S/N 2-10; padding as
    ModelGeneral::setPaddingSize with 10%."""
    nx, ny = int(rng.integers(nxy[0], nxy[1] + 1)), int(rng.integers(nxy[0], nxy[1] + 1))
    nz = int(rng.integers(nzr[0], nzr[1] + 1))
    nt = int(rng.integers(3, 9))
    nxp = find_closest_factorable(int(math.ceil(nx * 1.1)))
    nyp = find_closest_factorable(int(math.ceil(ny * 1.1)))
    nzp = find_closest_factorable(int(math.ceil(nz * 1.1)))
    th = list(np.sort(rng.uniform(2.0, 45.0, nt)))
    f = float(rng.uniform(20.0, 40.0))
    return Config(f"batch{idx}", nx, ny, nz, nxp, nyp, nzp, th, [f] * nt, list(rng.uniform(0.8, 1.2, nt)),
                  list(rng.uniform(2.0, 10.0, nt)), lateral_range=float(rng.uniform(300, 1500)),
                  vertical_range_ms=float(rng.uniform(20, 60)), seed=seed + idx)


@dataclass
class Problem:
    cfg: Config
    sigma0: np.ndarray          # (3,3) float64
    A: np.ndarray               # (nt,3) float32
    err_theta_cov: np.ndarray   # (nt,nt) float64
    wavelets: np.ndarray        # (nt,nzp) float32, FFT order, scaled
    wnorm: np.ndarray           # (nt,) float32
    corr_xy: np.ndarray         # (nyp,nxp) float32
    circ_corr_t: np.ndarray     # (nzp,) float32
    mu: list = field(default_factory=list)      # 3 real grids [nzp][nyp][rnxp] float32
    data: list = field(default_factory=list)    # nt grids
    cov: list = field(default_factory=list)     # 6 grids (aa, bb, rr, ab, ar, br)
    errcorr: torch.Tensor | None = None
    corr: torch.Tensor | None = None            # the correlation grid: cov[q] = corr * Sigma0_q
    true_m: list = field(default_factory=list)

    @property
    def ntheta(self):
        return len(self.cfg.thetas_deg)

    @property
    def lz(self) -> float:
        return float(self.cfg.nz * self.cfg.dt)


def sigma0_of(cfg: Config) -> np.ndarray:
    s = np.array(cfg.std, dtype=np.float64)
    rab, rar, rbr = cfg.corr
    R = np.array([[1.0, rab, rar], [rab, 1.0, rbr], [rar, rbr, 1.0]])
    return (s[:, None] * R * s[None, :]).astype(np.float64)


def reflection_matrix(thetas_deg, vsvp: float = 0.5) -> np.ndarray:
    A = np.zeros((len(thetas_deg), 3), dtype=np.float32)
    v2 = vsvp * vsvp
    for l, th in enumerate(thetas_deg):
        a = math.radians(float(th))
        s2 = math.sin(a) ** 2
        t2 = math.tan(a) * math.tan(a)
        A[l] = [np.float32((1.0 + t2) / 2.0), np.float32(-4 * v2 * s2), np.float32((1.0 - 4.0 * v2 * s2) / 2.0)]
    return A


def spherical(d: np.ndarray) -> np.ndarray:
    d = d.astype(np.float32)
    out = (1.0 + d.astype(np.float64) ** 3 / 2.0 - 1.5 * d.astype(np.float64)).astype(np.float32)
    return np.where(d >= 1.0, np.float32(0.0), out).astype(np.float32)


def corr_xy_grid(cfg: Config) -> np.ndarray:
    i = np.arange(cfg.nxp)
    j = np.arange(cfg.nyp)
    ri = np.where(i < cfg.nxp // 2 + 1, i, i - cfg.nxp).astype(np.float32)
    rj = np.where(j < cfg.nyp // 2 + 1, j, j - cfg.nyp).astype(np.float32)
    dx = (ri[None, :] * np.float32(cfg.dx)) / np.float32(cfg.lateral_range)
    dy = (rj[:, None] * np.float32(cfg.dy)) / np.float32(cfg.lateral_range)
    d = np.sqrt(dx * dx + dy * dy).astype(np.float32)
    return spherical(d)


def circ_corr_t(cfg: Config) -> np.ndarray:
    nzp = cfg.nzp
    lags = np.arange(nzp)
    refk = np.where(lags < nzp // 2 + 1, lags, nzp - lags)
    corr_t = spherical((refk * cfg.dt / cfg.vertical_range_ms).astype(np.float32))
    low_int_cut = int(math.floor(cfg.low_cut * (nzp * 0.001 * cfg.dt)))
    F = np.fft.rfft(corr_t.astype(np.float64))
    mag = np.abs(F)
    mag[: low_int_cut + 1] = 0.0
    c = np.fft.irfft(mag, n=nzp) * nzp
    if c[0] <= 1e-5:
        raise ValueError("circular temporal correlation undefined (increase nz)")
    return (c * (1.0 / c[0])).astype(np.float32)


def ricker_fft_order(nzp: int, dt: float, peak: float) -> np.ndarray:
    k = np.arange(nzp)
    t = np.where(k < nzp // 2 + 1, dt * k, dt * (k - nzp)).astype(np.float64)
    c = math.pi * math.pi * peak * peak * t * t * 1e-6
    return ((1 - 2 * c) * np.exp(-c)).astype(np.float32)


def kwd_table(wavelets: np.ndarray, nzp: int) -> np.ndarray:
    """kD(k) conj(W_l(k)), complex128 [nzp][nt] (generator use only)."""
    k = np.arange(nzp)
    kD = -1.0 + np.exp(2j * np.pi * k / nzp)
    W = np.fft.fft(wavelets.astype(np.float64), axis=1)  # [nt][nzp]
    return (kD[:, None] * np.conj(W.T))


def _padded(real_zyx: torch.Tensor, rnxp: int, fill: float) -> torch.Tensor:
    nzp, nyp, nxp = real_zyx.shape
    out = torch.full((nzp, nyp, rnxp), fill, dtype=torch.float32, device=real_zyx.device)
    out[:, :, :nxp] = real_zyx.to(torch.float32)
    return out


def make_problem(cfg: Config, device: str = "cpu", grids: bool = True) -> Problem:
    nt = len(cfg.thetas_deg)
    nxp, nyp, nzp = cfg.nxp, cfg.nyp, cfg.nzp
    rnxp = 2 * (nxp // 2 + 1)
    sig0 = sigma0_of(cfg)
    A = reflection_matrix(cfg.thetas_deg)
    cxy = corr_xy_grid(cfg)
    cct = circ_corr_t(cfg)
    wav = np.zeros((nt, nzp), dtype=np.float32)
    wnorm = np.zeros(nt, dtype=np.float32)
    for l in range(nt):
        w = ricker_fft_order(nzp, cfg.dt, float(cfg.peak_hz[l]))
        wnorm[l] = np.float32(math.sqrt(float(np.sum((w * w).astype(np.float64)))))
        wav[l] = (w * np.float32(cfg.wavelet_scale[l])).astype(np.float32)
    th = np.array(cfg.thetas_deg, dtype=np.float64)
    ang = np.exp(-3.0 * np.abs(th[:, None] - th[None, :]) / cfg.ang_range_deg)
    prob = Problem(cfg, sig0, A, np.zeros((nt, nt)), wav, wnorm, cxy, cct)
    if not grids:
        prob.err_theta_cov = ang * 1e-4
        return prob

    gen = torch.Generator(device=device)
    gen.manual_seed(int(cfg.seed))
    dev = torch.device(device)
    f64 = torch.float64
    base = torch.from_numpy(cct.astype(np.float64)).to(dev)[:, None, None] * \
        torch.from_numpy(cxy.astype(np.float64)).to(dev)[None, :, :]
    # covariance grids: (circCorrT x corrXY) in float, times Sigma0_ab in float
    base32 = (torch.from_numpy(cct).to(dev)[:, None, None] * torch.from_numpy(cxy).to(dev)[None, :, :])
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    prob.corr = _padded(base32, rnxp, RMISSING)
    prob.cov = []
    for (a, b) in pairs:
        s = np.float32(sig0[a, b])
        prob.cov.append(prob.corr * float(s))
    ec = torch.zeros((nzp, nyp, nxp), dtype=torch.float32, device=dev)
    ec[0] = torch.from_numpy(cxy).to(dev)
    prob.errcorr = _padded(ec, rnxp, RMISSING)

    # background trends (constant laterally), cosine return through the z padding
    k = torch.arange(nzp, device=dev, dtype=f64)
    nz = cfg.nz
    frac = torch.clamp(k / max(1, nz - 1), max=1.0)
    back = torch.where(k < nz, frac, 0.5 * (1.0 + torch.cos(math.pi * (k - nz + 1) / (nzp - nz + 1))))
    vp = 2500.0 + 1000.0 * back
    rho = 2.2 + 0.3 * back
    trends = [torch.log(vp), torch.log(0.5 * vp), torch.log(rho)]
    mu = [t[:, None, None].expand(nzp, nyp, nxp).contiguous() for t in trends]

    # Gaussian perturbation drawn from the prior: sqrt(eig) * chol(Sigma0) * white
    lam = torch.fft.rfftn(base, dim=(0, 1, 2)).real.abs()
    L0 = torch.from_numpy(np.linalg.cholesky(sig0)).to(dev)
    z = [torch.randn((nzp, nyp, nxp), generator=gen, device=dev, dtype=f64) for _ in range(3)]
    zh = [torch.fft.rfftn(zz, dim=(0, 1, 2), norm="ortho") for zz in z]
    del z
    sq = torch.sqrt(lam)
    m_true = []
    for a in range(3):
        acc = L0[a, 0] * zh[0]
        for b in range(1, a + 1):
            acc = acc + L0[a, b] * zh[b]
        dm = torch.fft.irfftn(acc * sq, s=(nzp, nyp, nxp), dim=(0, 1, 2), norm="ortho")
        m_true.append(mu[a] + dm)
    del zh, sq, lam
    mh = [torch.fft.rfftn(m, dim=(0, 1, 2), norm="ortho") for m in m_true]
    K = torch.from_numpy(kwd_table(wav, nzp)).to(dev)  # [nzp][nt]
    data = []
    sig_e = np.zeros(nt)
    for l in range(nt):
        acc = None
        for c in range(3):
            term = mh[c] * float(A[l, c])
            acc = term if acc is None else acc + term
        acc = acc * K[:, l][:, None, None]
        d = torch.fft.irfftn(acc, s=(nzp, nyp, nxp), dim=(0, 1, 2), norm="ortho")
        var = float(d[: cfg.nz, : cfg.ny, : cfg.nx].var())
        s = math.sqrt(var / float(cfg.snr[l]))
        sig_e[l] = s
        d = d + s * torch.randn((nzp, nyp, nxp), generator=gen, device=dev, dtype=f64)
        data.append(_padded(d, rnxp, 0.0))
    del mh
    prob.err_theta_cov = (sig_e[:, None] * sig_e[None, :]) * ang
    prob.mu = [_padded(m, rnxp, 0.0) for m in mu]
    prob.data = data
    prob.true_m = [_padded(m, rnxp, 0.0) for m in m_true]
    return prob


def post_params(prob: Problem, mixed_threshold: float = 1e-3):
    """Fill the C-ABI parameter struct of a problem."""
    from ._lib import PostParams
    cfg = prob.cfg
    p = PostParams()
    p.nxp, p.nyp, p.nzp, p.ntheta = cfg.nxp, cfg.nyp, cfg.nzp, prob.ntheta
    p.wnc = cfg.wnc
    for i, v in enumerate(prob.sigma0.reshape(-1)):
        p.sigma0[i] = float(v)
    for i, v in enumerate(prob.A.reshape(-1)):
        p.A[i] = float(v)
    for i, v in enumerate(prob.err_theta_cov.reshape(-1)):
        p.err_theta_cov[i] = float(v)
    p.mixed_threshold = mixed_threshold
    return p


def model_data(prob: Problem, tables: dict, mu_hat: np.ndarray, cov0_hat: np.ndarray, ec_hat: np.ndarray,
               scales, noise_var, chol0, seed: int = 7) -> list:
    """Fourier-domain data drawn from the likelihood model at given parameters:
    d(p) = K mu(p) + chol(M(p)) zeta(p), zeta ~ CN(0, I), at the in-band
    points (others are left at zero).  mu_hat [3][nzp][nyp][cnxp] complex,
    cov0_hat / ec_hat the Fourier-domain covariance grid 0 and errCorr.
    Returns nt complex128 arrays [nzp][nyp][cnxp]."""
    cfg = prob.cfg
    nt = prob.ntheta
    rng = np.random.default_rng(seed)
    shape = cov0_hat.shape
    out = [np.zeros(shape, np.complex128) for _ in range(nt)]
    L0 = np.zeros((3, 3))
    for q, (a, b) in enumerate([(0, 0), (1, 0), (1, 1), (2, 0), (2, 1), (2, 2)]):
        L0[a, b] = chol0[q]
    Sig0 = L0 @ L0.T
    bl = np.sqrt(np.diag(prob.err_theta_cov))
    rho = prob.err_theta_cov / np.outer(bl, bl)
    sd = np.sqrt(np.asarray(noise_var, np.float64))
    Se = np.outer(sd, sd) * rho
    wnc = float(np.float32(cfg.wnc))     # the interface carries wnc as a float
    c0 = 0.5 * (1 - wnc)
    for k in np.nonzero(tables["invert"])[0]:
        kw = (tables["kWD"][k, :, 0] + 1j * tables["kWD"][k, :, 1]) * np.asarray(scales)
        e1 = tables["e1"][k, :, 0] + 1j * tables["e1"][k, :, 1]
        e2 = tables["e2"][k, :, 0] + 1j * tables["e2"][k, :, 1]
        e3 = tables["e3"][k, :, 0].astype(np.float64)
        K = kw[:, None] * prob.A.astype(np.float64)
        g = np.outer(e1, e1.conj()) + np.outer(e2, e2.conj())
        Ec = c0 * Se * g
        Ec = Ec - 1j * np.diag(np.diag(Ec).imag)
        Ed = np.diag(wnc * np.diag(Se) * e3 ** 2)
        cabs = np.abs(cov0_hat[k].real / prob.sigma0[0, 0]).reshape(-1)
        lam = np.abs(ec_hat[k].real).reshape(-1)
        M = (cabs[:, None, None] * (K @ Sig0 @ K.conj().T)[None] + lam[:, None, None] * Ec[None] + Ed[None])
        M = 0.5 * (M + np.conj(np.swapaxes(M, 1, 2)))
        L = np.linalg.cholesky(M)
        zeta = (rng.standard_normal((M.shape[0], nt)) + 1j * rng.standard_normal((M.shape[0], nt))) / np.sqrt(2)
        mu_k = np.stack([m[k].reshape(-1) for m in mu_hat], -1)
        d = (mu_k @ K.T) + np.einsum("plm,pm->pl", L, zeta)
        for l in range(nt):
            out[l][k] = d[:, l].reshape(shape[1:])
    return out


# ---------------------------------------------------------------- non-constant thickness
def swdiv_inputs(nx, ny, nz, nxp, nyp, nzp, nt, seed, local=True, rel_range=(0.6, 1.0), wavelet="ricker"):
    """Synthetic inputs of the wavelet division: (positional arguments after
    the six sizes, keyword tables) of core.divide_by_scale_wavelet --
    ntheta grids, wavelets, norms, A, prior covariance, filtered temporal
    correlation, error variances, relThick, dz = 4, lz = 100, cuts 5 / 55 Hz,
    wnc 0.1; local shifts on even angles and gains on all but angle 1."""
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


def local_noise_problem(ntheta: int, nx: int, ny: int, seed: int = 0, uniform: float | None = None,
                        duplicate_angle: bool = False) -> dict:
    """A local-noise correction problem: angle error covariance E
    (variances ~1e-4, correlation exp(-3|dtheta|/30 deg)), prior variance Sm,
    its Bayesian update Smd = (Sm^-1 + K^T E^-1 K)^-1 for an Aki-Richards
    like K, and per-angle noise-scale tables [ny][nx] in [1, 3] (all equal
    to `uniform` when given).  duplicate_angle: the last two angles (and their
    scale tables) coincide, so that every trace's G Sm G^T + D E D is
    singular and its pseudo-inverse drops an eigenvalue."""
    import numpy as np
    rng = np.random.default_rng(seed)
    th = np.linspace(5.0, 40.0, ntheta)
    if duplicate_angle:  # the last two angles coincide: E is singular
        th[-1] = th[-2]
    ev = 1e-4 * (1.0 + rng.random(ntheta))
    rho = np.exp(-3.0 * np.abs(th[:, None] - th[None, :]) / 30.0)
    E = np.sqrt(ev)[:, None] * np.sqrt(ev)[None, :] * rho
    B = rng.normal(size=(3, 3))
    Sm = 1e-3 * (B @ B.T + 0.5 * np.eye(3))
    t = np.radians(th)
    K = np.stack([0.5 / np.cos(t) ** 2, -4.0 * 0.25 * np.sin(t) ** 2, 0.5 - 2.0 * 0.25 * np.sin(t) ** 2], axis=1)
    Smd = np.linalg.inv(np.linalg.inv(Sm) + K.T @ np.linalg.pinv(E) @ K)
    Smd = 0.5 * (Smd + Smd.T)
    if uniform is None:
        scales = [1.0 + 2.0 * rng.random((ny, nx)) for _ in range(ntheta)]
        if duplicate_angle:  # equal scales on the twin angles: G Sm G^T + D E D is singular
            scales[-1] = scales[-2].copy()
    else:
        scales = [np.full((ny, nx), float(uniform)) for _ in range(ntheta)]
    return dict(E=E, Sm=Sm, Smd=Smd, scales=scales)
