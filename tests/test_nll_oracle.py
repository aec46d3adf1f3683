"""CPU pins of the FP64 likelihood oracle (oracle/nll_torch.py)."""
import numpy as np
import torch

import oracle
from crava_amd import synth
from oracle import nll_torch
from tests.helpers import rel_l2


def _fourier64(prob):
    cfg = prob.cfg
    n = (cfg.nxp, cfg.nyp, cfg.nzp)
    f = lambda t: np.ascontiguousarray(t.numpy().astype(np.float64))
    mu, d, cv, ec = [f(g) for g in prob.mu], [f(g) for g in prob.data], [f(g) for g in prob.cov], f(prob.errcorr)
    for g in mu + d:
        oracle.fft_r2c_f64(g, *n, False)
    for g in cv + [ec]:
        oracle.fft_r2c_f64(g, *n, True)
    c = lambda g: g.view(np.complex128)
    return [c(g) for g in mu], [c(g) for g in d], [c(g) for g in cv], c(ec)


def theta0(prob):
    nt = prob.ntheta
    L0 = np.linalg.cholesky(prob.sigma0)
    return (torch.ones(nt, dtype=torch.float64), torch.from_numpy(np.diag(prob.err_theta_cov).copy()),
            torch.tensor([L0[0, 0], L0[1, 0], L0[1, 1], L0[2, 0], L0[2, 1], L0[2, 2]], dtype=torch.float64))


def test_nll_oracle_equals_dense_gaussian_tiny():
    """1/2 sum_p w [r^H M^-1 r + log det M] = 1/2 r^T C^-1 r + 1/2 log det C for
    the dense time-domain data covariance C on a 4x4x8 grid, 2 angles."""
    prob = synth.make_problem(synth.config("tiny"))
    cfg = prob.cfg
    nxp, nyp, nzp, nt = cfg.nxp, cfg.nyp, cfg.nzp, prob.ntheta
    N = nxp * nyp * nzp
    tabs = oracle.perk_tables(nt, cfg.nz, nzp, prob.lz, cfg.low_cut, cfg.high_cut, 1.0, prob.wavelets, prob.wnorm)
    tabs["invert"][:] = 1                # every plane, so the sum covers the whole spectrum
    mu, d, cv, ec = _fourier64(prob)
    P = nll_torch.prepare(tabs, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0[0, 0], mu, d, cv[0], ec, nxp)
    s, v, l0 = theta0(prob)
    val = float(nll_torch.nll(P, s, v, l0))
    # dense: full spectrum via Hermitian extension of the per-point blocks
    F = np.fft.fftn(np.eye(N).reshape(N, nzp, nyp, nxp), axes=(1, 2, 3), norm="ortho").reshape(N, N).T
    Fi = F.conj().T
    full = lambda t: np.fft.fftn(t.numpy().astype(np.float64)[:, :, :nxp], axes=(0, 1, 2))
    cabs = np.abs(full(prob.cov[0]).real / prob.sigma0[0, 0]).reshape(-1)
    lam = np.abs(full(prob.errcorr).real).reshape(-1)
    kz = np.repeat(np.arange(nzp), nyp * nxp)
    kfull = lambda key: (tabs[key][..., 0].astype(np.float64) + 1j * tabs[key][..., 1])[kz]
    kw, e1, e2 = kfull("kWD"), kfull("e1"), kfull("e2")
    Sig0 = prob.sigma0
    Se = prob.err_theta_cov
    Kb = lambda l, c: Fi @ np.diag(kw[:, l] * prob.A[l, c]) @ F
    G = np.block([[Kb(l, c) for c in range(3)] for l in range(nt)])
    Cm = np.block([[Fi @ np.diag(cabs * Sig0[a, b]) @ F for b in range(3)] for a in range(3)])
    wnc = float(np.float32(cfg.wnc))
    Eb = lambda l, m: (0.5 * (1 - wnc) * Se[l, m] * lam * (e1[:, l] * e1[:, m].conj() + e2[:, l] * e2[:, m].conj())
                       + (wnc * Se[l, l] if l == m else 0.0))
    Ce = np.block([[Fi @ np.diag(Eb(l, m)) @ F for m in range(nt)] for l in range(nt)])
    Cd = G @ Cm @ G.conj().T + Ce
    Cd = 0.5 * (Cd + Cd.conj().T)
    mvec = np.concatenate([g.numpy().astype(np.float64)[:, :, :nxp].reshape(-1) for g in prob.mu])
    dvec = np.concatenate([g.numpy().astype(np.float64)[:, :, :nxp].reshape(-1) for g in prob.data])
    r = dvec - (G @ mvec).real
    sign, logdet = np.linalg.slogdet(Cd.real)
    dense = 0.5 * r @ np.linalg.solve(Cd.real, r) + 0.5 * logdet
    assert sign > 0
    assert abs(val - dense) < 1e-8 * max(1.0, abs(dense))


def test_nll_oracle_gradient_finite_differences():
    prob = synth.make_problem(synth.config("small"))
    cfg = prob.cfg
    tabs = oracle.perk_tables(prob.ntheta, cfg.nz, cfg.nzp, prob.lz, cfg.low_cut, cfg.high_cut, 1.0,
                              prob.wavelets, prob.wnorm)
    mu, d, cv, ec = _fourier64(prob)
    P = nll_torch.prepare(tabs, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0[0, 0], mu, d, cv[0], ec, cfg.nxp)
    th = [t.clone().requires_grad_(True) for t in theta0(prob)]
    f = nll_torch.nll(P, *th)
    f.backward()
    for which in range(3):
        for i in range(th[which].numel()):
            h = 1e-6 * max(1.0, abs(float(th[which][i])))
            vals = []
            for sgn in (1, -1):
                t = [x.detach().clone() for x in th]
                t[which][i] += sgn * h
                vals.append(float(nll_torch.nll(P, *t)))
            fd = (vals[0] - vals[1]) / (2 * h)
            ad = float(th[which].grad[i])
            assert abs(fd - ad) <= 1e-5 * max(1.0, abs(ad)), (which, i, fd, ad)
