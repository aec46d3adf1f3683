"""FP64 oracle of the negative log marginal likelihood (test infrastructure).

An independent, dense restatement in torch float64 on the CPU of the model
the posterior loop assumes (src/crava.cpp:1087-1130, error model
src/crava.cpp:1298-1334): per in-band half-complex Fourier point,

    M = K S K^H + E,  K = diag(s kWD) A,  S = |c| L0 L0^T,
    E = lambda c0 (Sigma_e o g) + diag(wnc Sigma_e,ll e3^2),
    Sigma_e,lm = sqrt(v_l v_m) rho_lm,
    NLL = 1/2 sum_p w_p [ r^H M^-1 r + log det M ],  r = d - K mu,

with w_p = 1 on the self-conjugate x planes and 2 elsewhere.  torch's
autograd of this function is the reference gradient; finite differences
pin it in turn.  Only tests/ import this module.
"""
from __future__ import annotations

import numpy as np
import torch


def _c(t: np.ndarray) -> torch.Tensor:
    return torch.from_numpy(np.ascontiguousarray(t)).to(torch.complex128)


def prepare(tabs, A, err_theta_cov, wnc, base00, mu, data, cov0, errcorr, nxp):
    """Gather the in-band points of Fourier-domain grids (complex arrays
    [nzp][nyp][cnxp]) into flat tensors."""
    inv = tabs["invert"].astype(bool)
    ks = np.nonzero(inv)[0]
    cn = nxp // 2 + 1
    nt = A.shape[0]
    sel = lambda g: _c(np.asarray(g)[ks].reshape(len(ks), -1))       # [nk, plane]
    d = torch.stack([sel(g) for g in data], -1)                       # [nk, plane, nt]
    m = torch.stack([sel(g) for g in mu], -1)                         # [nk, plane, 3]
    cabs = (sel(cov0).real / base00).abs()
    lam = sel(errcorr).real.abs()
    plane = d.shape[1]
    i = torch.arange(plane) % cn
    w = torch.where((i == 0) | ((nxp % 2 == 0) & (i == cn - 1)), 1.0, 2.0).to(torch.float64)
    kw = lambda key: _c(tabs[key][ks, :, 0] + 1j * tabs[key][ks, :, 1])   # [nk, nt]
    bl = np.sqrt(np.diag(err_theta_cov))
    rho = err_theta_cov / np.outer(bl, bl)
    return dict(d=d, m=m, cabs=cabs, lam=lam, w=w, kwd=kw("kWD"), e1=kw("e1"), e2=kw("e2"),
                e3=torch.from_numpy(tabs["e3"][ks, :, 0].astype(np.float64)),
                A=torch.from_numpy(A.astype(np.float64)), rho=torch.from_numpy(rho),
                wnc=float(np.float32(wnc)), nt=nt)   # wnc is a float in the interface


def nll(P, scales: torch.Tensor, noise_var: torch.Tensor, chol0: torch.Tensor) -> torch.Tensor:
    nt = P["nt"]
    wnc = P["wnc"]
    z0 = torch.zeros((), dtype=torch.float64)
    L0 = torch.stack([torch.stack([chol0[0], z0, z0]), torch.stack([chol0[1], chol0[2], z0]),
                      torch.stack([chol0[3], chol0[4], chol0[5]])])
    Sig0 = L0 @ L0.T
    sd = torch.sqrt(noise_var)
    Se = sd[:, None] * sd[None, :] * P["rho"]
    K = (P["kwd"] * scales)[..., :, None] * P["A"]                       # [nk, nt, 3]
    S = P["cabs"][..., None, None] * Sig0                                 # [nk, plane, 3, 3]
    # elementwise products and sums only: batched complex BLAS paths lose ~1e-9
    Kc = K.to(torch.complex128)
    KS = (Kc[:, None, :, :, None] * S.to(torch.complex128)[:, :, None, :, :]).sum(-2)      # [nk, p, nt, 3]
    Ms = (KS[:, :, :, None, :] * Kc.conj()[:, None, None, :, :]).sum(-1)                  # [nk, p, nt, nt]
    g = P["e1"][..., :, None] * P["e1"].conj()[..., None, :] + P["e2"][..., :, None] * P["e2"].conj()[..., None, :]
    c0 = 0.5 * (1.0 - wnc)
    Ec = c0 * Se * g                                                      # [nk, nt, nt]
    eye = torch.eye(nt, dtype=torch.float64)
    Ec = Ec - 1j * Ec.imag * eye                                          # real diagonal
    Ed = wnc * torch.diagonal(Se) * P["e3"] ** 2                          # [nk, nt]
    M = Ms + P["lam"][..., None, None] * Ec[:, None] + torch.diag_embed(Ed)[:, None].to(torch.complex128)
    M = 0.5 * (M + M.conj().transpose(-1, -2))
    r = P["d"] - (Kc[:, None, :, :] * P["m"][:, :, None, :]).sum(-1)
    L = torch.linalg.cholesky(M)
    z = torch.linalg.solve_triangular(L, r[..., None], upper=False)[..., 0]
    quad = (z.abs() ** 2).sum(-1)
    logdet = 2.0 * torch.log(torch.diagonal(L, dim1=-2, dim2=-1).real).sum(-1)
    return 0.5 * (P["w"] * (quad + logdet)).sum()
