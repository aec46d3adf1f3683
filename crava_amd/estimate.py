"""Maximum-likelihood estimation of wavelet scales, noise variances and the
prior covariance factor with L-BFGS over the GPU likelihood adjoint.

No reference counterpart: CRAVA takes these as fixed inputs (wavelet
estimation and S/N from wells, src/crava.cpp:322-473).  The optimiser works on
unconstrained coordinates: log scales, log noise variances and the Cholesky
factor of Sigma0 with log-diagonal.
"""
from __future__ import annotations

import math

import torch

from .autograd import nll as nll_fn

DIAG = (0, 2, 5)   # positions of L00, L11, L22 in the packed factor


def pack(scales, noise_var, chol0) -> torch.Tensor:
    c = torch.as_tensor(chol0, dtype=torch.float64).clone()
    for i in DIAG:
        c[i] = math.log(float(c[i]))
    return torch.cat([torch.log(torch.as_tensor(scales, dtype=torch.float64)),
                      torch.log(torch.as_tensor(noise_var, dtype=torch.float64)), c])


def unpack(x: torch.Tensor, nt: int):
    s = torch.exp(x[:nt])
    v = torch.exp(x[nt:2 * nt])
    c = x[2 * nt:]
    parts = [torch.exp(c[i]) if i in DIAG else c[i] for i in range(6)]
    return s, v, torch.stack(parts)


def fit(problem, nt: int, scales0, noise_var0, chol00, steps: int = 50, history: int = 20, free=None):
    """L-BFGS from the given start; returns (scales, noise_var, chol0, trace)
    where trace lists (nll, |grad|) per function evaluation.  `free` is an
    optional boolean mask over the packed parameters (scales, noise
    variances, Sigma0 factor); masked-out entries stay at their start.

    The wavelet scales and Sigma0 enter K S K^H only through s_l s_m Sigma0,
    so a common factor between them is identified by the mean term K mu alone;
    fitting the scales with Sigma0 held (or the reverse) is well posed."""
    x = pack(scales0, noise_var0, chol00).requires_grad_(True)
    mask = None if free is None else torch.as_tensor(free, dtype=torch.float64)
    opt = torch.optim.LBFGS([x], lr=1.0, max_iter=steps, history_size=history, line_search_fn="strong_wolfe",
                            tolerance_grad=1e-12, tolerance_change=1e-14)
    trace = []

    def closure():
        opt.zero_grad()
        s, v, c = unpack(x, nt)
        f = nll_fn(s, v, c, problem)
        f.backward()
        if mask is not None:
            x.grad.mul_(mask)
        trace.append((float(f.detach()), float(x.grad.norm())))
        return f

    opt.step(closure)
    s, v, c = unpack(x.detach(), nt)
    return s, v, c, trace
