"""torch.autograd.Function over crava_hip_nll_grad.

The forward pass calls the GPU likelihood kernel, which returns the value and
the gradient together; backward scales the stored gradient.  Parameters are
float64 CPU tensors: wavelet scales [nt], noise variances [nt] and the
lower Cholesky factor of the prior covariance [6] (L00, L10, L11, L20, L21, L22).
"""
from __future__ import annotations

import ctypes as C

import torch

from . import _lib as L
from ._lib import check, lib, ptr_array
from .core import Posterior, _dev_ptr


def nll_grad(post: Posterior, precision: int, mu, data, cov, errcorr, scales, noise_var, chol0):
    """One call of crava_hip_nll_grad; returns (nll: float, grad: list[float])."""
    nt = len(data)
    s = (C.c_double * nt)(*[float(x) for x in scales])
    v = (C.c_double * nt)(*[float(x) for x in noise_var])
    l0 = (C.c_double * 6)(*[float(x) for x in chol0])
    out = C.c_double(0.0)
    g = (C.c_double * (2 * nt + 6))()
    check(lib().crava_hip_nll_grad(post.h, precision, ptr_array([_dev_ptr(t) for t in mu]),
                                   ptr_array([_dev_ptr(t) for t in data]), ptr_array([_dev_ptr(t) for t in cov]),
                                   _dev_ptr(errcorr), s, v, l0, C.byref(out), g), "crava_hip_nll_grad")
    return out.value, list(g)


class NLLFunction(torch.autograd.Function):
    """nll = NLLFunction.apply(scales, noise_var, chol0, problem) with
    problem = (post, precision, mu, data, cov, errcorr) of Fourier-domain GPU grids."""

    @staticmethod
    def forward(ctx, scales, noise_var, chol0, problem):
        post, precision, mu, data, cov, errcorr = problem
        val, grad = nll_grad(post, precision, mu, data, cov, errcorr, scales.tolist(), noise_var.tolist(),
                             chol0.tolist())
        nt = scales.numel()
        g = torch.tensor(grad, dtype=torch.float64)
        ctx.save_for_backward(g[:nt], g[nt:2 * nt], g[2 * nt:])
        return torch.tensor(val, dtype=torch.float64)

    @staticmethod
    def backward(ctx, gout):
        gs, gv, gc = ctx.saved_tensors
        return gout * gs, gout * gv, gout * gc, None


def nll(scales, noise_var, chol0, problem):
    return NLLFunction.apply(scales, noise_var, chol0, problem)
