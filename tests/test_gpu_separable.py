"""crava_hip_post_run_separable: the posterior of a separable prior
Sigma0 (x) corr from the one transformed correlation grid, against the
oracles run on the six grids corr_hat * Sigma0_ab (the form
SeismicParametersHolder::initializeCorrelations builds,
src/seismicparametersholder.cpp:197-221), where the cleaned prior
|Re(C_ab) / Sigma0_ab| Sigma0_ab equals |Re corr_hat| Sigma0_ab.

Tolerances: FP64 1e-8 (the FP64 form's bar against the FP64 oracle); FP32
against the float oracle as the six-grid FP32 tests (1e-6 mean, 2e-5 the
rest: the float prior differs by the rounding of corr * Sigma0_ab); MIXED
1e-5 against the FP64 oracle, in the 512x512x320 workload's noise regime.
"""
import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd import synth
from crava_amd.core import Posterior, problem_tables
from tests.helpers import rel_l2
from tests.test_gpu_mixed import precision_regime

PAIRS = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]


def _cplx(g):
    return g.view(np.complex64 if g.dtype == np.float32 else np.complex128)


def _fourier64(prob):
    """FP64 Fourier-domain inputs: means, data, correlation, error correlation."""
    c = prob.cfg
    n = (c.nxp, c.nyp, c.nzp)

    def f(g, cov):
        x = g.numpy().astype(np.float64)
        oracle.fft_r2c_f64(x, *n, cov)
        return x
    return [f(g, False) for g in prob.mu], [f(g, False) for g in prob.data], f(prob.corr, True), \
        f(prob.errcorr, True)


def _six(corr, sigma0, dtype):
    """The six covariance spectra corr * Sigma0_ab (complex layout as floats)."""
    cc = _cplx(corr)
    out = []
    for a, b in PAIRS:
        g = (cc * sigma0[a, b]).astype(np.complex64 if dtype == np.float32 else np.complex128)
        out.append(np.ascontiguousarray(g).view(dtype))
    return out


def _gpu(ctx, prob, precision, thr, mu, d, corr, ec):
    t = lambda g: torch.from_numpy(np.ascontiguousarray(g)).cuda()
    tm, td = [t(g.copy()) for g in mu], [t(g.copy()) for g in d]
    tc = [torch.full_like(tm[0], np.nan) for _ in range(6)]     # outputs only
    post = Posterior(ctx, synth.post_params(prob, thr), problem_tables(prob))
    nf, n64 = post.run(precision, tm, td, tc, t(ec.copy()), counts=True, corr=t(corr.copy()))
    torch.cuda.synchronize()
    post.close()
    return [x.cpu().numpy() for x in tm], [x.cpu().numpy() for x in td], [x.cpu().numpy() for x in tc], nf, n64


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small", "plain_cpu"])
def test_separable_f64_and_f32_vs_oracle(ctx, name):
    print(f"separable posterior {name}", flush=True)
    prob = synth.make_problem(synth.config(name), "cpu")
    cfg = prob.cfg
    n = (cfg.nxp, cfg.nyp, cfg.nzp)
    tabs = problem_tables(prob)
    mu, d, corr, ec = _fourier64(prob)
    # FP64: GPU separable vs the FP64 oracle on the six grids
    om, od, oc = [g.copy() for g in mu], [g.copy() for g in d], _six(corr, prob.sigma0, np.float64)
    nf_ref = oracle.posterior_f64(*n, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, tabs, om, od, oc, ec.copy())
    gm, gd, gc, nf, _ = _gpu(ctx, prob, L.F64, 1e-3, mu, d, corr, ec)
    assert nf == nf_ref
    for a, b in zip(gm + gd + gc, om + od + oc):
        assert rel_l2(_cplx(a), _cplx(b)) < 1e-8
    # FP32: GPU separable vs the float oracle on the six grids
    f32 = lambda g: np.ascontiguousarray(_cplx(g).astype(np.complex64)).view(np.float32)
    mu32, d32, corr32, ec32 = [f32(g) for g in mu], [f32(g) for g in d], f32(corr), f32(ec)
    om, od = [g.copy() for g in mu32], [g.copy() for g in d32]
    oc = _six(corr32, prob.sigma0.astype(np.float32), np.float32)
    nf_ref = oracle.posterior(*n, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, tabs, om, od, oc, ec32.copy())
    gm, gd, gc, nf, _ = _gpu(ctx, prob, L.F32, 1e-3, mu32, d32, corr32, ec32)
    assert nf == nf_ref
    for a, b in zip(gm, om):
        assert rel_l2(_cplx(a), _cplx(b)) < 1e-6
    for a, b in zip(gd + gc, od + oc):
        assert rel_l2(_cplx(a), _cplx(b)) < 2e-5


@pytest.mark.gpu
@pytest.mark.parametrize("thr", [1e-3, 3e-2], ids=["thr1e-3", "thr3e-2"])
def test_separable_mixed_vs_fp64_oracle(ctx, thr):
    print(f"separable mixed thr={thr:g}", flush=True)
    prob = synth.make_problem(precision_regime(), "cpu")
    cfg = prob.cfg
    n = (cfg.nxp, cfg.nyp, cfg.nzp)
    tabs = problem_tables(prob)
    mu, d, corr, ec = _fourier64(prob)
    f32 = lambda g: np.ascontiguousarray(_cplx(g).astype(np.complex64)).view(np.float32)
    mu32, d32, corr32, ec32 = [f32(g) for g in mu], [f32(g) for g in d], f32(corr), f32(ec)
    up = lambda g: np.ascontiguousarray(_cplx(g).astype(np.complex128)).view(np.float64)
    pm = [up(g) for g in mu32]
    om, od, oc = [g.copy() for g in pm], [up(g) for g in d32], _six(up(corr32), prob.sigma0, np.float64)
    oracle.posterior_f64(*n, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, tabs, om, od, oc, up(ec32))
    gm, gd, gc, nf, n64 = _gpu(ctx, prob, L.MIXED, thr, mu32, d32, corr32, ec32)
    assert nf == 0
    e = {"mean_update": max(rel_l2(_cplx(a).astype(np.complex128) - _cplx(p), _cplx(b) - _cplx(p))
                            for a, b, p in zip(gm, om, pm)),
         "mean": max(rel_l2(_cplx(a), _cplx(b)) for a, b in zip(gm, om)),
         "cov": max(rel_l2(_cplx(a), _cplx(b)) for a, b in zip(gc, oc)),
         "residual": max(rel_l2(_cplx(a), _cplx(b)) for a, b in zip(gd, od))}
    print(f"  FP64 points {n64}; errors vs FP64 oracle {e}", flush=True)
    for k, v in e.items():
        assert v < 1e-5, (k, v)
