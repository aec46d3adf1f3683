"""CRAVA_HIP_MIXED (FP32 square-root solve, FP64 for waves below the
noise/signal threshold; covariance transforms with double arithmetic inside
each pass, mean transforms with mean deflation) against the FP64 oracle, in
the noise regime of
the 512x512x320 precision workload: 6 angles at 4-40 deg, per-angle S/N
1e4, 3e3, 1e3, 300, 100, 20.  The grid is that workload's 320-sample z axis
with a 64 x 64 lateral FFT grid, so the oracle runs in seconds.

Tolerance: 1e-5 relative L2 on the posterior mean update (mu_post - mu_prior),
the posterior mean, every covariance grid and every residual grid -- the
accuracy bar of the mixed form.  The condition number of M = K S K^H + E
reached at the in-band points is printed (about 1e5 here).
"""
import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd import synth
from crava_amd.core import Posterior, problem_tables
from tests.helpers import as_f64, fourier_inputs, rel_l2

TOL = 1e-5


def precision_regime(nxp=64, seed=1):
    cfg = synth.config("precision", seed)
    cfg.name = f"precision_{nxp}"
    cfg.nx = cfg.ny = nxp * 7 // 8
    cfg.nxp = cfg.nyp = nxp
    return cfg


def _cplx(g):
    return g.view(np.complex64 if g.dtype == np.float32 else np.complex128)


def kappa_max(prob, tabs, cov, ec):
    """Largest condition number of M over the in-band points (FP64, host)."""
    cfg = prob.cfg
    nt = prob.ntheta
    A = prob.A.astype(np.float64)
    wnc = float(np.float32(cfg.wnc))
    c0 = 0.5 * (1.0 - wnc)
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    worst = 0.0
    for k in np.nonzero(tabs["invert"])[0]:
        kw = tabs["kWD"][k, :, 0].astype(np.float64) + 1j * tabs["kWD"][k, :, 1]
        e1 = tabs["e1"][k, :, 0] + 1j * tabs["e1"][k, :, 1]
        e2 = tabs["e2"][k, :, 0] + 1j * tabs["e2"][k, :, 1]
        e3 = tabs["e3"][k, :, 0].astype(np.float64)
        Ec = c0 * prob.err_theta_cov * (np.outer(e1, e1.conj()) + np.outer(e2, e2.conj()))
        Ec[np.diag_indices(nt)] = Ec.diagonal().real
        Ed = np.diag(wnc * np.diag(prob.err_theta_cov) * e3 ** 2)
        K = kw[:, None] * A
        S = np.zeros((cov[0][k].size // 2, 3, 3))
        for q, (a, b) in enumerate(pairs):
            v = _cplx(cov[q][k]).real.reshape(-1).astype(np.float64)
            S[:, a, b] = S[:, b, a] = np.abs(v / prob.sigma0[a, b]) * prob.sigma0[a, b]
        lam = np.abs(_cplx(ec[k]).real.reshape(-1).astype(np.float64))
        M = K[None] @ S.astype(np.complex128) @ K.conj().T[None] + lam[:, None, None] * Ec[None] + Ed[None]
        ev = np.linalg.eigvalsh(M)
        worst = max(worst, float((ev[:, -1] / ev[:, 0]).max()))
    return worst


@pytest.fixture(scope="module")
def regime():
    prob = synth.make_problem(precision_regime(), "cpu")
    cfg = prob.cfg
    tabs = problem_tables(prob)
    grids = fourier_inputs(prob)
    m64, d64, c64 = as_f64(grids[0]), as_f64(grids[1]), as_f64(grids[2])
    e64 = as_f64([grids[3]])[0]
    om, od, oc = [g.copy() for g in m64], [g.copy() for g in d64], [g.copy() for g in c64]
    nf = oracle.posterior_f64(cfg.nxp, cfg.nyp, cfg.nzp, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, tabs,
                              om, od, oc, e64.copy())
    assert nf == 0
    n_in = int(tabs["invert"].sum()) * (cfg.nxp // 2 + 1) * cfg.nyp
    return dict(prob=prob, tabs=tabs, grids=grids, prior_mu=m64, out=(om, od, oc), n_in=n_in)


def _errors(gm, gd, gc, prior_mu, out):
    om, od, oc = out
    return {
        "mean_update": max(rel_l2(_cplx(a).astype(np.complex128) - _cplx(p), _cplx(b) - _cplx(p))
                           for a, b, p in zip(gm, om, prior_mu)),
        "mean": max(rel_l2(_cplx(a), _cplx(b)) for a, b in zip(gm, om)),
        "cov": max(rel_l2(_cplx(a), _cplx(b)) for a, b in zip(gc, oc)),
        "residual": max(rel_l2(_cplx(a), _cplx(b)) for a, b in zip(gd, od)),
    }


@pytest.mark.gpu
@pytest.mark.parametrize("thr", [0.0, 1e-3, 3e-2, 1e30], ids=["fp32", "thr1e-3", "thr3e-2", "fp64"])
def test_mixed_posterior_vs_fp64_oracle(ctx, regime, thr):
    """The posterior alone, on the same float Fourier-domain inputs: every
    output within TOL of the FP64 oracle, at thresholds that send no wave,
    some waves or every wave to FP64."""
    prob, tabs = regime["prob"], regime["tabs"]
    print(f"mixed posterior thr={thr:g}", flush=True)
    mu, d, cv, ec = regime["grids"]
    tm = [torch.from_numpy(g.copy()).cuda() for g in mu]
    td = [torch.from_numpy(g.copy()).cuda() for g in d]
    tc = [torch.from_numpy(g.copy()).cuda() for g in cv]
    te = torch.from_numpy(ec.copy()).cuda()
    post = Posterior(ctx, synth.post_params(prob, thr), tabs)
    nf, n64 = post.run(L.MIXED, tm, td, tc, te, counts=True)
    torch.cuda.synchronize()
    post.close()
    assert nf == 0
    n_in = regime["n_in"]
    if thr == 0.0:
        assert n64 == 0
    elif thr > 1.0:
        assert n64 == n_in
    elif thr == 3e-2:
        assert 0 < n64 < n_in
    e = _errors([t.cpu().numpy() for t in tm], [t.cpu().numpy() for t in td], [t.cpu().numpy() for t in tc],
                regime["prior_mu"], regime["out"])
    print(f"  FP64 points {n64} of {n_in}; errors vs FP64 oracle {e}", flush=True)
    for k, v in e.items():
        assert v < TOL, (k, v)


@pytest.mark.gpu
def test_condition_number_of_the_regime(regime):
    """The regime is the ill-conditioned one the mixed form is for."""
    mu, d, cv, ec = regime["grids"]
    kap = kappa_max(regime["prob"], regime["tabs"], cv, ec)
    print(f"largest cond(M) over the in-band points: {kap:.3g}", flush=True)
    assert kap > 1e4


@pytest.mark.gpu
def test_mixed_path_with_transforms_vs_fp64_oracle(ctx, regime):
    """The whole path in the time domain: forward transforms (data and error
    correlation in float, means in float with mean deflation, covariances
    with double arithmetic in each pass), MIXED posterior at threshold 1e-3,
    inverse transforms (means deflated, residuals and covariances in float)
    -- against FP64 transforms and the FP64 posterior of the oracle."""
    prob = regime["prob"]
    cfg = prob.cfg
    n = (cfg.nxp, cfg.nyp, cfg.nzp)
    print("mixed path with transforms", flush=True)
    tm = [g.clone().cuda() for g in prob.mu]
    td = [g.clone().cuda() for g in prob.data]
    tc = [g.clone().cuda() for g in prob.cov]
    te = prob.errcorr.clone().cuda()
    ctx.r2c_many(td, *n, L.DATA)
    ctx.r2c(te, *n, L.COVARIANCE)
    ctx.r2c_many(tm, *n, L.DATA, L.F32_DEFLATE)
    ctx.r2c_many(tc, *n, L.COVARIANCE, L.MIXED)
    post = Posterior(ctx, synth.post_params(prob, 1e-3), regime["tabs"])
    nf, n64 = post.run(L.MIXED, tm, td, tc, te, counts=True)
    post.close()
    ctx.c2r_many(tm, *n, L.DATA, L.F32_DEFLATE)
    ctx.c2r_many(td, *n, L.DATA)
    ctx.c2r_many(tc, *n, L.COVARIANCE)
    torch.cuda.synchronize()

    def f64(g):
        return g.numpy().astype(np.float64)
    om, od, oc = [f64(g) for g in prob.mu], [f64(g) for g in prob.data], [f64(g) for g in prob.cov]
    oe = f64(prob.errcorr)
    for g in om + od:
        oracle.fft_r2c_f64(g, *n, False)
    for g in oc + [oe]:
        oracle.fft_r2c_f64(g, *n, True)
    oracle.posterior_f64(*n, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, regime["tabs"], om, od, oc, oe)
    for g in om + od:
        oracle.fft_c2r_f64(g, *n, False)
    for g in oc:
        oracle.fft_c2r_f64(g, *n, True)
    x = cfg.nxp
    gm = [t.cpu().numpy()[:, :, :x].astype(np.float64) for t in tm]
    e = {"mean_update": max(rel_l2(a - p.numpy()[:, :, :x], b[:, :, :x] - p.numpy()[:, :, :x])
                            for a, b, p in zip(gm, om, prob.mu)),
         "mean": max(rel_l2(a, b[:, :, :x]) for a, b in zip(gm, om)),
         "cov": max(rel_l2(t.cpu().numpy()[:, :, :x], b[:, :, :x]) for t, b in zip(tc, oc)),
         "residual": max(rel_l2(t.cpu().numpy()[:, :, :x], b[:, :, :x]) for t, b in zip(td, od))}
    print(f"  FP64 points {n64}; time-domain errors vs FP64 oracle {e}", flush=True)
    assert nf == 0
    for k, v in e.items():
        assert v < TOL, (k, v)
