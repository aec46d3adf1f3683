"""GPU per-point posterior against the float oracle and the FP64 oracle."""
import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd import synth
from crava_amd.core import Posterior, problem_tables
from tests.helpers import as_f64, fourier_inputs, rel_l2


def _run_gpu(ctx, prob, grids, precision, thr=1e-3):
    mu, d, cv, ec = grids
    want = np.float64 if precision == L.F64 else np.float32
    def up(g):
        assert g.dtype == want
        return torch.from_numpy(g.copy()).cuda()
    tm = [up(g) for g in mu]
    td = [up(g) for g in d]
    tc = [up(g) for g in cv]
    te = up(ec)
    post = Posterior(ctx, synth.post_params(prob, thr), problem_tables(prob))
    nf, n64 = post.run(precision, tm, td, tc, te, counts=True)
    torch.cuda.synchronize()
    post.close()
    return [t.cpu().numpy() for t in tm], [t.cpu().numpy() for t in td], [t.cpu().numpy() for t in tc], nf, n64


def _cplx(g):
    return g.view(np.complex64 if g.dtype == np.float32 else np.complex128)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small", "plain_cpu"])
def test_posterior_f32_vs_oracle(ctx, name):
    print(f"posterior f32 {name}", flush=True)
    prob = synth.make_problem(synth.config(name), "cpu")
    cfg = prob.cfg
    grids = fourier_inputs(prob)
    tabs = problem_tables(prob)
    mu, d, cv, ec = [x if not isinstance(x, list) else [g.copy() for g in x] for x in grids]
    ec = ec.copy()
    nf_ref = oracle.posterior(cfg.nxp, cfg.nyp, cfg.nzp, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0,
                              tabs, mu, d, cv, ec)
    gm, gd, gc, nf, _ = _run_gpu(ctx, prob, grids, L.F32)
    assert nf == nf_ref
    for a, b in zip(gm, mu):
        assert rel_l2(_cplx(a), _cplx(b)) < 1e-6
    for a, b in zip(gd, d):
        assert rel_l2(_cplx(a), _cplx(b)) < 2e-5
    for a, b in zip(gc, cv):
        assert rel_l2(_cplx(a), _cplx(b)) < 2e-5


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small", "plain_cpu"])
def test_posterior_f64_and_mixed_vs_fp64_oracle(ctx, name):
    print(f"posterior f64/mixed {name}", flush=True)
    prob = synth.make_problem(synth.config(name), "cpu")
    cfg = prob.cfg
    grids = fourier_inputs(prob)
    tabs = problem_tables(prob)
    m64, d64, c64 = as_f64(grids[0]), as_f64(grids[1]), as_f64(grids[2])
    e64 = as_f64([grids[3]])[0]
    g64 = (m64, d64, c64, e64)
    om, od, oc = [g.copy() for g in m64], [g.copy() for g in d64], [g.copy() for g in c64]
    nf_ref = oracle.posterior_f64(cfg.nxp, cfg.nyp, cfg.nzp, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0,
                                  tabs, om, od, oc, e64.copy())
    gm, gd, gc, nf, _ = _run_gpu(ctx, prob, g64, L.F64)
    assert nf == nf_ref
    # the GPU's Gram form S - Y^H Y and the oracle's substitution form differ by
    # FP64 roundings amplified by cond(M) and by the cancellation in S - R
    # (observed <= 1.3e-9 normwise); a dropped or mis-signed term is O(1)
    for a, b in zip(gm + gd + gc, om + od + oc):
        assert rel_l2(_cplx(a), _cplx(b)) < 1e-8
    # mixed: float grids, every point in FP64 (threshold huge) -> FP64 accuracy up to float rounding
    gm, gd, gc, nf, n64 = _run_gpu(ctx, prob, grids, L.MIXED, thr=1e30)
    assert nf == nf_ref
    assert n64 > 0
    for a, b in zip(gc, oc):
        assert rel_l2(_cplx(a), _cplx(b)) < 1e-6


@pytest.mark.gpu
def test_posterior_f32_vs_golden_reference_outputs(ctx):
    """GPU posterior on the golden fixture's inputs against the reference's
    lib_matr outputs stored with them (tests/golden)."""
    import os
    G = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "reference_cases.npz"))
    nxp, nyp, nzp, nt = (int(x) for x in G["post_dims"])
    p = L.PostParams()
    p.nxp, p.nyp, p.nzp, p.ntheta = nxp, nyp, nzp, nt
    p.wnc = float(G["post_wnc"])
    for i, v in enumerate(G["post_sigma0"].reshape(-1)):
        p.sigma0[i] = float(v)
    for i, v in enumerate(G["post_A"].reshape(-1)):
        p.A[i] = float(v)
    for i, v in enumerate(G["post_etc"].reshape(-1)):
        p.err_theta_cov[i] = float(v)
    p.mixed_threshold = 1e-3
    tabs = {k[len("post_tab_"):]: G[k] for k in G.files if k.startswith("post_tab_")}
    post = Posterior(ctx, p, tabs)
    tm = [torch.from_numpy(g.copy()).cuda() for g in G["post_in_mu"]]
    td = [torch.from_numpy(g.copy()).cuda() for g in G["post_in_d"]]
    tc = [torch.from_numpy(g.copy()).cuda() for g in G["post_in_cov"]]
    te = torch.from_numpy(G["post_in_ec"].copy()).cuda()
    nf, _ = post.run(L.F32, tm, td, tc, te, counts=True)
    torch.cuda.synchronize()
    post.close()
    assert nf == int(G["post_nfail"])
    for got, ref in zip(tm, G["post_out_mu"]):
        assert rel_l2(_cplx(got.cpu().numpy()), _cplx(ref)) < 1e-6
    for got, ref in zip(td + tc, list(G["post_out_d"]) + list(G["post_out_cov"])):
        assert rel_l2(_cplx(got.cpu().numpy()), _cplx(ref)) < 2e-5


@pytest.mark.gpu
@pytest.mark.parametrize("nt", [1, 2, 5, 6, 7, 8])
def test_posterior_every_ntheta_vs_oracle(ctx, nt):
    """Every angle count the kernel is instantiated for (1..8; 3 and 4 are the
    small/plain_cpu cases above), F32 against the float oracle and MIXED with
    every point in FP64 against the FP64 oracle, on an odd-x grid."""
    print(f"posterior ntheta={nt}", flush=True)
    th = list(np.linspace(5.0, 40.0, nt))
    cfg = synth.Config(f"nt{nt}", 12, 10, 20, 15, 12, 24, th, [35.0] * nt, list(np.linspace(0.9, 1.1, nt)),
                       list(np.linspace(8.0, 3.0, nt)), low_cut=5.0, high_cut=100.0, seed=nt)
    prob = synth.make_problem(cfg, "cpu")
    grids = fourier_inputs(prob)
    tabs = problem_tables(prob)
    mu, d, cv, ec = [g.copy() for g in grids[0]], [g.copy() for g in grids[1]], [g.copy() for g in grids[2]], \
        grids[3].copy()
    nf_ref = oracle.posterior(cfg.nxp, cfg.nyp, cfg.nzp, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0,
                              tabs, mu, d, cv, ec)
    gm, gd, gc, nf, _ = _run_gpu(ctx, prob, grids, L.F32)
    assert nf == nf_ref
    for a, b in zip(gm, mu):
        assert rel_l2(_cplx(a), _cplx(b)) < 1e-6
    for a, b in zip(gd + gc, d + cv):
        assert rel_l2(_cplx(a), _cplx(b)) < 2e-5
    m64, d64, c64 = as_f64(grids[0]), as_f64(grids[1]), as_f64(grids[2])
    e64 = as_f64([grids[3]])[0]
    om, od, oc = [g.copy() for g in m64], [g.copy() for g in d64], [g.copy() for g in c64]
    oracle.posterior_f64(cfg.nxp, cfg.nyp, cfg.nzp, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, tabs,
                         om, od, oc, e64.copy())
    gm, gd, gc, nf, n64 = _run_gpu(ctx, prob, grids, L.MIXED, thr=1e30)
    assert n64 > 0
    for a, b in zip(gc, oc):
        assert rel_l2(_cplx(a), _cplx(b)) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("gate", ["none", "all"])
def test_posterior_frequency_gate_extremes(ctx, gate):
    """The solve kernel runs over the in-band k-planes and a pass-through
    kernel over the others: with no plane in band every covariance is the
    cleaned prior and means and data are untouched (float, mixed and FP64);
    with every plane in band no pass-through launch happens (float against
    the float oracle on the same gate, failed factorisations included)."""
    print(f"posterior gate {gate}", flush=True)
    prob = synth.make_problem(synth.config("small"), "cpu")
    cfg = prob.cfg
    grids = fourier_inputs(prob)
    tabs = dict(problem_tables(prob))
    tabs["invert"] = np.full_like(tabs["invert"], 0 if gate == "none" else 1)
    mu, d, cv, ec = [x if not isinstance(x, list) else [g.copy() for g in x] for x in grids]
    ec = ec.copy()
    nf_ref = oracle.posterior(cfg.nxp, cfg.nyp, cfg.nzp, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0,
                              tabs, mu, d, cv, ec)
    for prec in ((L.F32, L.MIXED, L.F64) if gate == "none" else (L.F32,)):
        src = grids if prec != L.F64 else [as_f64(grids[0]), as_f64(grids[1]), as_f64(grids[2]),
                                           as_f64([grids[3]])[0]]
        tm = [torch.from_numpy(g.copy()).cuda() for g in src[0]]
        td = [torch.from_numpy(g.copy()).cuda() for g in src[1]]
        tc = [torch.from_numpy(g.copy()).cuda() for g in src[2]]
        te = torch.from_numpy(src[3].copy()).cuda()
        post = Posterior(ctx, synth.post_params(prob, 1e-3), tabs)
        nf, _ = post.run(prec, tm, td, tc, te, counts=True)
        torch.cuda.synchronize()
        post.close()
        if gate == "none":
            assert nf == 0 and nf_ref == 0
            for t, g in zip(tm + td, src[0] + src[1]):
                assert np.array_equal(t.cpu().numpy(), g)
        else:
            assert nf == nf_ref
        tol = 2e-5 if prec != L.F64 else 1e-6
        for t, b in zip(tm, mu):
            assert rel_l2(_cplx(t.cpu().numpy()).astype(np.complex128), _cplx(b)) < tol, prec
        for t, b in zip(tc, cv):
            assert rel_l2(_cplx(t.cpu().numpy()).astype(np.complex128), _cplx(b)) < tol, prec
