"""The non-constant-thickness inversion path on the GPU against the oracle:
the per-k tables, wavelet resampling and energy gain from the C ABI against
the oracle's restatements; then the whole path -- wavelet division
(crava_hip_divide_by_scale_wavelet), forward transforms, the float
posterior with the non-constant-thickness tables (K = D A), inverse
transforms -- against the oracle's division, transforms and per-point loop
on the same inputs (src/crava.cpp:1058-1083 branch of
computePostMeanResidAndFFTCov).

Tolerances: the tables, the resampled wavelet and the energy variances
share the oracle's float arithmetic; the spectra are double DFTs rounded to
float in both, summed in a different order, so they agree to 2e-7 of their
largest value (bit-exact elsewhere).

The path: division and forward transforms agree with the oracle's to 4e-7
and the posterior alone, on identical inputs, to 2e-7 of the FP64 oracle.
The residual r = E M^-1 (d - K mu) is ill-conditioned in the transform
roundoff here: e3 = 1/|W| spans 0.3-25 over the band, so E is far from a
multiple of the identity and E M^-1 amplifies the 1e-7 differences of two
float transforms by about 100.  The reference's own chain with FFTW 2.1.2 in
float (oracle/_ref) differs from the oracle's chain by 3.2e-5 (small case)
and 7.8e-5 (64x64x128) in the residuals; the GPU by 3.6e-5 and 6.9e-5.
Bounds: means 2e-6, covariances 4e-6, residuals 1.2e-4 relative L2, and
where oracle/_ref is built, the residual within 1.5x of that reference
spread.
"""
import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd import synth
from crava_amd.core import (Posterior, divide_by_scale_wavelet, perk_tables_nonconst, wavelet_energy,
                            wavelet_resample)
from tests.helpers import rel_l2


@pytest.mark.gpu
@pytest.mark.parametrize("nzp", [24, 90, 160])
def test_host_nonconst_tables_vs_oracle(nzp):
    rng = np.random.default_rng(nzp)
    nt = 4
    w = rng.standard_normal((nt, nzp)).astype(np.float32)
    w[3] = 0.0                                   # the vanishing-wavelet branch
    wn = (1.0 + rng.random(nt)).astype(np.float32)
    args = (nt, int(nzp * 0.8), nzp, 3.2 * nzp, 5.0, 55.0, 0.75)
    a = perk_tables_nonconst(*args, w, wn)
    b = oracle.perk_tables_nonconst(*args, w, wn)
    assert np.array_equal(a["invert"], b["invert"])
    for k in ("kWD", "e1", "e2", "e3"):
        d = np.abs(a[k][:, :3].astype(np.float64) - b[k][:, :3]).max()
        assert d <= 2e-7 * np.abs(b[k][:, :3]).max(), (k, d)
    assert np.array_equal(a["e1"], b["e1"]) and np.array_equal(a["kWD"], b["kWD"])
    # the zero wavelet: e3 = |w| nzp^2 100; e2 multiplies kD3 by 1/0 as the
    # reference does (inf or nan, by the roundoff-level sign of kD3's parts)
    assert np.array_equal(a["e3"][:, 3], b["e3"][:, 3])
    for t in (a, b):
        v = t["e2"][:, 3]
        assert not np.any(np.isfinite(v) & (v != 0))


@pytest.mark.gpu
@pytest.mark.parametrize("dzw,dz,nzp", [(4.0, 4.0, 64), (4.0, 1.7, 96), (2.0, 3.1, 50)])
def test_host_wavelet_resample_vs_oracle(dzw, dz, nzp):
    rng = np.random.default_rng(int(dz * 10))
    w = rng.standard_normal(41).astype(np.float32)
    got, gn = wavelet_resample(w, 20, dzw, dz, nzp)
    want, wn = oracle.wavelet_resample(w, 20, dzw, dz, nzp)
    assert np.array_equal(got, want) and gn == wn
    with pytest.raises(L.CravaHipError):
        wavelet_resample(np.ones(9, np.float32), 4, 0.01, 0.005, 16)


@pytest.mark.gpu
def test_host_wavelet_energy_vs_oracle():
    rng = np.random.default_rng(9)
    nt, nzp = 5, 72
    w = rng.standard_normal((nt, nzp)).astype(np.float32)
    A = rng.standard_normal((nt, 3)).astype(np.float32)
    s = rng.standard_normal((3, 3))
    s0 = s @ s.T + np.eye(3)
    ct = np.exp(-np.arange(nzp) / 5.0).astype(np.float32)
    ev = (0.1 + rng.random(nt)).astype(np.float32)
    sn = (2.0 + 4 * rng.random(nt)).astype(np.float32)
    for match in (None, np.array([1, 0, 1, 0, 1], np.uint8)):
        got = wavelet_energy(w, A, s0, ct, ev, sn, match)
        want = oracle.wavelet_energy(w, A, s0, ct, ev, sn, match)
        for k in want:
            assert np.array_equal(got[k], want[k]), k


def _cplx(g):
    return g.view(np.complex64)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small", "plain_cpu"])
def test_nonconst_thickness_path_vs_oracle(ctx, name):
    print(f"non-constant-thickness path {name}", flush=True)
    prob = synth.make_problem(synth.config(name), "cpu")
    cfg = prob.cfg
    nt = prob.ntheta
    n6 = (cfg.nx, cfg.ny, cfg.nz, cfg.nxp, cfg.nyp, cfg.nzp)
    n3 = n6[3:]
    args, kw = synth.swdiv_inputs(*n6, nt, seed=cfg.nxp + nt)
    args[0] = [g.numpy().copy() for g in prob.data]
    args[1], args[2], args[3], args[4] = prob.wavelets, prob.wnorm, prob.A, prob.sigma0
    args[8] = cfg.dt
    args[9] = cfg.nz * cfg.dt                   # lz
    args[10], args[11] = cfg.low_cut, cfg.high_cut
    args[12] = cfg.wnc
    lz, rel = args[9], args[7]
    targs = (nt, cfg.nz, cfg.nzp, lz, cfg.low_cut, cfg.high_cut, float(rel.min()), prob.wavelets, prob.wnorm)
    tabs = perk_tables_nonconst(*targs)
    otabs = oracle.perk_tables_nonconst(*targs)
    assert 0 < tabs["invert"].sum() < cfg.nzp

    # oracle: division, forward transforms, float per-point loop
    od = oracle.divide_by_scale_wavelet(*n6, *args, **kw)
    om = [g.numpy().copy() for g in prob.mu]
    oc = [g.numpy().copy() for g in prob.cov]
    oe = prob.errcorr.numpy().copy()
    for g in om + od:
        oracle.fft_r2c(g, *n3, False)
    for g in oc + [oe]:
        oracle.fft_r2c(g, *n3, True)
    nf_ref = oracle.posterior(*n3, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, otabs, om, od, oc, oe)

    # GPU
    td = [torch.from_numpy(g.copy()).cuda() for g in args[0]]
    divide_by_scale_wavelet(ctx, *n6, td, *args[1:], **kw)
    tm = [g.clone().cuda() for g in prob.mu]
    tc = [g.clone().cuda() for g in prob.cov]
    te = prob.errcorr.clone().cuda()
    ctx.r2c_many(tm + td, *n3, L.DATA)
    ctx.r2c_many(tc + [te], *n3, L.COVARIANCE)
    post = Posterior(ctx, synth.post_params(prob, 0.0), tabs)
    nf, _ = post.run(L.F32, tm, td, tc, te, counts=True)
    post.close()
    torch.cuda.synchronize()
    assert nf == nf_ref
    e = {"mean": max(rel_l2(_cplx(a.cpu().numpy()), _cplx(b)) for a, b in zip(tm, om)),
         "residual": max(rel_l2(_cplx(a.cpu().numpy()), _cplx(b)) for a, b in zip(td, od)),
         "cov": max(rel_l2(_cplx(a.cpu().numpy()), _cplx(b)) for a, b in zip(tc, oc))}
    print(f"  errors vs oracle {e}", flush=True)
    assert e["mean"] < 2e-6
    assert e["residual"] < 1.2e-4
    assert e["cov"] < 4e-6
    if oracle.ref_available():
        rd = oracle.divide_by_scale_wavelet(*n6, *args, **kw, which="ref")
        rm = [g.numpy().copy() for g in prob.mu]
        rc = [g.numpy().copy() for g in prob.cov]
        re_ = prob.errcorr.numpy().copy()
        for g in rm + rd:
            oracle.fft_r2c(g, *n3, False, "ref")
        for g in rc + [re_]:
            oracle.fft_r2c(g, *n3, True, "ref")
        oracle.posterior(*n3, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, otabs, rm, rd, rc, re_, which="ref")
        spread = max(rel_l2(_cplx(a), _cplx(b)) for a, b in zip(rd, od))
        print(f"  reference FFTW chain vs oracle chain: residual {spread:.3g}", flush=True)
        assert e["residual"] < 1.5 * spread
