"""GPU posterior simulation and device noise against the oracle."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd import synth
from crava_amd._lib import check, lib, ptr_array
from crava_amd.core import problem_tables
from tests.helpers import fourier_inputs, rel_l2


def _posterior_cov(name="small"):
    prob = synth.make_problem(synth.config(name))
    cfg = prob.cfg
    tabs = problem_tables(prob)
    mu, d, cv, ec = fourier_inputs(prob)
    oracle.posterior(cfg.nxp, cfg.nyp, cfg.nzp, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, tabs, mu, d, cv, ec)
    return prob, mu, cv


def _ptrs(ts):
    return ptr_array([int(t.data_ptr()) for t in ts])


@pytest.mark.gpu
def test_simulate_points_vs_oracle(ctx):
    print("simulate points", flush=True)
    prob, mu, cv = _posterior_cov()
    cfg = prob.cfg
    seeds = [oracle.fill_complex_noise(cfg.nxp, cfg.nyp, cfg.nzp, 23665 + s)[0] for s in range(3)]
    cv_bad = [g.copy() for g in cv]
    cv_bad[1].view(np.complex64)[2, 3, 4] = -1.0      # not positive definite -> zero seed there
    ref_seed = [s.copy() for s in seeds]
    nf_ref = oracle.simulate_points([g.view(np.complex64).reshape(-1).copy() for g in cv_bad],
                                    [s.view(np.complex64).reshape(-1) for s in ref_seed])
    tc = [torch.from_numpy(g.copy()).cuda() for g in cv_bad]
    ts = [torch.from_numpy(s.copy()).cuda() for s in seeds]
    nf = C.c_longlong(0)
    npts = (cfg.nxp // 2 + 1) * cfg.nyp * cfg.nzp
    check(lib().crava_hip_simulate_points(ctx.h, npts, _ptrs(tc), _ptrs(ts), C.byref(nf)), "simulate_points")
    torch.cuda.synchronize()
    assert nf.value == nf_ref >= 1
    for a, b in zip(ts, ref_seed):
        got = a.cpu().numpy().view(np.complex64)
        assert rel_l2(got, b.view(np.complex64)) < 2e-6
        assert got[2, 3, 4] == 0


@pytest.mark.gpu
def test_simulate_full_realisation_vs_oracle(ctx):
    print("simulate realisation", flush=True)
    prob, mu, cv = _posterior_cov()
    cfg = prob.cfg
    n = (cfg.nxp, cfg.nyp, cfg.nzp)
    seeds = [oracle.fill_complex_noise(*n, 7 + s)[0] for s in range(3)]
    mean_t = [g.copy() for g in mu]
    for g in mean_t:
        oracle.fft_c2r(g, *n, False)
    ref = [s.copy() for s in seeds]
    oracle.simulate_points([g.view(np.complex64).reshape(-1).copy() for g in cv],
                           [s.view(np.complex64).reshape(-1) for s in ref])
    for s, m in zip(ref, mean_t):
        oracle.fft_c2r(s, *n, False)
        s[:, :, : cfg.nxp] += m[:, :, : cfg.nxp]
    tc = [torch.from_numpy(g.copy()).cuda() for g in cv]
    ts = [torch.from_numpy(s.copy()).cuda() for s in seeds]
    tm = [torch.from_numpy(m.copy()).cuda() for m in mean_t]
    check(lib().crava_hip_simulate(ctx.h, *n, _ptrs(tc), _ptrs(ts), _ptrs(tm), None), "simulate")
    torch.cuda.synchronize()
    for a, b in zip(ts, ref):
        assert rel_l2(a.cpu().numpy()[:, :, : cfg.nxp], b[:, :, : cfg.nxp]) < 2e-6


@pytest.mark.gpu
def test_device_noise_statistics_and_symmetry(ctx):
    print("device noise", flush=True)
    nxp, nyp, nzp = 64, 48, 40
    cn = nxp // 2 + 1
    g = torch.zeros((nzp, nyp, 2 * cn), device="cuda")
    check(lib().crava_hip_fill_complex_noise(ctx.h, int(g.data_ptr()), nxp, nyp, nzp, 1234, 0), "noise")
    h = torch.zeros_like(g)
    check(lib().crava_hip_fill_complex_noise(ctx.h, int(h.data_ptr()), nxp, nyp, nzp, 1234, 0), "noise")
    k = torch.zeros_like(g)
    check(lib().crava_hip_fill_complex_noise(ctx.h, int(k.data_ptr()), nxp, nyp, nzp, 1234, 1), "noise")
    torch.cuda.synchronize()
    z = g.cpu().numpy().view(np.complex64)
    assert np.array_equal(g.cpu().numpy(), h.cpu().numpy())          # deterministic
    assert not np.array_equal(g.cpu().numpy(), k.cpu().numpy())      # streams differ
    inner = z[:, :, 1:cn - 1]
    assert abs(inner.real.var() - 0.5) < 0.01 and abs(inner.imag.var() - 0.5) < 0.01
    assert abs(np.mean(inner.real * inner.imag)) < 0.005
    for i in (0, cn - 1):                                            # Hermitian planes
        for kk in range(nzp):
            for j in range(nyp):
                assert z[kk, j, i] == np.conj(z[(nzp - kk) % nzp, (nyp - j) % nyp, i])
    # the realisation of a real white field: c2r gives unit variance per cell
    check(lib().crava_hip_fft_c2r(ctx.h, int(g.data_ptr()), nxp, nyp, nzp, L.PARAMETER, L.F32), "c2r")
    torch.cuda.synchronize()
    x = g.cpu().numpy()[:, :, :nxp]
    assert abs(x.var() - 1.0) < 0.02 and abs(x.mean()) < 0.01


@pytest.mark.gpu
def test_simulated_draws_have_posterior_covariance(ctx):
    """Draws L z with device noise, whitened by numpy's own Cholesky factor of
    each point's Sigma, have identity covariance (statistical, 5 sigma)."""
    print("draw covariance", flush=True)
    rng = np.random.default_rng(5)
    nxp, nyp, nzp = 30, 20, 16            # interior x bins are ordinary CN(0,1) points
    cn = nxp // 2 + 1
    shape = (nzp, nyp, cn)
    B = rng.standard_normal(shape + (3, 3)) + 1j * rng.standard_normal(shape + (3, 3))
    Sig = B @ np.conj(np.swapaxes(B, -1, -2)) + 0.5 * np.eye(3)
    Sig *= np.exp(rng.uniform(-3, 3, shape))[..., None, None]
    order = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    cov = [np.ascontiguousarray(Sig[..., a, b].astype(np.complex64)) for a, b in order]
    tc = [torch.from_numpy(g.view(np.float32).copy()).cuda() for g in cov]
    Lnp = np.linalg.cholesky(Sig.astype(np.complex64).astype(np.complex128))[:, :, 1:cn - 1]
    acc = np.zeros((3, 3), dtype=np.complex128)
    reps, count = 30, 0
    for r in range(reps):
        ts = [torch.zeros_like(tc[0]) for _ in range(3)]
        for c in range(3):
            check(lib().crava_hip_fill_complex_noise(ctx.h, int(ts[c].data_ptr()), nxp, nyp, nzp, 99, 3 * r + c),
                  "noise")
        check(lib().crava_hip_simulate_points(ctx.h, nzp * nyp * cn, _ptrs(tc), _ptrs(ts), None), "simulate_points")
        torch.cuda.synchronize()
        s = np.stack([t.cpu().numpy().view(np.complex64)[:, :, 1:cn - 1] for t in ts], axis=-1).astype(np.complex128)
        w = np.linalg.solve(Lnp, s[..., None])[..., 0].reshape(-1, 3)
        acc += w.T @ np.conj(w)
        count += w.shape[0]
    acc /= count
    assert np.abs(acc - np.eye(3)).max() < 5 * 1.0 / np.sqrt(count)
