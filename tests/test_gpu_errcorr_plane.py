"""The error-correlation spectrum as one Fourier plane
(crava_hip_err_corr_plane, crava_hip_post_set_errcorr_plane).

With zero correlation gradients FFTGrid::fillInErrCorr (src/fftgrid.cpp:
1065-1116) puts priorCorrXY in plane k = 0 and zeros elsewhere, so the
grid's unscaled 3-D transform is the 2-D transform of corrXY on every kz
plane.  Checks:
  * the plane against numpy's float64 2-D transform of corrXY (the
    mathematics: a delta along z transforms to a constant) and against every
    kz plane of the GPU's 3-D transform of the filled grid, F32 / MIXED / F64,
    odd and even lengths, a one-plane grid;
  * the posterior (F32, MIXED with FP64 waves, F64, separable), the
    likelihood and the batch (every launch form, graph replay) with the plane
    against the same calls reading a grid that holds the plane on every kz
    plane: bit for bit, since the same values are read;
  * a batch refuses to run after a volume's form changed.
"""
import numpy as np
import pytest
import torch

from crava_amd import _lib as L
from crava_amd import synth
from crava_amd.autograd import nll_grad
from crava_amd.core import Batch, Posterior, err_corr_plane, fill_err_corr, problem_tables
from tests.helpers import fourier_inputs


def _corr_xy(nxp, nyp, seed):
    rng = np.random.default_rng(seed)
    return rng.uniform(-0.2, 1.0, size=(nyp, nxp)).astype(np.float32)


def _plane(ctx, cxy, nxp, nyp, nzp, precision):
    dt = torch.float64 if precision == L.F64 else torch.float32
    p = torch.full((nyp, 2 * (nxp // 2 + 1)), 7.0, dtype=dt, device="cuda")
    err_corr_plane(ctx, p, nxp, nyp, nzp, torch.from_numpy(cxy).cuda(), precision)
    return p


def _cplx(t):
    a = t.cpu().numpy()
    return a.view(np.complex128 if a.dtype == np.float64 else np.complex64)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(80, 64, 160), (15, 12, 24), (9, 7, 5), (4, 7, 1)])
@pytest.mark.parametrize("precision", [L.F32, L.MIXED, L.F64], ids=["f32", "mixed", "f64"])
def test_plane_equals_grid_transform_and_numpy(ctx, n, precision):
    nxp, nyp, nzp = n
    print(f"errcorr plane {n} precision {precision}", flush=True)
    cxy = _corr_xy(nxp, nyp, nxp * 1000 + nyp)
    plane = _cplx(_plane(ctx, cxy, nxp, nyp, nzp, precision))
    ref = np.fft.rfft2(cxy.astype(np.float64))
    scale = np.abs(ref).max()
    tol = 1e-13 if precision == L.F64 else 2e-6
    assert plane.shape == ref.shape
    assert np.abs(plane - ref).max() <= tol * scale
    # the whole grid filled and transformed in 3-D: every kz plane is the plane
    grid = torch.zeros((nzp, nyp, 2 * (nxp // 2 + 1)), dtype=torch.float32, device="cuda")
    fill_err_corr(ctx, grid, nxp, nyp, nzp, torch.from_numpy(cxy).cuda())
    if precision == L.F64:
        grid = grid.double()
    ctx.r2c(grid, nxp, nyp, nzp, L.COVARIANCE, precision)
    g = _cplx(grid)
    for k in range(nzp):
        assert np.abs(g[k] - ref).max() <= tol * scale, k


def _posterior_pair(ctx, prob, grids, precision, thr, separable):
    """The posterior with the plane and with the plane on every kz plane."""
    cfg = prob.cfg
    n = (cfg.nxp, cfg.nyp, cfg.nzp)
    plane = _plane(ctx, np.ascontiguousarray(prob.corr_xy, dtype=np.float32), *n, precision)
    full = plane.unsqueeze(0).expand(cfg.nzp, -1, -1).contiguous()
    dt = np.float64 if precision == L.F64 else np.float32
    mu, d, cv, _ = grids
    out = []
    for ec, flat in ((full, False), (plane, True)):
        up = lambda g: torch.from_numpy(np.ascontiguousarray(g.view(np.complex64).astype(
            np.complex128 if dt == np.float64 else np.complex64)).view(dt)).cuda()
        tm, td, tc = [up(g) for g in mu], [up(g) for g in d], [up(g) for g in cv]
        post = Posterior(ctx, synth.post_params(prob, thr), problem_tables(prob))
        post.set_errcorr_plane(flat)
        corr = tc[0].clone() if separable else None
        counts = post.run(precision, tm, td, tc, ec, counts=True, corr=corr)
        torch.cuda.synchronize()
        post.close()
        out.append((tm + td + tc, counts))
    return out


@pytest.mark.gpu
@pytest.mark.parametrize("precision,thr,separable", [(L.F32, 1e-3, False), (L.MIXED, 1e-3, False),
                                                     (L.MIXED, 1e30, False), (L.F64, 1e-3, False),
                                                     (L.MIXED, 1e-3, True)],
                         ids=["f32", "mixed", "mixed-fp64-waves", "f64", "mixed-separable"])
def test_posterior_with_plane_matches_grid(ctx, precision, thr, separable):
    print(f"posterior with errcorr plane, precision {precision} thr {thr} separable {separable}", flush=True)
    prob = synth.make_problem(synth.config("small"), "cpu")
    grids = fourier_inputs(prob)
    (a, ca), (b, cb) = _posterior_pair(ctx, prob, grids, precision, thr, separable)
    assert ca == cb
    if precision == L.MIXED and thr > 1.0:
        assert ca[1] > 0   # every wave took the FP64 solve
    for x, y in zip(a, b):
        assert torch.equal(x, y)


@pytest.mark.gpu
def test_likelihood_with_plane_matches_grid(ctx):
    print("nll with errcorr plane", flush=True)
    from tests.test_gpu_nll import _setup
    from tests.test_nll_oracle import theta0
    prob, gpu, _ = _setup(ctx, "small", True)
    post, prec, mu, d, cv, ec = gpu
    plane = ec[0].clone()
    full = plane.unsqueeze(0).expand(ec.shape[0], -1, -1).contiguous()
    th = [t.tolist() for t in theta0(prob)]
    v_full, g_full = nll_grad(post, prec, mu, d, cv, full, *th)
    post.set_errcorr_plane(True)
    v_plane, g_plane = nll_grad(post, prec, mu, d, cv, plane, *th)
    post.close()
    assert v_full == v_plane
    assert g_full == g_plane


class _Vol:
    """One volume's device grids (time domain) and its error-correlation plane."""

    def __init__(self, ctx, prob, precision, thr):
        self.prob = prob
        dt = torch.float64 if precision == L.F64 else torch.float32
        self.src = [g.to(dt).cuda() for g in prob.mu + prob.data + prob.cov]
        self.grids = [g.clone() for g in self.src]
        nt = prob.ntheta
        self.mu, self.data, self.cov = self.grids[:3], self.grids[3:3 + nt], self.grids[3 + nt:]
        c = prob.cfg
        self.plane = _plane(ctx, np.ascontiguousarray(prob.corr_xy, dtype=np.float32), c.nxp, c.nyp, c.nzp,
                            precision)
        self.post = Posterior(ctx, synth.post_params(prob, thr), problem_tables(prob))
        self.post.set_errcorr_plane(True)

    def reload(self):
        for d, s in zip(self.grids, self.src):
            d.copy_(s)

    def outputs(self):
        return [g.clone() for g in self.grids]


@pytest.mark.gpu
@pytest.mark.parametrize("precision", [L.MIXED, L.F64], ids=["mixed", "f64"])
def test_batch_with_planes_matches_single_path(ctx, precision):
    print(f"batch with errcorr planes, precision {precision}", flush=True)
    rng = np.random.default_rng(31)
    probs = [synth.make_problem(synth.random_config(rng, i, seed=31, nxy=(24, 80), nzr=(32, 96)), "cpu")
             for i in range(5)]
    vols = [_Vol(ctx, p, precision, 0.1) for p in probs]
    b = Batch(ctx, precision, [(v.post, v.mu, v.data, v.cov, v.plane) for v in vols])
    mc = L.MIXED if precision == L.MIXED else None
    md = L.F32_DEFLATE if precision == L.MIXED else None
    ref = []
    for v in vols:
        c = v.prob.cfg
        n = (c.nxp, c.nyp, c.nzp)
        ctx.r2c_many(v.data, *n, L.DATA)
        ctx.r2c_many(v.mu, *n, L.DATA, md)
        ctx.r2c_many(v.cov, *n, L.COVARIANCE, mc)
        v.post.run(precision, v.mu, v.data, v.cov, v.plane)
        ctx.c2r_many(v.mu, *n, L.DATA, md)
        ctx.c2r_many(v.data, *n, L.DATA)
        ctx.c2r_many(v.cov, *n, L.COVARIANCE)
        ref.append(v.outputs())
        v.reload()
    torch.cuda.synchronize()
    planes = [v.plane.clone() for v in vols]
    for ppb, blocks, form in [(128, 0, Batch.FORM_PER_VOLUME), (128, 1024, Batch.FORM_SINGLE),
                              (256, 0, Batch.FORM_PER_CLASS)]:
        b.set_launch(ppb, blocks, form)
        for graph in (False, True):
            for v in vols:
                v.reload()
            torch.cuda.synchronize()
            b.run(graph=graph)
            torch.cuda.synchronize()
            for i, v in enumerate(vols):
                assert torch.equal(v.plane, planes[i])   # read only
                for a, r in zip(v.outputs(), ref[i]):
                    if form == Batch.FORM_PER_VOLUME:
                        assert torch.equal(a, r), (form, graph, i)
                    rel = float(torch.linalg.vector_norm((a - r).double()) / torch.linalg.vector_norm(r.double()))
                    assert rel < (1e-12 if precision == L.F64 else 1e-6), (form, graph, i)
    # the batch was built for planes: a volume switched back to grids is refused
    vols[0].post.set_errcorr_plane(False)
    with pytest.raises(RuntimeError, match="error-correlation form changed"):
        b.run(graph=False)
    vols[0].post.set_errcorr_plane(True)
    b.close()
    for v in vols:
        v.post.close()


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(80, 64, 160), (15, 12, 24), (9, 7, 5)])
@pytest.mark.parametrize("precision", [L.MIXED, L.F64], ids=["mixed", "f64"])
def test_separable_corr_spectrum_equals_fill_and_transform(ctx, n, precision):
    """crava_hip_corr_spectrum_separable against fillInParamCorr (zero
    gradients) + the 3-D transform on the GPU, and against numpy's float64
    3-D transform of corrXY(i, j) circT(k)."""
    from crava_amd.core import corr_spectrum_separable, fill_param_corr
    nxp, nyp, nzp = n
    print(f"separable correlation spectrum {n} precision {precision}", flush=True)
    cxy = _corr_xy(nxp, nyp, nzp)
    ct = np.random.default_rng(nzp + 7).uniform(-0.3, 1.0, nzp).astype(np.float32)
    cxy_d, ct_d = torch.from_numpy(cxy).cuda(), torch.from_numpy(ct).cuda()
    rnxp = 2 * (nxp // 2 + 1)
    g = torch.zeros((nzp, nyp, rnxp), dtype=torch.float32, device="cuda")
    fill_param_corr(ctx, [g], [1.0], nxp, nyp, nzp, cxy_d, ct_d)
    if precision == L.F64:
        g = g.double()
    ctx.r2c(g, nxp, nyp, nzp, L.COVARIANCE, precision)
    out = torch.full((nzp, nyp, rnxp), 5.0, dtype=g.dtype, device="cuda")
    corr_spectrum_separable(ctx, out, nxp, nyp, nzp, cxy_d, ct_d, precision)
    a, r = _cplx(out), _cplx(g)
    exact = np.fft.rfftn(ct.astype(np.float64)[:, None, None] * cxy.astype(np.float64)[None, :, :])
    scale = np.abs(exact).max()
    # the fill rounds each product to float: the grid path differs from the
    # exact product of the two transforms by about eps_float / sqrt(N)
    assert np.abs(a - r).max() <= (2e-6 if precision == L.MIXED else 1e-7) * scale
    assert np.abs(a - exact).max() <= (2e-6 if precision == L.MIXED else 1e-12) * scale


@pytest.mark.gpu
def test_posterior_on_separable_corr_spectrum(ctx):
    """The separable mixed posterior on the factor spectrum and on the
    filled and transformed correlation grid agree to the FP64 bar."""
    from crava_amd.core import corr_spectrum_separable, fill_param_corr
    print("posterior on the separable correlation spectrum", flush=True)
    prob = synth.make_problem(synth.config("small"), "cpu")
    c = prob.cfg
    n = (c.nxp, c.nyp, c.nzp)
    mu, d, cv, ec = fourier_inputs(prob)
    cxy = torch.from_numpy(np.ascontiguousarray(prob.corr_xy, dtype=np.float32)).cuda()
    ct = torch.from_numpy(np.ascontiguousarray(prob.circ_corr_t, dtype=np.float32)).cuda()
    g = torch.zeros((c.nzp, c.nyp, 2 * (c.nxp // 2 + 1)), dtype=torch.float32, device="cuda")
    fill_param_corr(ctx, [g], [1.0], *n, cxy, ct)
    ctx.r2c(g, *n, L.COVARIANCE, L.MIXED)
    f = torch.empty_like(g)
    corr_spectrum_separable(ctx, f, *n, cxy, ct, L.MIXED)
    outs = []
    for corr in (g, f):
        tm = [torch.from_numpy(x.copy()).cuda() for x in mu]
        td = [torch.from_numpy(x.copy()).cuda() for x in d]
        tc = [torch.from_numpy(x.copy()).cuda() for x in cv]
        post = Posterior(ctx, synth.post_params(prob, 1e-3), problem_tables(prob))
        post.run(L.MIXED, tm, td, tc, torch.from_numpy(ec.copy()).cuda(), corr=corr)
        torch.cuda.synchronize()
        post.close()
        outs.append(tm + td + tc)
    for a, r in zip(*outs):
        a64, r64 = a.double(), r.double()
        assert float(torch.linalg.vector_norm(a64 - r64) / torch.linalg.vector_norm(r64)) < 1e-5
