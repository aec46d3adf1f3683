"""Batched inversion of uneven volumes (crava_hip_batch_*): one persistent
posterior launch over every volume, per-volume transform sets, hipGraph
capture and replay, and the autotuner.

Checks:
  * small uneven volumes (odd and 7-smooth lengths, ntheta 3..8) against the
    oracle's forward FFT + float posterior loop + inverse FFT per volume;
  * the batch against the single-volume C-ABI path at the batching
    workload's sizes, for every posterior launch form: the transforms are the
    same kernels (bit for bit); the per-volume form runs the single-volume
    posterior kernel (bit for bit); the persistent forms run the same
    per-point code compiled into another kernel, where the compiler may fuse
    multiply-adds differently (observed: ulp-level differences at ntheta >= 6
    only);
  * graph replays, each from freshly reloaded inputs with eager work and
    synchronisations between them, bit for bit against the eager run of the
    same launch form;
  * the autotuner leaves the registered grids unchanged.
"""
import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd import synth
from crava_amd.core import Batch, Posterior, problem_tables
from tests.helpers import rel_l2


def _problems(n, seed, nxy, nzr):
    rng = np.random.default_rng(seed)
    return [synth.make_problem(synth.random_config(rng, i, seed=seed, nxy=nxy, nzr=nzr), "cpu") for i in range(n)]


class _Vol:
    """Device grids of one volume plus their pristine time-domain inputs."""

    def __init__(self, ctx, prob, thr, separable=False):
        self.prob = prob
        self.separable = separable
        self.src = [g.cuda() for g in prob.mu + prob.data + prob.cov + [prob.errcorr, prob.corr.to(prob.errcorr.dtype)]]
        self.grids = [g.clone() for g in self.src]
        self.post = Posterior(ctx, synth.post_params(prob, thr), problem_tables(prob))
        nt = prob.ntheta
        self.mu, self.data = self.grids[:3], self.grids[3:3 + nt]
        self.cov, self.ec, self.corr = self.grids[3 + nt:9 + nt], self.grids[9 + nt], self.grids[10 + nt]

    def reload(self):
        for d, s in zip(self.grids, self.src):
            d.copy_(s)

    def entry(self):
        return (self.post, self.mu, self.data, self.cov, self.ec) + ((self.corr,) if self.separable else ())

    def outputs(self):
        return [g.clone() for g in self.mu + self.data + self.cov]


def _single_path(ctx, v, precision):
    """The same sequence volume by volume through the single-volume entry
    points: the mixed form transforms the covariances forward with double
    arithmetic in each pass and the means with mean deflation, the rest in
    the grids' own precision."""
    c = v.prob.cfg
    n = (c.nxp, c.nyp, c.nzp)
    mc = L.MIXED if precision == L.MIXED else None
    md = L.F32_DEFLATE if precision == L.MIXED else None
    ctx.r2c_many(v.data, *n, L.DATA)
    ctx.r2c(v.ec, *n, L.COVARIANCE)
    ctx.r2c_many(v.mu, *n, L.DATA, md)
    ctx.r2c_many([v.corr] if v.separable else v.cov, *n, L.COVARIANCE, mc)
    nf, n64 = v.post.run(precision, v.mu, v.data, v.cov, v.ec, counts=True, corr=v.corr if v.separable else None)
    ctx.c2r_many(v.mu, *n, L.DATA, md)
    ctx.c2r_many(v.data, *n, L.DATA)
    ctx.c2r_many(v.cov, *n, L.COVARIANCE)
    return nf, n64


@pytest.mark.gpu
def test_batch_small_uneven_vs_oracle(ctx):
    print("batch: 6 small uneven volumes vs oracle", flush=True)
    probs = _problems(6, 11, (9, 40), (12, 48))
    vols = [_Vol(ctx, p, 1e-3) for p in probs]
    b = Batch(ctx, L.F32, [v.entry() for v in vols])
    b.run(graph=False)
    torch.cuda.synchronize()
    nf_gpu, _ = b.counts()
    for v, nf in zip(vols, nf_gpu):
        p, c = v.prob, v.prob.cfg
        n = (c.nxp, c.nyp, c.nzp)
        tabs = problem_tables(p)
        mu = [g.numpy().copy() for g in p.mu]
        d = [g.numpy().copy() for g in p.data]
        cv = [g.numpy().copy() for g in p.cov]
        ec = p.errcorr.numpy().copy()
        for g in mu + d:
            oracle.fft_r2c(g, *n, False)
        for g in cv + [ec]:
            oracle.fft_r2c(g, *n, True)
        nf_ref = oracle.posterior(*n, p.A, p.err_theta_cov, c.wnc, p.sigma0, tabs, mu, d, cv, ec)
        for g in mu + d:
            oracle.fft_c2r(g, *n, False)
        for g in cv:
            oracle.fft_c2r(g, *n, True)
        assert nf == nf_ref
        got = [g.cpu().numpy() for g in v.mu + v.data + v.cov]
        for a, r in zip(got, mu + d + cv):
            # float FFTs of two different algorithms plus the float loop (as the smoke check)
            assert rel_l2(a[:, :, : c.nxp], r[:, :, : c.nxp]) < 5e-5, (n, p.ntheta)
    b.close()
    for v in vols:
        v.post.close()


@pytest.mark.gpu
@pytest.mark.parametrize("precision,separable", [(L.MIXED, False), (L.F64, False), (L.MIXED, True), (L.F64, True)],
                         ids=["mixed", "f64", "mixed-separable", "f64-separable"])
def test_batch_matches_single_volume_path_and_graph_replays(ctx, precision, separable):
    print(f"batch vs single path, precision {precision}, separable {separable}", flush=True)
    probs = _problems(8, 5, (32, 192), (64, 256))
    if precision == L.F64:
        for p in probs:
            for lst in (p.mu, p.data, p.cov):
                lst[:] = [g.double() for g in lst]
            p.errcorr = p.errcorr.double()
    vols = [_Vol(ctx, p, 0.1, separable) for p in probs]
    b = Batch(ctx, precision, [v.entry() for v in vols])

    # reference: volume by volume through the single-volume entry points
    # (the same context, hence the same transform plans and variants)
    ref, ref_counts = [], []
    for v in vols:
        ref_counts.append(_single_path(ctx, v, precision))
        ref.append(v.outputs())
        v.reload()
    torch.cuda.synchronize()

    def rel(a, r):
        a, r = a.double(), r.double()
        return float(torch.linalg.vector_norm(a - r) / torch.linalg.vector_norm(r))

    forms = [(256, 0, Batch.FORM_PER_CLASS), (64, 512, Batch.FORM_PER_CLASS), (128, 2048, Batch.FORM_SINGLE),
             (256, 0, Batch.FORM_PER_VOLUME)]
    for ppb, blocks, form in forms:
        b.set_launch(ppb, blocks, form)
        b.run(graph=False)
        torch.cuda.synchronize()
        eager = [v.outputs() for v in vols]
        counts = b.counts()
        for i, v in enumerate(vols):
            assert (counts[0][i], counts[1][i]) == ref_counts[i], (ppb, blocks, form, i)
            for a, r in zip(eager[i], ref[i]):
                if form == Batch.FORM_PER_VOLUME:     # the single-volume kernel itself
                    assert torch.equal(a, r), (form, i)
                # rounding level: a dropped term or a wrong index is O(1e-2) or more
                assert rel(a, r) < (1e-12 if precision == L.F64 else 1e-6), (ppb, blocks, form, i)
        for rep in range(2):
            for v in vols:
                v.reload()
            torch.cuda.synchronize()
            b.run(graph=True)
            torch.cuda.synchronize()
            for i, v in enumerate(vols):
                for a, r in zip(v.outputs(), eager[i]):
                    assert torch.equal(a, r), (ppb, blocks, form, rep, i)
            assert b.counts() == counts
            _ = torch.randn(1000, device="cuda").sum().item()   # eager work between replays
        for v in vols:
            v.reload()
    b.close()
    for v in vols:
        v.post.close()


@pytest.mark.gpu
def test_batch_autotune_restores_grids(ctx):
    print("batch autotune", flush=True)
    probs = _problems(4, 23, (32, 96), (64, 128))
    vols = [_Vol(ctx, p, 0.1) for p in probs]
    b = Batch(ctx, L.MIXED, [v.entry() for v in vols])
    before = [[g.clone() for g in v.grids] for v in vols]
    ppb, blocks, form, ms = b.autotune()
    torch.cuda.synchronize()
    assert ppb in (64, 128, 256) and blocks >= 0 and form in (0, 1, 2) and ms > 0
    for v, bf in zip(vols, before):
        for a, r in zip(v.grids, bf):
            assert torch.equal(a, r)
    b.run(graph=True)
    torch.cuda.synchronize()
    for v in vols:
        for g in v.mu + v.data + v.cov:
            assert torch.isfinite(g).all()
    b.close()
    for v in vols:
        v.post.close()


@pytest.mark.gpu
def test_batch_rejects_bad_launch(ctx):
    probs = _problems(1, 3, (16, 24), (16, 24))
    v = _Vol(ctx, probs[0], 0.1)
    b = Batch(ctx, L.F32, [v.entry()])
    with pytest.raises(L.CravaHipError):
        b.set_launch(100, 0, 1)
    with pytest.raises(L.CravaHipError):
        b.set_launch(64, -1, 0)
    with pytest.raises(L.CravaHipError):
        b.set_launch(64, 0, 3)
    b.close()
    v.post.close()
