"""GPU likelihood and adjoint gradient against the FP64 oracle, finite
differences and an L-BFGS parameter recovery."""
import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd import synth
from crava_amd.autograd import NLLFunction, nll_grad
from crava_amd.core import Posterior, problem_tables
from oracle import nll_torch
from tests.test_nll_oracle import _fourier64, theta0


def _setup(ctx, name, dbl, data_override=None):
    prob = synth.make_problem(synth.config(name))
    cfg = prob.cfg
    tabs = problem_tables(prob)
    mu, d, cv, ec = _fourier64(prob)
    if data_override is not None:
        d = data_override(prob, tabs, mu, cv, ec)
    dt = np.complex128 if dbl else np.complex64
    to_t = lambda g: torch.from_numpy(np.ascontiguousarray(g.astype(dt)).view(np.float64 if dbl else np.float32)).cuda()
    gpu = (Posterior(ctx, synth.post_params(prob), tabs), L.F64 if dbl else L.F32,
           [to_t(g) for g in mu], [to_t(g) for g in d], [to_t(g) for g in cv], to_t(ec))
    # the oracle sees exactly the values the GPU sees
    rt = lambda g: g.astype(dt).astype(np.complex128)
    P = nll_torch.prepare(tabs, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0[0, 0],
                          [rt(g) for g in mu], [rt(g) for g in d], rt(cv[0]), rt(ec), cfg.nxp)
    return prob, gpu, P


@pytest.mark.gpu
@pytest.mark.parametrize("dbl", [True, False])
def test_nll_value_and_gradient_vs_oracle(ctx, dbl):
    print(f"nll dbl={dbl}", flush=True)
    prob, gpu, P = _setup(ctx, "small", dbl)
    s, v, l0 = theta0(prob)
    s = s * torch.tensor([1.05, 0.9, 1.1, 0.97], dtype=torch.float64)
    v = v * 1.3
    th = [t.clone().requires_grad_(True) for t in (s, v, l0)]
    ref = nll_torch.nll(P, *th)
    ref.backward()
    val, grad = nll_grad(*gpu, s.tolist(), v.tolist(), l0.tolist())
    assert abs(val - float(ref)) <= 1e-9 * abs(float(ref))
    gref = torch.cat([t.grad for t in th]).numpy()
    assert np.abs(np.array(grad) - gref).max() <= 1e-7 * np.abs(gref).max()
    gpu[0].close()


@pytest.mark.gpu
def test_nll_gradient_vs_central_differences(ctx):
    """FP64 central differences (h = 1e-6 relative) of the GPU likelihood."""
    print("nll fd", flush=True)
    prob, gpu, _ = _setup(ctx, "small", True)
    s, v, l0 = [t.tolist() for t in theta0(prob)]
    _, grad = nll_grad(*gpu, s, v, l0)
    x = s + v + l0
    nt = prob.ntheta
    for i in range(len(x)):
        h = 1e-6 * max(abs(x[i]), 1e-12)
        f = []
        for sg in (1, -1):
            y = list(x)
            y[i] += sg * h
            f.append(nll_grad(*gpu, y[:nt], y[nt:2 * nt], y[2 * nt:])[0])
        fd = (f[0] - f[1]) / (2 * h)
        scale = max(abs(grad[i]), 1e-3 * max(abs(g) * abs(xx) for g, xx in zip(grad, x)) / max(abs(x[i]), 1e-30))
        assert abs(fd - grad[i]) <= 2e-5 * scale, (i, fd, grad[i])
    gpu[0].close()


@pytest.mark.gpu
def test_autograd_function_backward(ctx):
    print("nll autograd", flush=True)
    prob, gpu, _ = _setup(ctx, "small", True)
    th = [t.clone().requires_grad_(True) for t in theta0(prob)]
    f = NLLFunction.apply(*th, gpu)
    (2.0 * f).backward()
    _, grad = nll_grad(*gpu, *[t.tolist() for t in theta0(prob)])
    got = torch.cat([t.grad for t in th]).numpy()
    assert np.allclose(got, 2.0 * np.array(grad), rtol=1e-12, atol=0)
    gpu[0].close()


@pytest.mark.gpu
def test_lbfgs_recovers_perturbed_parameters(ctx):
    """Data drawn from the model at known parameters.  (a) Wavelet scales and
    noise variances perturbed by 10-100 %, Sigma0 held at the truth: 50 L-BFGS
    iterations recover both to a few percent.  (b) All 16 parameters free
    from a perturbed start: the fit reaches at least the likelihood of the
    truth (scales and Sigma0 trade off through s_l s_m Sigma0) and the noise
    variances to a few percent."""
    from crava_amd.autograd import nll as nll_fn
    from crava_amd.estimate import fit
    print("lbfgs", flush=True)
    truth = {}

    def draw(prob, tabs, mu, cv, ec):
        s, v, l0 = [t.numpy() for t in theta0(prob)]
        truth.update(s=s, v=v, l0=l0)
        return synth.model_data(prob, tabs, mu, cv[0], ec, s, v, l0, seed=3)

    prob, gpu, _ = _setup(ctx, "plain_cpu", True, data_override=draw)
    nt = prob.ntheta
    s0 = truth["s"] * np.array([1.15, 0.85, 1.1])[:nt]
    v0 = truth["v"] * 2.0
    rel = lambda a, b: np.abs(a / b - 1).max()
    free = [1.0] * (2 * nt) + [0.0] * 6
    s, v, c, trace = fit(gpu, nt, s0, v0, truth["l0"], steps=50, free=free)
    assert trace[-1][0] < trace[0][0]
    assert rel(s.numpy(), truth["s"]) < 0.03
    assert rel(v.numpy(), truth["v"]) < 0.03
    f_truth = float(nll_fn(*[torch.from_numpy(np.asarray(x, np.float64)) for x in (truth["s"], truth["v"], truth["l0"])], gpu))
    s, v, c, trace = fit(gpu, nt, s0, v0, truth["l0"] * 1.2, steps=50)
    assert trace[-1][0] <= f_truth + 1e-6 * abs(f_truth)
    assert rel(v.numpy(), truth["v"]) < 0.05
    gpu[0].close()
