"""crava_hip_post_run_static: the posterior entry point that allocates nothing
(its mixed-form queue lives in the library's static device array) against
crava_hip_post_run, bit for bit, on two calls in a row -- the second on a
smaller problem -- with thresholds that queue some waves and then every
wave for the FP64 kernel; and its refusal of a problem whose queue would not
fit the array."""
import ctypes as C

import numpy as np
import pytest
import torch

from crava_amd import _lib as L
from crava_amd import synth
from crava_amd._lib import lib, ptr_array
from crava_amd.core import Posterior, problem_tables
from tests.helpers import fourier_inputs


def _grids(prob):
    mu, d, cv, ec = fourier_inputs(prob)
    return [torch.from_numpy(g.copy()).cuda() for g in mu], [torch.from_numpy(g.copy()).cuda() for g in d], \
        [torch.from_numpy(g.copy()).cuda() for g in cv], torch.from_numpy(ec.copy()).cuda()


def _run(fn, post, grids):
    tm, td, tc, te = [[t.clone() for t in g] if isinstance(g, list) else g.clone() for g in grids]
    nf, n64 = C.c_longlong(-1), C.c_longlong(-1)
    rc = fn(post.h, L.MIXED, ptr_array([int(t.data_ptr()) for t in tm]), ptr_array([int(t.data_ptr()) for t in td]),
            ptr_array([int(t.data_ptr()) for t in tc]), int(te.data_ptr()), C.byref(nf), C.byref(n64))
    L.check(rc, fn.__name__)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in tm + td + tc], nf.value, n64.value


def _case(ctx, cfg, thr):
    prob = synth.make_problem(cfg, "cpu")
    grids = _grids(prob)
    post = Posterior(ctx, synth.post_params(prob, thr), problem_tables(prob))
    ref, nf0, n640 = _run(lib().crava_hip_post_run, post, grids)
    got, nf1, n641 = _run(lib().crava_hip_post_run_static, post, grids)
    post.close()
    assert (nf1, n641) == (nf0, n640)
    for a, b in zip(got, ref):
        assert np.array_equal(a.view(np.uint32), b.view(np.uint32))
    return n640


@pytest.mark.gpu
def test_static_queue_matches_allocated_queue_two_calls(ctx):
    print("no-alloc posterior: 64x64x320 precision regime, then the small grid", flush=True)
    big = synth.config("precision")
    big.nx = big.ny = 56
    big.nxp = big.nyp = 64
    n64 = _case(ctx, big, 3e-2)
    assert n64 > 0                      # some waves went through the static queue
    n64 = _case(ctx, synth.config("small"), 1e30)
    assert n64 > 0                      # every in-band wave, smaller problem, same array


@pytest.mark.gpu
def test_static_queue_rejects_a_problem_too_large(ctx):
    cfg = synth.config("precision")
    cfg.nx = cfg.ny = 1800
    cfg.nxp = cfg.nyp = 2048
    prob = synth.make_problem(cfg, "cpu", grids=False)
    post = Posterior(ctx, synth.post_params(prob, 1e-3), problem_tables(prob))
    dummy = torch.zeros(16, device="cuda")
    p = int(dummy.data_ptr())
    rc = lib().crava_hip_post_run_static(post.h, L.MIXED, ptr_array([p] * 3), ptr_array([p] * 6), ptr_array([p] * 6),
                                         p, None, None)
    post.close()
    assert rc == L.E_INVALID
    assert b"static queue" in lib().crava_hip_last_error()
