"""GPU grid fills, element-wise operators, data variance and facies
geomodel update against the oracle: bit-exact (the kernels reproduce the
reference's float operations one by one)."""
import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd import synth
from crava_amd.core import (data_variance, facies_geomodel, fill_err_corr, fill_param_corr, grid_op)

SIZES = [(9, 6, 8), (80, 64, 50), (105, 45, 98), (512, 512, 320)]


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES, ids=lambda n: "x".join(map(str, n)))
def test_fill_param_corr_bitexact(ctx, n):
    nxp, nyp, nzp = n
    print(f"fill_param_corr {n}", flush=True)
    rng = np.random.default_rng(1)
    cxy = rng.uniform(0.0, 1.0, (nyp, nxp)).astype(np.float32)
    ct = rng.uniform(-0.2, 1.0, nzp).astype(np.float32)
    rnxp = 2 * (nxp // 2 + 1)
    scales = [0.0064, 0.0144, 0.0016, -0.00672, 0.0016, 0.00192]
    for gi, gj in ((0.0, 0.0), (0.37, -0.21)):
        grids = [torch.empty((nzp, nyp, rnxp), dtype=torch.float32, device="cuda") for _ in scales]
        fill_param_corr(ctx, grids, scales, nxp, nyp, nzp, torch.from_numpy(cxy).cuda(), torch.from_numpy(ct).cuda(),
                        gi, gj)
        torch.cuda.synchronize()
        for g, s in zip(grids, scales):
            ref = oracle.fill_param_corr(nxp, nyp, nzp, cxy, ct, gi, gj, np.float32(s))
            assert np.array_equal(g.cpu().numpy(), ref), (gi, gj, s)


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES[:3], ids=lambda n: "x".join(map(str, n)))
def test_fill_err_corr_bitexact(ctx, n):
    nxp, nyp, nzp = n
    rng = np.random.default_rng(2)
    cxy = rng.uniform(0.0, 1.0, (nyp, nxp)).astype(np.float32)
    rnxp = 2 * (nxp // 2 + 1)
    for gi, gj in ((0.0, 0.0), (0.04, -0.03), (0.3, 0.2)):
        g = torch.empty((nzp, nyp, rnxp), dtype=torch.float32, device="cuda")
        fill_err_corr(ctx, g, nxp, nyp, nzp, torch.from_numpy(cxy).cuda(), gi, gj)
        torch.cuda.synchronize()
        assert np.array_equal(g.cpu().numpy(), oracle.fill_err_corr(nxp, nyp, nzp, cxy, gi, gj)), (gi, gj)


@pytest.mark.gpu
def test_synthetic_covariance_matches_fill(ctx):
    """The generator's prior covariance grids equal the fill kernel's."""
    prob = synth.make_problem(synth.config("small"), "cpu")
    c = prob.cfg
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    scales = [float(np.float32(prob.sigma0[a, b])) for a, b in pairs]
    rnxp = 2 * (c.nxp // 2 + 1)
    grids = [torch.empty((c.nzp, c.nyp, rnxp), dtype=torch.float32, device="cuda") for _ in pairs]
    fill_param_corr(ctx, grids, scales, c.nxp, c.nyp, c.nzp, torch.from_numpy(prob.corr_xy).cuda(),
                    torch.from_numpy(prob.circ_corr_t).cuda())
    ec = torch.empty_like(grids[0])
    fill_err_corr(ctx, ec, c.nxp, c.nyp, c.nzp, torch.from_numpy(prob.corr_xy).cuda())
    torch.cuda.synchronize()
    for g, ref in zip(grids, prob.cov):
        assert torch.equal(g[:, :, : c.nxp].cpu(), ref[:, :, : c.nxp])
    assert torch.equal(ec[:, :, : c.nxp].cpu(), prob.errcorr[:, :, : c.nxp])


@pytest.mark.gpu
@pytest.mark.parametrize("transformed", [False, True])
def test_grid_ops_bitexact(ctx, transformed):
    nxp, nyp, nzp = 45, 32, 20
    rnxp = 2 * (nxp // 2 + 1)
    rng = np.random.default_rng(3)
    a0 = rng.standard_normal((nzp, nyp, rnxp)).astype(np.float32)
    b = rng.standard_normal((nzp, nyp, rnxp)).astype(np.float32)
    ops = [L.OP_ADD, L.OP_SUBTRACT, L.OP_MULTIPLY, L.OP_CHANGE_SIGN, L.OP_COPY]
    ops += [L.OP_CONJUGATE] if transformed else [L.OP_MULTIPLY_SCALAR, L.OP_ADD_SCALAR, L.OP_FILL_CONSTANT]
    tb = torch.from_numpy(b).cuda()
    for op in ops:
        ta = torch.from_numpy(a0.copy()).cuda()
        grid_op(ctx, op, ta, tb, nxp, nyp, nzp, transformed, 0.731)
        torch.cuda.synchronize()
        ref = oracle.grid_op(op, a0.copy(), b, transformed, 0.731)
        assert np.array_equal(ta.cpu().numpy(), ref), op
    with pytest.raises(L.CravaHipError):
        grid_op(ctx, L.OP_MULTIPLY_SCALAR, tb, None, nxp, nyp, nzp, True, 2.0)
    with pytest.raises(L.CravaHipError):
        grid_op(ctx, L.OP_CONJUGATE, tb, None, nxp, nyp, nzp, False)
    with pytest.raises(L.CravaHipError):
        grid_op(ctx, L.OP_ADD, tb, None, nxp, nyp, nzp, False)


@pytest.mark.gpu
def test_data_variance_bitexact(ctx):
    prob = synth.make_problem(synth.config("plain_cpu"), "cpu")
    c = prob.cfg
    for g in prob.data:
        v, n = data_variance(ctx, g.cuda(), c.nx, c.ny, c.nz, c.nxp, c.nyp, c.nzp)
        rv, rn = oracle.data_variance(g.numpy(), c.nx, c.ny, c.nz, c.nxp, c.nyp, c.nzp)
        assert n == rn and np.float32(v) == np.float32(rv)


@pytest.mark.gpu
def test_facies_geomodel_bitexact(ctx):
    nx, ny, nz, nfac = 37, 20, 11, 4
    rnx = 2 * (nx // 2 + 1)
    rng = np.random.default_rng(7)
    prob = [rng.uniform(0, 1, (nz, ny, rnx)).astype(np.float32) for _ in range(nfac + 1)]
    prior = [0.1, 0.4, 0.3, 0.2]
    tp = [torch.from_numpy(p.copy()).cuda() for p in prob]
    facies_geomodel(ctx, nx, ny, nz, prior, tp)
    torch.cuda.synchronize()
    oracle.facies_geomodel(nx, ny, nz, prior, prob)
    for a, r in zip(tp, prob):
        assert np.array_equal(a.cpu().numpy(), r)


@pytest.mark.gpu
def test_grid_read_trace(ctx):
    """crava_hip_grid_read_trace returns grid[k0:k0+nk, j, i] bit for bit."""
    import ctypes as C
    nxp, nyp, nzp = 105, 45, 98
    rnxp = 2 * (nxp // 2 + 1)
    g = torch.randn((nzp, nyp, rnxp), device="cuda")
    h = g.cpu().numpy()
    lib = L.lib()
    for i, j, k0, nk in ((0, 0, 0, nzp), (104, 44, 3, 50), (rnxp - 1, 7, 97, 1), (17, 30, 10, 0)):
        out = np.full(max(nk, 1), np.nan, np.float32)
        rc = lib.crava_hip_grid_read_trace(ctx.h, C.c_void_p(g.data_ptr()), nxp, nyp, nzp, i, j, k0, nk,
                                           out.ctypes.data)
        assert rc == L.OK
        assert np.array_equal(out[:nk], h[k0:k0 + nk, j, i])
    out = np.zeros(4, np.float32)
    for bad in ((rnxp, 0, 0, 1), (0, nyp, 0, 1), (0, 0, nzp - 1, 2), (0, 0, -1, 1)):
        assert lib.crava_hip_grid_read_trace(ctx.h, C.c_void_p(g.data_ptr()), nxp, nyp, nzp, *bad,
                                             out.ctypes.data) == L.E_INVALID
