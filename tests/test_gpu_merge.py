"""GPU time-lapse merge (crava_hip_merge_mean_4d / crava_hip_merge_cov_4d)
against the oracle's restatement of State4D::merge / mergeCov: bit-exact
(float additions in the reference's order), also with the outputs aliasing
inputs."""
import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd.core import merge_cov_4d, merge_mean_4d

SIZES = [(9, 6, 8), (80, 64, 50), (256, 128, 96)]


def _grids(rng, n, shape):
    nxp, nyp, nzp = n
    cn = nxp // 2 + 1
    return [(rng.standard_normal((nzp, nyp, cn)) + 1j * rng.standard_normal((nzp, nyp, cn))).astype(np.complex64)
            for _ in range(shape)]


def _dev(a):
    return torch.from_numpy(a.view(np.float32).copy()).cuda()


def _host(t):
    return t.cpu().numpy().view(np.complex64)


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES, ids=lambda n: "x".join(map(str, n)))
def test_merge_cov_bitexact(ctx, n):
    print(f"merge cov {n}", flush=True)
    rng = np.random.default_rng(sum(n))
    ss, dd, sd = _grids(rng, n, 6), _grids(rng, n, 6), _grids(rng, n, 9)
    want = oracle.merge_cov_4d(ss, dd, sd)
    dss, ddd, dsd = [_dev(x) for x in ss], [_dev(x) for x in dd], [_dev(x) for x in sd]
    out = [torch.empty_like(dss[0]) for _ in range(6)]
    merge_cov_4d(ctx, *n, dss, ddd, dsd, out)
    torch.cuda.synchronize()
    for q in range(6):
        assert np.array_equal(_host(out[q]), want[q].reshape(_host(out[q]).shape)), q
    # in place over the static-static grids
    merge_cov_4d(ctx, *n, dss, ddd, dsd, dss)
    torch.cuda.synchronize()
    for q in range(6):
        assert np.array_equal(_host(dss[q]), want[q].reshape(_host(dss[q]).shape)), q


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES[:2], ids=lambda n: "x".join(map(str, n)))
def test_merge_mean_bitexact(ctx, n):
    rng = np.random.default_rng(7)
    ms, md = _grids(rng, n, 3), _grids(rng, n, 3)
    want = oracle.merge_mean_4d(ms, md)
    dms, dmd = [_dev(x) for x in ms], [_dev(x) for x in md]
    merge_mean_4d(ctx, *n, dms, dmd, dms)      # in place over the static means
    torch.cuda.synchronize()
    for l in range(3):
        assert np.array_equal(_host(dms[l]), want[l].reshape(_host(dms[l]).shape)), l


@pytest.mark.gpu
def test_merge_rejects_null(ctx):
    g = torch.zeros(2 * 5 * 4 * 4, device="cuda")
    p = g.data_ptr()
    sd = L.ptr_array([p] * 8 + [None])          # a NULL static-dynamic grid
    rc = L.lib().crava_hip_merge_cov_4d(ctx.h, 8, 4, 4, L.ptr_array([p] * 6), L.ptr_array([p] * 6), sd,
                                        L.ptr_array([p] * 6))
    assert rc == L.E_INVALID
    rc = L.lib().crava_hip_merge_mean_4d(ctx.h, 8, 4, 4, None, L.ptr_array([p] * 3), L.ptr_array([p] * 3))
    assert rc == L.E_INVALID
