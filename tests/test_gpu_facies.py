"""GPU facies probabilities: bit-exact against the oracle."""
import ctypes as C

import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd._lib import check, lib, ptr_array


def facies_case(seed, nx=20, ny=12, nz=14, nxp=24, nyp=16, nzp=18, nfac=3, tn=(100, 100, 50)):
    """Seeded elastic grids and smoothed-histogram-like tables (Gaussian blobs
    with small negative ripples, so the max(0, .) clamp is exercised)."""
    rng = np.random.default_rng(seed)
    rnxp = 2 * (nxp // 2 + 1)
    centres = np.array([[8.05, 7.35, 0.85], [8.12, 7.45, 0.88], [7.98, 7.30, 0.80]])[:nfac]
    grids = []
    pick = rng.integers(0, nfac, size=(nzp, nyp, rnxp))
    for c in range(3):
        v = centres[pick, c] + rng.normal(0, [0.06, 0.08, 0.03][c], size=(nzp, nyp, rnxp))
        grids.append(v.astype(np.float32))
    x0, y0, z0 = 7.7, 7.0, 0.7
    lx, ly, lz = 0.7, 0.8, 0.3
    n1, n2, n3 = tn
    vol = np.array([x0, y0, lx / n1, ly / n2, z0, z0 + lz], np.float64)
    rn = 2 * (n1 // 2 + 1)
    xs = x0 + (np.arange(n1) + 0.5) * lx / n1
    ys = y0 + (np.arange(n2) + 0.5) * ly / n2
    zs = z0 + (np.arange(n3) + 0.5) * lz / n3
    tabs = np.zeros((nfac, n3, n2, rn), np.float32)
    for f in range(nfac):
        cx, cy, cz = centres[f]
        g = np.exp(-((xs[None, None, :] - cx) ** 2 / 0.01 + (ys[None, :, None] - cy) ** 2 / 0.015 +
                     (zs[:, None, None] - cz) ** 2 / 0.002))
        g = g / g.sum() + rng.normal(0, 2e-7, size=g.shape)
        tabs[f, :, :, :n1] = g.astype(np.float32)
    prior = np.array([0.5, 0.3, 0.2][:nfac], np.float32)
    vn = np.array([n1, n2, n3], np.int32)
    return dict(nx=nx, ny=ny, nz=nz, nxp=nxp, nyp=nyp, nzp=nzp, grids=grids, tabs=tabs, tn=tn, vol=vol, vn=vn,
                prior=prior, p_undef=0.01)


@pytest.mark.gpu
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_facies_bitexact_vs_oracle(ctx, seed):
    print(f"facies seed={seed}", flush=True)
    cse = facies_case(seed)
    nx, ny, nz = cse["nx"], cse["ny"], cse["nz"]
    nfac = len(cse["prior"])
    ref = oracle.facies(nx, ny, nz, cse["nxp"], cse["nyp"], cse["nzp"], *cse["grids"], cse["tabs"], cse["tn"],
                        cse["vol"], cse["vn"], cse["prior"], cse["p_undef"])
    tg = [torch.from_numpy(g).cuda() for g in cse["grids"]]
    tt = torch.from_numpy(cse["tabs"]).cuda()
    sm = 2 * (nx // 2 + 1)
    out = [torch.zeros((nz, ny, sm), device="cuda") for _ in range(nfac + 1)]
    am = torch.zeros((nz, ny, nx), dtype=torch.int32, device="cuda")
    cnt = torch.zeros(nfac, dtype=torch.int64, device="cuda")
    vol = np.ascontiguousarray(cse["vol"])
    vn = np.ascontiguousarray(cse["vn"])
    pr = np.ascontiguousarray(cse["prior"])
    check(lib().crava_hip_facies(ctx.h, nx, ny, nz, cse["nxp"], cse["nyp"], cse["nzp"],
                                 *[int(t.data_ptr()) for t in tg], nfac, int(tt.data_ptr()), *cse["tn"],
                                 vol.ctypes.data, vn.ctypes.data, pr.ctypes.data, cse["p_undef"],
                                 ptr_array([int(o.data_ptr()) for o in out]), int(am.data_ptr()),
                                 int(cnt.data_ptr())), "facies")
    torch.cuda.synchronize()
    for a, b in zip(out, ref):
        assert np.array_equal(a.cpu().numpy().view(np.uint32), b.view(np.uint32))
    probs = np.stack([r[:, :, :nx] for r in ref[:nfac]])
    want = np.argmax(probs, axis=0)
    assert np.array_equal(am.cpu().numpy(), want)
    assert np.array_equal(cnt.cpu().numpy(), np.bincount(want.ravel(), minlength=nfac))
    # probabilities of every voxel sum to one with the undefined class
    tot = probs.sum(0) + ref[nfac][:, :, :nx]
    assert np.allclose(tot, 1.0, atol=1e-5)
