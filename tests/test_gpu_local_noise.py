"""GPU local-noise correction against the oracle
(crava_hip_local_noise_correct, crava_hip_local_noise_realisation).

G comes from the oracle's computeG (LAPACK dgeev's eigenvectors), as a
reference build would pass its own.

Tolerances: the per-trace matrices are matrix functions computed by Jacobi
rotations on the device and by dgeev in the oracle, so the realisation
factors agree to 1e-10 relative; the corrected float grids then agree to
within one float ulp at every point.  The realisation transform is bit-exact
for a given factor, checked with the device's own factors.  At the full
grid size, uniform noise scales make every trace's matrix the same, whose
corrected grid is computed in numpy from one oracle matrix (one ulp).
"""
import numpy as np
import pytest
import torch

from crava_amd import _lib as L
from crava_amd import synth
from crava_amd.core import local_noise_correct, local_noise_realisation
from oracle import local_noise as ln

CASES = [
    # nx, ny, nz, nxp, nyp, nzp, ntheta
    (13, 9, 21, 16, 12, 24, 3),
    (7, 11, 40, 9, 12, 45, 1),
    (20, 6, 33, 25, 8, 40, 6),
    (9, 8, 70, 12, 9, 80, 8),
    (11, 7, 30, 12, 8, 32, 4, "duplicate"),  # singular G Sm G^T + D E D: the eigen route drops an eigenvalue
]


def _grids(nx, ny, nz, nxp, nyp, nzp, seed):
    rnxp = 2 * (nxp // 2 + 1)
    rng = np.random.default_rng(seed)
    base = np.array([8.0, 7.2, 7.8], dtype=np.float32)
    mean = [(b + 0.05 * rng.standard_normal((nzp, nyp, rnxp))).astype(np.float32) for b in base]
    post = [(m + 0.03 * rng.standard_normal(m.shape)).astype(np.float32) for m in mean]
    return post, mean


def _ulps(got, want):
    g = got.astype(np.float64)
    w = want.astype(np.float64)
    return float(np.max(np.abs(g - w) / np.spacing(np.abs(want)).astype(np.float64)))


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES)
def test_local_noise_correction_vs_oracle(ctx, case):
    nx, ny, nz, nxp, nyp, nzp, nt = case[:7]
    print("local noise", case, flush=True)
    p = synth.local_noise_problem(nt, nx, ny, seed=nt + nx, duplicate_angle=len(case) > 7)
    G = ln.compute_g(p["E"], p["Sm"], p["Smd"])
    post, mean = _grids(nx, ny, nz, nxp, nyp, nzp, nx * ny)
    sn, tc = ln.trace_matrices(G, p["scales"], p["E"], p["Sm"], p["Smd"])
    want = [g.copy() for g in post]
    ln.apply_correction(want, mean, sn, nx, ny, nz)

    tp = [torch.from_numpy(g.copy()).cuda() for g in post]
    tm = [torch.from_numpy(g.copy()).cuda() for g in mean]
    ttc = torch.zeros(ny * nx * 9, dtype=torch.float64, device="cuda")
    local_noise_correct(ctx, nx, ny, nz, nxp, nyp, nzp, G, p["E"], p["Sm"], p["Smd"], p["scales"], tp, tm, ttc)
    torch.cuda.synchronize()
    got_tc = ttc.cpu().numpy().reshape(ny, nx, 3, 3)
    assert np.allclose(got_tc, tc, rtol=1e-10, atol=1e-10 * np.abs(tc).max())
    for c in range(3):
        g = tp[c].cpu().numpy()
        assert _ulps(g[:nz, :ny, :nx], want[c][:nz, :ny, :nx]) <= 1.0, c
        # outside the volume nothing changes
        assert np.array_equal(g[nz:], post[c][nz:])
        assert np.array_equal(g[:, ny:], post[c][:, ny:])
        assert np.array_equal(g[:, :, nx:], post[c][:, :, nx:])

    # the realisation transform, bit-exact for the device's factors
    seed = [(0.01 * np.random.default_rng(c).standard_normal(post[0].shape)).astype(np.float32) for c in range(3)]
    ref = [s.copy() for s in seed]
    ln.transform_realisation(ref, got_tc, nx, ny, nz)
    ts = [torch.from_numpy(s.copy()).cuda() for s in seed]
    local_noise_realisation(ctx, nx, ny, nz, nxp, nyp, nzp, ttc, ts)
    torch.cuda.synchronize()
    for c in range(3):
        assert np.array_equal(ts[c].cpu().numpy(), ref[c]), c


@pytest.mark.gpu
def test_local_noise_full_grid_uniform_scales(ctx):
    nx, ny, nz, nxp, nyp, nzp, nt = 480, 480, 300, 512, 512, 320, 6
    print("local noise full grid", flush=True)
    p = synth.local_noise_problem(nt, nx, ny, seed=11, uniform=2.5)
    # angles 1.. are uniform (relative scale 1); angle 0's table holds its
    # minimum 1.0 at trace (0, 0) and 3.0 elsewhere, so every other trace has
    # the relative scales (3, 1, ..., 1) and the correction is not the identity
    p["scales"][0] = np.where((np.arange(nx)[None, :] + np.arange(ny)[:, None]) == 0, 1.0, 3.0)
    G = ln.compute_g(p["E"], p["Sm"], p["Smd"])
    sc_rest = [3.0] + [1.0] * (nt - 1)
    S, T = ln.new_posterior_cov_pointwise(G, sc_rest, p["E"], p["Sm"], p["Smd"])
    post, mean = _grids(nx, ny, nz, nxp, nyp, nzp, 5)
    tp = [torch.from_numpy(g).cuda() for g in post]
    tm = [torch.from_numpy(g).cuda() for g in mean]
    ttc = torch.zeros(ny * nx * 9, dtype=torch.float64, device="cuda")
    local_noise_correct(ctx, nx, ny, nz, nxp, nyp, nzp, G, p["E"], p["Sm"], p["Smd"], p["scales"], tp, tm, ttc)
    torch.cuda.synchronize()
    tc = ttc.cpu().numpy().reshape(ny, nx, 3, 3)
    want_tc = T @ ln.inv_sqrt_post(p["Smd"])
    assert np.allclose(tc[1:, 1:], want_tc, rtol=1e-10, atol=1e-10 * np.abs(want_tc).max())
    d = [(post[c][:nz, :ny, :nx] - mean[c][:nz, :ny, :nx]).astype(np.float64) for c in range(3)]
    for r in range(3):
        v = mean[r][:nz, :ny, :nx].astype(np.float64) + S[r, 0] * d[0]
        v = v + S[r, 1] * d[1]
        v = v + S[r, 2] * d[2]
        want = v.astype(np.float32)
        got = tp[r][:nz, :ny, :nx].cpu().numpy()
        assert _ulps(got[:, 1:, 1:], want[:, 1:, 1:]) <= 1.0, r


@pytest.mark.gpu
def test_local_noise_rejects(ctx):
    nx, ny, nz, nxp, nyp, nzp = 4, 3, 5, 4, 4, 8
    p = synth.local_noise_problem(3, nx, ny)
    G = ln.compute_g(p["E"], p["Sm"], p["Smd"])
    post, mean = _grids(nx, ny, nz, nxp, nyp, nzp, 1)
    tp = [torch.from_numpy(g).cuda() for g in post]
    tm = [torch.from_numpy(g).cuda() for g in mean]
    with pytest.raises(L.CravaHipError):
        local_noise_correct(ctx, nx + 1, ny, nz, nxp, nyp, nzp, G, p["E"], p["Sm"], p["Smd"], p["scales"], tp, tm)
    with pytest.raises(L.CravaHipError):
        local_noise_correct(ctx, nx, ny, nz, nxp, nyp, nzp, np.zeros((9, 3)), np.eye(9), p["Sm"], p["Smd"],
                            p["scales"] * 3, tp, tm)
    ts = [torch.zeros_like(g) for g in tp]
    with pytest.raises(L.CravaHipError):
        local_noise_realisation(ctx, nx, ny, nz + 9, nxp, nyp, nzp, torch.zeros(9 * nx * ny, dtype=torch.float64,
                                                                                 device="cuda"), ts)
    assert L.lib().crava_hip_local_noise_realisation(ctx.h, nx, ny, nz, nxp, nyp, nzp, None, None) != 0
