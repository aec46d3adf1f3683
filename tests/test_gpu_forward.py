"""GPU forward modelling (crava_hip_synthetic_seismic) against the oracle.

Tolerance: inside the volume the two sides use different float transforms
(double DFTs rounded to float in the oracle, two traces per complex Stockham
column on the GPU): every sample agrees to 4e-6 of the angle's largest value
and the relative L2 error is below 1e-6; the padding traces hold the
impedance and agree bit for bit.
"""
import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from crava_amd.core import synthetic_seismic

CASES = [
    # nx, ny, nz, nxp, nyp, nzp, ntheta, tables
    (7, 5, 20, 9, 6, 32, 1, False),      # odd nx: a complex column pairs a volume and a padding trace
    (12, 9, 60, 16, 12, 80, 3, True),    # specialised z length
    (15, 11, 41, 18, 14, 50, 8, True),   # runtime-length z (50 = 2 5 5), 8 angles
    (30, 20, 130, 36, 24, 160, 6, True),
    (20, 9, 250, 24, 10, 320, 6, True),   # the 512x512x320 workload's z length
    (10, 6, 200, 12, 8, 288, 2, False),   # radix-3 passes
    (9, 7, 400, 10, 8, 512, 3, True),
]


def _inputs(n, nt, tables, seed):
    nx, ny, nz, nxp, nyp, nzp = n
    rng = np.random.default_rng(seed)
    rnxp = 2 * (nxp // 2 + 1)
    g = [rng.standard_normal((nzp, nyp, rnxp)).astype(np.float32) for _ in range(3)]
    A = rng.standard_normal((nt, 3)).astype(np.float32)
    t = np.arange(nzp)
    t = np.minimum(t, nzp - t).astype(np.float64)
    w = np.stack([np.exp(-(t / 3.0) ** 2) * np.cos((0.4 + 0.1 * l) * t) for l in range(nt)]).astype(np.float32)
    kw = {}
    if tables:
        kw["shift"] = [rng.uniform(-12, 12, (ny, nx)).astype(np.float32) if l % 2 == 0 else None for l in range(nt)]
        kw["gain"] = [rng.uniform(0.5, 1.5, (ny, nx)).astype(np.float32) if l % 3 != 1 else None for l in range(nt)]
        kw["stretch"] = [rng.uniform(0.7, 1.3, (ny, nx)).astype(np.float32) if l % 2 == 1 else None
                         for l in range(nt)]
    return g, A, w, kw


def _check(got, want, n):
    nx, ny = n[0], n[1]
    inside = np.zeros(got.shape[1:], bool)
    inside[:ny, :nx] = True
    a, b = got[:, inside].astype(np.float64), want[:, inside].astype(np.float64)
    assert np.abs(a - b).max() <= 4e-6 * np.abs(b).max()
    assert np.linalg.norm(a - b) <= 1e-6 * np.linalg.norm(b)
    assert np.array_equal(got[:, ~inside], want[:, ~inside])


@pytest.mark.gpu
@pytest.mark.parametrize("case", CASES, ids=lambda c: "x".join(map(str, c[:6])) + f"_nt{c[6]}")
def test_synthetic_seismic_vs_oracle(ctx, case):
    *n, nt, tables = case
    print(f"forward {n} nt={nt} tables={tables}", flush=True)
    g, A, w, kw = _inputs(n, nt, tables, seed=sum(n))
    want = oracle.synthetic_seismic(*n, A, w, 4.0, *g, **kw)
    dg = [torch.from_numpy(x).cuda() for x in g]
    out = [torch.full_like(dg[0], float("nan")) for _ in range(nt)]
    synthetic_seismic(ctx, *n, A, w, 4.0, *dg, out, **kw)
    torch.cuda.synchronize()
    for l in range(nt):
        _check(out[l].cpu().numpy(), want[l], n)
    # a second call on the same outputs gives the same bits
    first = [o.clone() for o in out]
    synthetic_seismic(ctx, *n, A, w, 4.0, *dg, out, **kw)
    torch.cuda.synchronize()
    assert all(torch.equal(a, b) for a, b in zip(first, out))


@pytest.mark.gpu
def test_synthetic_seismic_log_levels(ctx):
    """Parameters as log(Vp), log(Vs), log(rho) levels with small variations:
    the impedance's float rounding is of the order of the reflection
    coefficients' own size, so agreement at the usual bound needs the
    impedance, taper and difference rounded as the reference rounds them."""
    n = (20, 9, 250, 24, 10, 320)
    nt = 6
    print("forward log levels", flush=True)
    rng = np.random.default_rng(5)
    rnxp = 2 * (n[3] // 2 + 1)
    g = [(lvl + 0.01 * rng.standard_normal((n[5], n[4], rnxp))).astype(np.float32) for lvl in (8.0, 7.3, 0.8)]
    th = np.deg2rad(np.linspace(4.0, 40.0, nt))
    k2 = 0.25 * np.sin(th) ** 2 * 4
    A = np.stack([0.5 * (1 + np.tan(th) ** 2), -k2 * 2, 0.5 * (1 - k2)], axis=1).astype(np.float32)
    t = np.arange(n[5])
    t = np.minimum(t, n[5] - t).astype(np.float64)
    w = np.stack([np.exp(-(t / 3.0) ** 2) * np.cos(0.4 * t) for _ in range(nt)]).astype(np.float32)
    want = oracle.synthetic_seismic(*n, A, w, 4.0, *g)
    dg = [torch.from_numpy(x).cuda() for x in g]
    out = [torch.empty_like(dg[0]) for _ in range(nt)]
    synthetic_seismic(ctx, *n, A, w, 4.0, *dg, out)
    torch.cuda.synchronize()
    for l in range(nt):
        _check(out[l].cpu().numpy(), want[l], n)


@pytest.mark.gpu
def test_synthetic_seismic_full_grid_sampled_traces(ctx):
    """512x512x320 grid (448x448x256 volume), 6 angles: sampled traces of the
    GPU result against the oracle run on a small volume made of those traces
    (each trace depends on its own samples and the wavelet only)."""
    nx, ny, nz, nxp, nyp, nzp, nt = 448, 448, 256, 512, 512, 320, 6
    print("forward full grid", flush=True)
    rnxp = 2 * (nxp // 2 + 1)
    gen = torch.Generator(device="cuda").manual_seed(11)
    g = [torch.randn((nzp, nyp, rnxp), device="cuda", generator=gen) for _ in range(3)]
    rng = np.random.default_rng(12)
    A = rng.standard_normal((nt, 3)).astype(np.float32)
    t = np.arange(nzp)
    t = np.minimum(t, nzp - t).astype(np.float64)
    w = np.stack([np.exp(-(t / 4.0) ** 2) * np.cos(0.3 * t) for _ in range(nt)]).astype(np.float32)
    out = [torch.empty_like(g[0]) for _ in range(nt)]
    synthetic_seismic(ctx, nx, ny, nz, nxp, nyp, nzp, A, w, 4.0, *g, out)
    torch.cuda.synchronize()
    ii = np.array([0, 1, 2, 77, 200, 301, 446, 447])
    jj = np.array([0, 5, 100, 255, 390, 447])
    sub_in = [x[:, jj][:, :, ii].cpu().numpy() for x in g]
    sx, sy = len(ii), len(jj)
    srnxp = 2 * (sx // 2 + 1)
    pad = [np.zeros((nzp, sy, srnxp), np.float32) for _ in range(3)]
    for p, s in zip(pad, sub_in):
        p[:, :, :sx] = s
    want = oracle.synthetic_seismic(sx, sy, nz, sx, sy, nzp, A, w, 4.0, *pad)
    for l in range(nt):
        got = out[l][:, jj][:, :, ii].cpu().numpy()
        ref = want[l][:, :, :sx]
        assert np.abs(got - ref).max() <= 4e-6 * np.abs(ref).max(), l


@pytest.mark.gpu
def test_synthetic_seismic_rejects_bad_arguments(ctx):
    n = (7, 5, 20, 9, 6, 32)
    g, A, w, _ = _inputs(n, 2, False, 0)
    dg = [torch.from_numpy(x).cuda() for x in g]
    out = [torch.empty_like(dg[0]) for _ in range(9)]
    with pytest.raises(L.CravaHipError):
        synthetic_seismic(ctx, *n, np.zeros((9, 3), np.float32), np.zeros((9, 32), np.float32), 4.0, *dg, out)
    with pytest.raises(L.CravaHipError):   # volume larger than the padded grid
        synthetic_seismic(ctx, 10, 5, 20, 9, 6, 32, A, w, 4.0, *dg, out[:2])
    with pytest.raises(L.CravaHipError):   # shifts need a positive dz
        synthetic_seismic(ctx, *n, A, w, 0.0, *dg, out[:2], shift=[np.zeros((5, 7), np.float32), None])
    bad = np.ones((5, 7), np.float32)
    bad[2, 3] = 0.0
    with pytest.raises(L.CravaHipError):   # stretch factors must be positive
        synthetic_seismic(ctx, *n, A, w, 4.0, *dg, out[:2], stretch=[None, bad])
