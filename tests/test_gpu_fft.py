"""GPU 3-D real transforms against the oracle (through the C ABI)."""
import numpy as np
import pytest
import torch

import oracle
from crava_amd import _lib as L
from tests.helpers import rel_l2, seeded_real_grid, to_complex

SIZES = [
    (15, 12, 24),     # odd nxp, radices 5,3 / 4,3 / 8,3
    (80, 80, 160),    # the plain CPU-runnable grid
    (21, 49, 7),      # radix 7 everywhere
    (250, 9, 10),     # radix 2 + 5^3, radix 3, radix 2*5
    (96, 32, 45),
    (512, 8, 320),    # the large-grid x and z lengths
    (2, 1, 1),        # degenerate
]


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES, ids=lambda n: "x".join(map(str, n)))
@pytest.mark.parametrize("cubetype", [L.DATA, L.COVARIANCE])
def test_r2c_c2r_f32(ctx, n, cubetype):
    nxp, nyp, nzp = n
    print(f"fft f32 {n} cubetype={cubetype}", flush=True)
    g = seeded_real_grid(nxp, nyp, nzp, seed=sum(n) + cubetype)
    ref = g.copy()
    oracle.fft_r2c(ref, nxp, nyp, nzp, cubetype == L.COVARIANCE)
    t = torch.from_numpy(g.copy()).cuda()
    ctx.r2c(t, nxp, nyp, nzp, cubetype)
    torch.cuda.synchronize()
    got = t.cpu().numpy()
    assert rel_l2(to_complex(got), to_complex(ref)) < 1e-6
    # back: c2r of the oracle spectrum
    ref2 = ref.copy()
    oracle.fft_c2r(ref2, nxp, nyp, nzp, cubetype == L.COVARIANCE)
    ctx.c2r(t, nxp, nyp, nzp, cubetype)
    torch.cuda.synchronize()
    back = t.cpu().numpy()[:, :, :nxp]
    assert rel_l2(back, ref2[:, :, :nxp]) < 1e-6
    scale = 1.0 if cubetype != L.COVARIANCE else 1.0 / (nxp * nyp * nzp)
    assert rel_l2(back, g[:, :, :nxp] * scale) < 1e-6


@pytest.mark.gpu
@pytest.mark.parametrize("n", SIZES[:5], ids=lambda n: "x".join(map(str, n)))
def test_r2c_c2r_f64(ctx, n):
    nxp, nyp, nzp = n
    print(f"fft f64 {n}", flush=True)
    g = seeded_real_grid(nxp, nyp, nzp, seed=7 + sum(n), dtype=np.float64)
    ref = g.copy()
    oracle.fft_r2c_f64(ref, nxp, nyp, nzp, False)
    t = torch.from_numpy(g.copy()).cuda()
    ctx.r2c(t, nxp, nyp, nzp, L.DATA)
    torch.cuda.synchronize()
    assert rel_l2(to_complex(t.cpu().numpy()), to_complex(ref)) < 1e-13
    ctx.c2r(t, nxp, nyp, nzp, L.DATA)
    torch.cuda.synchronize()
    assert rel_l2(t.cpu().numpy()[:, :, :nxp], g[:, :, :nxp]) < 1e-13


@pytest.mark.gpu
def test_c2r_ignores_imag_of_self_conjugate_bins(ctx):
    """A halfcomplex c2r reads only Re of the kx=0 and kx=n/2 bins."""
    nxp, nyp, nzp = 16, 4, 6
    rng = np.random.default_rng(3)
    cn = nxp // 2 + 1
    spec = (rng.standard_normal((nzp, nyp, cn)) + 1j * rng.standard_normal((nzp, nyp, cn))).astype(np.complex64)
    g = np.ascontiguousarray(spec).view(np.float32).copy()
    ref = g.copy()
    oracle.fft_c2r(ref, nxp, nyp, nzp, False)
    t = torch.from_numpy(g).cuda()
    ctx.c2r(t, nxp, nyp, nzp, L.DATA)
    torch.cuda.synchronize()
    assert rel_l2(t.cpu().numpy()[:, :, :nxp], ref[:, :, :nxp]) < 1e-6


@pytest.mark.gpu
def test_fft_rejects_bad_sizes(ctx):
    t = torch.zeros(11 * 11 * 14, device="cuda")
    with pytest.raises(L.CravaHipError):
        ctx.r2c(t, 11, 11, 11, L.DATA)   # 11 is not 7-smooth
