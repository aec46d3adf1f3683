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
    (6, 1, 5),        # one-element y axis (a copy between storages)
    (4, 7, 1),        # one-element z axis
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
    assert rel_l2(back, g[:, :, :nxp]) < 1e-6


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


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(512, 16, 320), (288, 12, 216), (80, 64, 160)], ids=lambda n: "x".join(map(str, n)))
@pytest.mark.parametrize("dbl", [False, True])
def test_every_launch_variant_and_autotuner(ctx, n, dbl):
    """Each specialised launch variant and each runtime-length kernel variant
    (radix order x pencils per workgroup, codes -1..-6) against the oracle."""
    from crava_amd.core import autotune_fft, set_fft_variants
    nxp, nyp, nzp = n
    print(f"variants {n} dbl={dbl}", flush=True)
    dt = np.float64 if dbl else np.float32
    g = seeded_real_grid(nxp, nyp, nzp, seed=11, dtype=dt)
    ref = g.copy()
    (oracle.fft_r2c_f64 if dbl else lambda a, *r: oracle.fft_r2c(a, *r))(ref, nxp, nyp, nzp, False)
    tol = 1e-13 if dbl else 1e-6
    prec = L.F64 if dbl else L.F32
    # static variants: strided 0..11, rows 0..5 (strided 4..7 and rows 3..5
    # load the first pass and store the last straight from/to global memory,
    # strided 8..9 do so in persistent workgroups prefetching the next tile,
    # strided 10..11 on pencil-fastest passes)
    for v in range(-12, 12):
        try:
            set_fft_variants(ctx, nxp, nyp, nzp, prec, [min(v, 5), v, v])
        except L.CravaHipError:
            continue      # variant does not fit (LDS) for this length/precision
        t = torch.from_numpy(g.copy()).cuda()
        ctx.r2c(t, nxp, nyp, nzp, L.DATA)
        torch.cuda.synchronize()
        assert rel_l2(to_complex(t.cpu().numpy()), to_complex(ref)) < tol, v
        ctx.c2r(t, nxp, nyp, nzp, L.DATA)
        torch.cuda.synchronize()
        assert rel_l2(t.cpu().numpy()[:, :, :nxp], g[:, :, :nxp]) < tol, v
    scratch = torch.zeros(g.shape, dtype=torch.float64 if dbl else torch.float32, device="cuda")
    var, ms = autotune_fft(ctx, scratch, nxp, nyp, nzp)
    assert all(-12 <= x <= 11 for x in var) and all(m > 0 for m in ms)
    t = torch.from_numpy(g.copy()).cuda()
    ctx.r2c(t, nxp, nyp, nzp, L.DATA)
    torch.cuda.synchronize()
    assert rel_l2(to_complex(t.cpu().numpy()), to_complex(ref)) < tol


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(80, 80, 160), (15, 12, 24), (512, 8, 320), (6, 1, 5), (4, 7, 1)],
                         ids=lambda n: "x".join(map(str, n)))
def test_r2c_float_storage_double_arithmetic(ctx, n):
    """CRAVA_HIP_MIXED transforms run double arithmetic inside each pass and
    round each pass's outputs once to float: three roundings of at most
    2^-24 relative each, so the spectrum is within 3e-7 (normwise) of the
    double-arithmetic oracle and no worse than the float transform; the
    inverse returns the input to the same accuracy."""
    nxp, nyp, nzp = n
    print(f"fft mixed {n}", flush=True)
    rnxp = 2 * (nxp // 2 + 1)
    # a smooth, strongly decaying covariance-like field: spectrum spans many decades
    x = np.minimum(np.arange(nxp), nxp - np.arange(nxp))[None, None, :]
    y = np.minimum(np.arange(nyp), nyp - np.arange(nyp))[None, :, None]
    z = np.minimum(np.arange(nzp), nzp - np.arange(nzp))[:, None, None]
    g = np.zeros((nzp, nyp, rnxp), np.float32)
    g[:, :, :nxp] = np.exp(-(x / 3.0) ** 2 - (y / 3.0) ** 2 - (z / 4.0) ** 2).astype(np.float32)
    ref = g.copy()
    oracle.fft_r2c(ref, nxp, nyp, nzp, True)
    from crava_amd.core import set_fft_variants
    want = to_complex(ref)
    t32 = torch.from_numpy(g.copy()).cuda()
    ctx.r2c(t32, nxp, nyp, nzp, L.COVARIANCE, L.F32)
    torch.cuda.synchronize()
    err32 = rel_l2(to_complex(t32.cpu().numpy()), want)
    for v in (None, 0, 3, 4, 7, 8, 9, 10, 11):
        if v is not None:
            try:
                set_fft_variants(ctx, nxp, nyp, nzp, L.MIXED, [min(v, 5), v, v])
            except L.CravaHipError:
                continue      # a length without specialised kernels
        t = torch.from_numpy(g.copy()).cuda()
        ctx.r2c(t, nxp, nyp, nzp, L.COVARIANCE, L.MIXED)
        torch.cuda.synchronize()
        err = rel_l2(to_complex(t.cpu().numpy()), want)
        assert err < 3e-7 and err <= max(err32, 1e-7), (v, err, err32)
        # inverse with double arithmetic (COVARIANCE: forward unscaled,
        # inverse 1/N): the round trip returns the input
        ctx.c2r(t, nxp, nyp, nzp, L.COVARIANCE, L.MIXED)
        torch.cuda.synchronize()
        assert rel_l2(t.cpu().numpy()[:, :, :nxp], g[:, :, :nxp]) < 3e-7, v


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(105, 45, 98), (36, 126, 75)], ids=lambda n: "x".join(map(str, n)))
def test_generic_variants_odd_lengths(ctx, n):
    """Runtime-length kernels on odd and 3/5/7-rich lengths: every generic
    variant, forward against the oracle and the round trip."""
    from crava_amd.core import set_fft_variants
    nxp, nyp, nzp = n
    print(f"generic variants {n}", flush=True)
    g = seeded_real_grid(nxp, nyp, nzp, seed=5)
    ref = g.copy()
    oracle.fft_r2c(ref, nxp, nyp, nzp, False)
    # -7..-12 are -1..-6 with the first radix pass loading straight from
    # global memory; float and float-storage/double-arithmetic plans
    for prec in (L.F32, L.MIXED):
        for v in range(-12, 0):
            set_fft_variants(ctx, nxp, nyp, nzp, prec, [v, v, v])
            t = torch.from_numpy(g.copy()).cuda()
            ctx.r2c(t, nxp, nyp, nzp, L.DATA, prec)
            torch.cuda.synchronize()
            assert rel_l2(to_complex(t.cpu().numpy()), to_complex(ref)) < 1e-6, (prec, v)
            ctx.c2r(t, nxp, nyp, nzp, L.DATA, prec)
            torch.cuda.synchronize()
            assert rel_l2(t.cpu().numpy()[:, :, :nxp], g[:, :, :nxp]) < 1e-6, (prec, v)


@pytest.mark.gpu
@pytest.mark.parametrize("n,ng", [((80, 64, 160), 5), ((64, 64, 320), 5), ((64, 64, 64), 1)],
                         ids=lambda p: "x".join(map(str, p)) if isinstance(p, tuple) else str(p))
def test_persistent_strided_variants_on_sets(ctx, n, ng):
    """Strided variants 8 and 9 (persistent workgroups over XCD-contiguous
    tile ranges, the next tile's first pass prefetched) and 10 and 11
    (pencil-fastest passes) on a set of grids:
    partial column blocks (41 complex columns), sets of 5 grids, and a
    launch with more workgroups than tiles.  Spectra against the FP64
    oracle, within rounding of the one-workgroup-per-tile form of the same
    shape (variant 5), and the round trip; float, float storage with double
    arithmetic, double."""
    from crava_amd.core import set_fft_variants
    nxp, nyp, nzp = n
    print(f"persistent strided {n} x{ng}", flush=True)
    for prec, dt, tol in ((L.F32, np.float32, 1e-6), (L.MIXED, np.float32, 3e-7), (L.F64, np.float64, 1e-13)):
        gs = [seeded_real_grid(nxp, nyp, nzp, seed=20 + s, dtype=dt) for s in range(ng)]
        refs = []
        for a in gs:
            r = a.astype(np.float64)
            oracle.fft_r2c_f64(r, nxp, nyp, nzp, False)
            refs.append(to_complex(r))
        base = None
        for v in (5, 8, 9, 10, 11):
            set_fft_variants(ctx, nxp, nyp, nzp, prec, [5, v, v])
            ts = [torch.from_numpy(a.copy()).cuda() for a in gs]
            ctx.r2c_many(ts, nxp, nyp, nzp, L.DATA, prec)
            torch.cuda.synchronize()
            spec = [to_complex(t.cpu().numpy()) for t in ts]
            for x, r in zip(spec, refs):
                assert rel_l2(x, r) < tol, (prec, v)
            if base is None:
                base = spec
            else:
                for x, b in zip(spec, base):
                    assert rel_l2(x, b) < tol, (prec, v)
            ctx.c2r_many(ts, nxp, nyp, nzp, L.DATA, prec)
            torch.cuda.synchronize()
            for t, a in zip(ts, gs):
                assert rel_l2(t.cpu().numpy()[:, :, :nxp], a[:, :, :nxp]) < tol, (prec, v)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(105, 70, 45), (96, 35, 243), (189, 126, 50), (32, 315, 36)],
                         ids=lambda n: "x".join(map(str, n)))
def test_extended_lengths_direct_io_variants(ctx, n):
    """The specialised kernels of the extended 7-smooth lengths (32..320
    beyond the main ones; odd x lengths included): every direct-I/O variant
    they carry (strided 4, 5, 10, 11; rows 3, 5), float, float storage with double
    arithmetic and double, on a set of 3 grids, against the FP64 oracle and
    through the round trip; variants they do not carry are refused."""
    from crava_amd.core import set_fft_variants
    nxp, nyp, nzp = n
    print(f"extended lengths {n}", flush=True)
    with pytest.raises(L.CravaHipError):
        set_fft_variants(ctx, nxp, nyp, nzp, L.F32, [0, 1, 1])
    for prec, dt, tol in ((L.F32, np.float32, 1e-6), (L.MIXED, np.float32, 3e-7), (L.F64, np.float64, 1e-13)):
        gs = [seeded_real_grid(nxp, nyp, nzp, seed=40 + s, dtype=dt) for s in range(3)]
        refs = []
        for a in gs:
            r = a.astype(np.float64)
            oracle.fft_r2c_f64(r, nxp, nyp, nzp, False)
            refs.append(to_complex(r))
        for vr, vs in ((3, 4), (5, 5), (3, 5), (5, 4), (3, 10), (5, 11)):
            set_fft_variants(ctx, nxp, nyp, nzp, prec, [vr, vs, vs])
            ts = [torch.from_numpy(a.copy()).cuda() for a in gs]
            ctx.r2c_many(ts, nxp, nyp, nzp, L.DATA, prec)
            torch.cuda.synchronize()
            for t, r in zip(ts, refs):
                assert rel_l2(to_complex(t.cpu().numpy()), r) < tol, (prec, vr, vs)
            ctx.c2r_many(ts, nxp, nyp, nzp, L.DATA, prec)
            torch.cuda.synchronize()
            for t, a in zip(ts, gs):
                assert rel_l2(t.cpu().numpy()[:, :, :nxp], a[:, :, :nxp]) < tol, (prec, vr, vs)


@pytest.mark.gpu
def test_many_grids_one_launch_equals_one_by_one(ctx):
    """crava_hip_fft_r2c_many / c2r_many: a set of grids per launch gives the
    same bits as transforming the grids one at a time (float, float storage
    with double arithmetic, double)."""
    nxp, nyp, nzp = 80, 45, 64
    print("many grids", flush=True)
    for prec, dt in ((L.F32, np.float32), (L.MIXED, np.float32), (L.F64, np.float64)):
        gs = [seeded_real_grid(nxp, nyp, nzp, seed=s, dtype=dt) for s in range(7)]
        one = [torch.from_numpy(a.copy()).cuda() for a in gs]
        many = [torch.from_numpy(a.copy()).cuda() for a in gs]
        for t in one:
            ctx.r2c(t, nxp, nyp, nzp, L.COVARIANCE, prec)
        ctx.r2c_many(many, nxp, nyp, nzp, L.COVARIANCE, prec)
        torch.cuda.synchronize()
        for a, b in zip(one, many):
            assert torch.equal(a, b), prec
        for t in one:
            ctx.c2r(t, nxp, nyp, nzp, L.DATA, prec)
        ctx.c2r_many(many, nxp, nyp, nzp, L.DATA, prec)
        torch.cuda.synchronize()
        for a, b in zip(one, many):
            assert torch.equal(a, b), prec
    with pytest.raises(L.CravaHipError):
        ctx.r2c_many([torch.zeros(2 * 41 * 45 * 64, device="cuda")] * 17, nxp, nyp, nzp, L.DATA)


@pytest.mark.gpu
@pytest.mark.parametrize("n", [(80, 80, 160), (15, 12, 24), (512, 8, 320), (105, 45, 98)],
                         ids=lambda n: "x".join(map(str, n)))
def test_mean_deflated_float_transform(ctx, n):
    """CRAVA_HIP_F32_DEFLATE on a field whose level dominates its variation
    (8 + a trend of 0.3 + small noise, like ln Vp): the spectrum away from
    the DC bin keeps float accuracy relative to the variation (the plain
    float transform's error there scales with the level), the DC bin is the
    FP64 one rounded once, and the inverse returns the input to float
    rounding of the level."""
    nxp, nyp, nzp = n
    print(f"deflated transform {n}", flush=True)
    rnxp = 2 * (nxp // 2 + 1)
    rng = np.random.default_rng(3)
    z = np.arange(nzp)[:, None, None] / max(1, nzp - 1)
    g = np.zeros((nzp, nyp, rnxp), np.float32)
    g[:, :, :nxp] = (8.0 + 0.3 * z + 0.01 * rng.standard_normal((nzp, nyp, nxp))).astype(np.float32)
    ref = g.astype(np.float64)
    oracle.fft_r2c_f64(ref, nxp, nyp, nzp, False)
    want = ref.view(np.complex128).copy()
    errs = {}
    for prec in (L.F32, L.F32_DEFLATE):
        t = torch.from_numpy(g.copy()).cuda()
        ctx.r2c(t, nxp, nyp, nzp, L.DATA, prec)
        torch.cuda.synchronize()
        got = to_complex(t.cpu().numpy()).astype(np.complex128)
        d = got - want
        dc_err = abs(d[0, 0, 0]) / abs(want[0, 0, 0])
        d[0, 0, 0] = 0
        w0 = want.copy()
        w0[0, 0, 0] = 0
        errs[prec] = float(np.linalg.norm(d) / np.linalg.norm(w0))
        if prec == L.F32_DEFLATE:
            assert dc_err < 1.2e-7, dc_err
            ctx.c2r(t, nxp, nyp, nzp, L.DATA, prec)
            torch.cuda.synchronize()
            back = t.cpu().numpy()[:, :, :nxp].astype(np.float64)
            # float storage of the level bounds the round trip: within 2 ulp
            # of 8 (2^-20) root-mean-square, 4 ulp at most
            dev = np.abs(back - g[:, :, :nxp])
            assert np.sqrt(np.mean(dev ** 2)) < 2 * 2.0 ** -20 and dev.max() < 4 * 2.0 ** -20
    print(f"  non-DC spectrum error: float {errs[L.F32]:.2e}, deflated {errs[L.F32_DEFLATE]:.2e}", flush=True)
    assert errs[L.F32_DEFLATE] < 1e-6
    assert errs[L.F32_DEFLATE] < 0.25 * errs[L.F32]
