"""The C++ host mirror (crava_amd/host: FFTGrid, SeismicParametersHolder,
Crava over the C ABI) driven like the reference's inversion sequence by
tests/cpp/host_mirror_main, against the oracle:

  prior covariance fill (initializeCorrelations), cursor I/O of every grid,
  computeDataVariance (bit-exact), setupErrorCorrelation, the forward
  transforms, the per-point posterior and the inverse transforms
  (computePostMeanResidAndFFTCov), one posterior realisation.

Tolerances: circCorrT from the mirror's double DFT vs numpy's FFT, 1e-6;
the inversion outputs as the smoke check (float FFTs of two different
algorithms plus the float loop), 5e-5 relative L2; the synthetic seismic of
the mirror's own posterior means against the oracle's forward model, 1e-6
relative L2 inside the volume and bit-exact impedance in the padding.
"""
import math
import os
import struct
import subprocess

import numpy as np
import pytest

import oracle
from oracle import local_noise as lnoise
from crava_amd import _lib as L
from crava_amd import synth
from tests.helpers import rel_l2

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(ROOT, "tests", "cpp", "host_mirror_main")


def _write(path, a):
    np.ascontiguousarray(a, dtype=np.float32).tofile(path)


def _read(path, shape):
    return np.fromfile(path, dtype=np.float32).reshape(shape)


@pytest.mark.gpu
@pytest.mark.parametrize("name", ["small", "plain_cpu"])
def test_host_mirror_inversion_vs_oracle(tmp_path, name):
    print(f"host mirror {name}", flush=True)
    assert os.path.exists(BIN), "tests/cpp/host_mirror_main is built by crava_amd.build"
    prob = synth.make_problem(synth.config(name), "cpu")
    c = prob.cfg
    nt = prob.ntheta
    nxp, nyp, nzp = c.nxp, c.nyp, c.nzp
    rnxp = 2 * (nxp // 2 + 1)
    shape = (nzp, nyp, rnxp)
    low_int_cut = int(math.floor(c.low_cut * (nzp * 0.001 * c.dt)))
    lags = np.arange(nzp // 2 + 1)
    corr_t = synth.spherical((lags * c.dt / c.vertical_range_ms).astype(np.float32))
    seed = 12345
    d = str(tmp_path)
    with open(os.path.join(d, "params.txt"), "w") as f:
        f.write(f"{c.nx} {c.ny} {c.nz} {nxp} {nyp} {nzp} {nt} {c.dt} {c.low_cut} {c.high_cut} {c.wnc} "
                f"{L.F32} 0.1 0 0 {low_int_cut} {len(corr_t)} {c.ang_range_deg} {seed}\n")
        f.write(" ".join(repr(float(x)) for x in c.thetas_deg) + "\n")
        f.write(" ".join(repr(float(x)) for x in c.snr) + "\n")
        f.write(" ".join(repr(float(x)) for x in prob.A.reshape(-1)) + "\n")
        f.write(" ".join(repr(float(x)) for x in prob.wnorm) + "\n")
        f.write(" ".join(repr(float(x)) for x in prob.sigma0.reshape(-1)) + "\n")
    _write(os.path.join(d, "wavelets.bin"), prob.wavelets)
    _write(os.path.join(d, "corrxy.bin"), prob.corr_xy)
    _write(os.path.join(d, "corrt.bin"), corr_t)
    for i, g in enumerate(prob.mu):
        _write(os.path.join(d, f"mu{i}.bin"), g.numpy())
    for i, g in enumerate(prob.data):
        _write(os.path.join(d, f"data{i}.bin"), g.numpy())

    # a CRAVA binary file as the reference writes it: text line, simbox
    # geometry (4 doubles, 2 ints, 7 doubles), rnxp, nyp, nzp, the padded
    # reals; NRLib writes big-endian
    crava_in = np.random.default_rng(5).standard_normal(shape).astype(np.float32)
    with open(os.path.join(d, "in.crava"), "wb") as f:
        f.write(b"crava_fftgrid_binary\n")
        f.write(struct.pack(">4d2i7d3i", 1.0, 2.0, 25.0, 25.0, c.nx, c.ny, 0, 0, 1, 0, 0, 1, 0.0, rnxp, nyp, nzp))
        f.write(crava_in.astype(">f4").tobytes())

    # State4D parts (small case only): 3 + 3 means, 6 + 6 + 9 covariances
    cn = nxp // 2 + 1
    if name == "small":
        rng4 = np.random.default_rng(44)
        parts = [(rng4.standard_normal((nzp, nyp, cn)) + 1j * rng4.standard_normal((nzp, nyp, cn)))
                 .astype(np.complex64) for _ in range(27)]
        np.concatenate([p.reshape(-1) for p in parts]).view(np.float32).tofile(os.path.join(d, "state4d_in.bin"))

    r = subprocess.run([BIN, d], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr

    # computeCircCorrT
    circ = np.fromfile(os.path.join(d, "circ.bin"), dtype=np.float32)
    assert np.max(np.abs(circ - prob.circ_corr_t)) < 1e-6

    # computeDataVariance (bit-exact), setupErrorCorrelation
    lines = open(os.path.join(d, "out_scalars.txt")).read().split("\n")
    nfail, simfail, _ = (int(x) for x in lines[0].split())
    dv = np.array(lines[1].split(), dtype=np.float64).astype(np.float32)
    ev = np.array(lines[2].split(), dtype=np.float64).astype(np.float32)
    etc = np.array(lines[3].split(), dtype=np.float64).reshape(nt, nt)
    v_rand, v_out = (float(x) for x in lines[4].split())
    dv_ref = np.array([oracle.data_variance(g.numpy(), c.nx, c.ny, c.nz, nxp, nyp, nzp)[0] for g in prob.data],
                      dtype=np.float32)
    assert np.array_equal(dv, dv_ref)
    ev_ref = (dv_ref / np.array(c.snr, dtype=np.float32)).astype(np.float32)
    assert np.array_equal(ev, ev_ref)
    th = np.array(c.thetas_deg, dtype=np.float32)
    ang = np.exp(-3.0 * np.abs(th[:, None] - th[None, :]).astype(np.float64) / c.ang_range_deg)
    etc_ref = (np.sqrt(ev_ref)[:, None] * np.sqrt(ev_ref)[None, :] * ang).astype(np.float32).astype(np.float64)
    assert np.allclose(etc, etc_ref, rtol=1e-6)
    assert v_out == -99999.0

    # the same inversion on the oracle: covariance fill, FFT, float loop, inverse FFT
    pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
    cv = [oracle.fill_param_corr(nxp, nyp, nzp, prob.corr_xy, prob.circ_corr_t, 0.0, 0.0,
                                 np.float32(prob.sigma0[a, b])) for a, b in pairs]
    ec = oracle.fill_err_corr(nxp, nyp, nzp, prob.corr_xy, 0.0, 0.0)
    mu = [g.numpy().copy() for g in prob.mu]
    dd = [g.numpy().copy() for g in prob.data]
    n3 = (nxp, nyp, nzp)
    for g in mu + dd:
        oracle.fft_r2c(g, *n3, False)
    for g in cv + [ec]:
        oracle.fft_r2c(g, *n3, True)
    tabs = oracle.perk_tables(nt, c.nz, nzp, prob.lz, c.low_cut, c.high_cut, 1.0, prob.wavelets, prob.wnorm)
    nf_ref = oracle.posterior(*n3, prob.A, etc_ref, c.wnc, prob.sigma0, tabs, mu, dd, cv, ec)
    for g in mu + dd:
        oracle.fft_c2r(g, *n3, False)
    assert nfail == nf_ref
    for i in range(3):
        got = _read(os.path.join(d, f"out_mu{i}.bin"), shape)
        assert rel_l2(got[:, :, :nxp], mu[i][:, :, :nxp]) < 5e-5, i
    for l in range(nt):
        got = _read(os.path.join(d, f"out_res{l}.bin"), shape)
        assert rel_l2(got[:, :, :nxp], dd[l][:, :, :nxp]) < 5e-5, l
    cn = nxp // 2 + 1
    for q in range(6):
        got = _read(os.path.join(d, f"out_cov{q}.bin"), shape).view(np.complex64)
        ref = cv[q].view(np.complex64)
        assert rel_l2(got[:, :, :cn], ref[:, :, :cn]) < 5e-5, q
    # the realisation: points whose posterior covariance has no float Cholesky
    # factor get a zero seed.  They lie where the covariance is at rounding
    # level (the low-cut k-planes, where the prior is ~0), so which of those
    # points fail depends on rounding: both counts stay within that set.
    # Elsewhere the field is finite: the posterior mean plus a field no
    # larger than the prior.
    zeros = [np.zeros(cn * nyp * nzp, dtype=np.complex64) for _ in range(3)]
    sim_ref = oracle.simulate_points([g.view(np.complex64)[:, :, :cn].reshape(-1).copy() for g in cv], zeros)
    caa = np.abs(cv[0].view(np.complex64)[:, :, :cn].real)
    noise_level = int((caa <= 1e-5 * caa.max()).sum())
    assert simfail <= noise_level and sim_ref <= noise_level, (simfail, sim_ref, noise_level)
    # STORM files: header of Simbox::getStormHeader (NRLib::ToString(x, 6) is
    # fixed with 6 decimals), values k, j, i, big-endian or "%.6f", then "0"
    mu0 = _read(os.path.join(d, "out_mu0.bin"), shape)
    ctype = int(open(os.path.join(d, "storm_type.txt")).read())
    f6 = lambda v: "%.6f" % v

    def storm_head(kind, geom, n3):
        return (f"storm_petro_{kind}\n0 {ctype} {f6(-99999.0)}\nFFTGrid\n{geom}\n"
                f"{f6(1024.0)} {f6(0.3 * 180 / math.pi)}\n\n{n3[0]} {n3[1]} {n3[2]}\n")
    geo = f"{f6(456000.5)} {f6(11200.0)} {f6(6780000.25)} {f6(5600.0)}"
    want = storm_head("binary", geo + " top_surface base_surface 0.0 0.0", (c.nx, c.ny, c.nz)).encode()
    want += mu0[:c.nz, :c.ny, :c.nx].astype(">f4").tobytes() + b"0\n"
    assert open(os.path.join(d, "mu0.storm"), "rb").read() == want
    want = storm_head("binary", geo + " top_surface base_surface 0.0 0.0", (nxp, nyp, nzp)).encode()
    want += mu0[:, :, :nxp].astype(">f4").tobytes() + b"0\n"
    assert open(os.path.join(d, "mu0_pad.storm"), "rb").read() == want
    lines = ["%s\n" % " ".join(f6(float(x)) for x in mu0[k, j, :c.nx]) for k in range(c.nz) for j in range(c.ny)]
    want = storm_head("ascii", geo + f" 0.0 {f6(1024.0)} 0.0 0.0", (c.nx, c.ny, c.nz)) + "".join(lines) + "0\n"
    assert open(os.path.join(d, "mu0_flat.txt")).read() == want
    if oracle.ref_segy_available():
        # and byte for byte as the reference's writer on NRLib's ToString and
        # WriteBinaryFloat writes them
        head = dict(xmin=456000.5, lx=11200.0, ymin=6780000.25, ly=5600.0, lz=1024.0, angle=0.3,
                    top="top_surface", bot="base_surface")
        oracle.write_storm_ref(os.path.join(d, "ref_mu0"), mu0, c.nx, c.ny, c.nz, ctype, head)
        oracle.write_storm_ref(os.path.join(d, "ref_mu0_pad"), mu0, nxp, nyp, nzp, ctype, head)
        oracle.write_storm_ref(os.path.join(d, "ref_mu0_flat"), mu0, c.nx, c.ny, c.nz, ctype, head, ascii=True,
                               flat=True)
        for a, b in (("mu0.storm", "ref_mu0.storm"), ("mu0_pad.storm", "ref_mu0_pad.storm"),
                     ("mu0_flat.txt", "ref_mu0_flat.txt")):
            assert open(os.path.join(d, a), "rb").read() == open(os.path.join(d, b), "rb").read(), a

    # State4D::merge
    if name == "small":
        got = np.fromfile(os.path.join(d, "state4d_out.bin"), dtype=np.complex64).reshape(9, -1)
        mu_ref = oracle.merge_mean_4d(parts[0:3], parts[3:6])
        cov_ref = oracle.merge_cov_4d(parts[6:12], parts[12:18], parts[18:27])
        for q, ref in enumerate(mu_ref + cov_ref):
            assert np.array_equal(got[q], ref.reshape(-1)), q

    # CRAVA binary files
    assert np.array_equal(_read(os.path.join(d, "crava_read.bin"), shape), crava_in)
    raw = open(os.path.join(d, "mu0.crava"), "rb").read()
    head = b"crava_fftgrid_binary\n"
    assert raw.startswith(head)
    fields = struct.unpack_from(">4d2i7d3i", raw, len(head))
    assert fields == (456000.5, 6780000.25, 25.0, 12.5, c.nx, c.ny, 1001.0, 2002.0, 0.5, -0.25, 0.125, 2.0, 0.3,
                      rnxp, nyp, nzp)
    body = np.frombuffer(raw, dtype=">f4", offset=len(head) + struct.calcsize(">4d2i7d3i"))
    assert np.array_equal(body.astype(np.float32).reshape(shape), _read(os.path.join(d, "out_mu0.bin"), shape))
    err = open(os.path.join(d, "crava_err.txt")).read()
    assert "Grid dimension is wrong" in err and "missing.crava" in err

    # computeSyntSeismic of the posterior means the mirror wrote
    means = [_read(os.path.join(d, f"out_mu{i}.bin"), shape) for i in range(3)]
    sf = np.fromfile(os.path.join(d, "stretch.bin"), dtype=np.float32).reshape(c.ny, c.nx)
    ts = np.fromfile(os.path.join(d, "shift.bin"), dtype=np.float32).reshape(c.ny, c.nx)
    stretch = [sf] + [None] * (nt - 1)
    shift = [None] * (nt - 1) + [ts]
    syn_ref = oracle.synthetic_seismic(c.nx, c.ny, c.nz, nxp, nyp, nzp, prob.A, prob.wavelets, c.dt, *means,
                                       shift=shift, stretch=stretch)
    for l in range(nt):
        got = _read(os.path.join(d, f"out_syn{l}.bin"), shape)
        assert rel_l2(got[:, :c.ny, :c.nx], syn_ref[l][:, :c.ny, :c.nx]) < 1e-6, l
        pad = np.ones(shape[1:], bool)
        pad[:c.ny, :c.nx] = False
        assert np.array_equal(got[:, pad], syn_ref[l][:, pad]), l
    # computeVariances on that synthetic seismic with energy matching on
    # angle 0: the unfiltered temporal correlation of covAlpha, the
    # variances and the gain against the oracle, the scaled wavelet
    corr_u = np.fromfile(os.path.join(d, "nc_corrt.bin"), dtype=np.float32)
    circ_full = prob.circ_corr_t
    refk = np.minimum(np.arange(nzp), nzp - np.arange(nzp))
    assert np.abs(corr_u - circ_full[np.minimum(refk, len(circ_full) - 1)] / circ_full[0]).max() < 1e-6
    ev_nc = np.fromfile(os.path.join(d, "nc_errvar.bin"), dtype=np.float32)
    match = np.zeros(nt, np.uint8)
    match[0] = 1
    en = oracle.wavelet_energy(prob.wavelets, prob.A, prob.sigma0, corr_u, ev_nc, np.array(c.snr, np.float32), match)
    got_e = np.fromfile(os.path.join(d, "nc_energy.bin"), dtype=np.float32).reshape(3, nt)
    assert np.array_equal(got_e[0], en["model_var"])
    assert np.array_equal(got_e[1], en["signal_var"])
    assert np.array_equal(got_e[2], en["theo_sn"])
    assert abs(en["theo_sn"][0] / np.float32(c.snr[0]) - 1) < 1e-5 and en["gain"][0] != 1.0
    wav_nc = np.fromfile(os.path.join(d, "nc_wavelets.bin"), dtype=np.float32).reshape(nt, nzp)
    want_w = prob.wavelets.copy()
    want_w[0] = want_w[0] * en["gain"][0]
    assert np.array_equal(wav_nc, want_w)
    # divideDataByScaleWavelet + interpolateSeismic of that synthetic seismic
    # (traces (2, 1) and (3, 1) emptied, a gain table on angle 0, the shift
    # table on the last angle, relThick 0.7 + 0.3 ((5i + 3j) mod 13) / 12),
    # with the energy-matched wavelets
    ev = np.fromfile(os.path.join(d, "swd_errvar.bin"), dtype=np.float32)
    gain = np.fromfile(os.path.join(d, "swd_gain.bin"), dtype=np.float32).reshape(c.ny, c.nx)
    corr_f = np.fromfile(os.path.join(d, "swd_corrt.bin"), dtype=np.float32)
    circ_full = prob.circ_corr_t
    refk = np.minimum(np.arange(nzp), nzp - np.arange(nzp))
    want_corr = np.where(refk < c.nz, circ_full[np.minimum(refk, len(circ_full) - 1)] / circ_full[0], 0.0)
    assert np.abs(corr_f - want_corr).max() < 1e-6
    rel = np.array([[0.7 + 0.3 * ((i * 5 + j * 3) % 13) / 12.0 for i in range(nxp)] for j in range(nyp)])
    syn_in = []
    for l in range(nt):
        g = _read(os.path.join(d, f"out_syn{l}.bin"), shape).copy()
        g[:, 1, 2:4] = 0.0
        syn_in.append(g)
    div = oracle.divide_by_scale_wavelet(c.nx, c.ny, c.nz, nxp, nyp, nzp, syn_in, wav_nc, prob.wnorm, prob.A,
                                         prob.sigma0, corr_f, ev, rel, c.dt, c.nz * c.dt, c.low_cut, c.high_cut,
                                         c.wnc, shift=shift, gain=[gain] + [None] * (nt - 1))
    for l in range(nt):
        want, nf, n0 = oracle.interpolate_seismic(c.nx, c.ny, c.nz, nxp, nyp, nzp, 0.05, div[l])
        assert n0 >= 2, l
        got = _read(os.path.join(d, f"out_swd{l}.bin"), shape)[:, :, :nxp].astype(np.float64)
        ref = want[:, :, :nxp].astype(np.float64)
        assert np.abs(got - ref).max() <= 5e-6 * np.abs(ref).max(), l
    # the inversion of the divided data on the non-constant-thickness per-k
    # vectors, against the oracle's transforms and loop on the same divided
    # data.  The residual of this path amplifies transform roundoff (E M^-1
    # with e3 = 1/|W| far from flat): its bound is twice the spread between
    # the reference's own float chain (FFTW 2.1.2, lib_matr) and the
    # oracle's where oracle/_ref is built, else 5e-4
    etc_nc = (np.sqrt(ev_nc)[:, None] * np.sqrt(ev_nc)[None, :] * ang).astype(np.float32).astype(np.float64)
    ntabs = oracle.perk_tables_nonconst(nt, c.nz, nzp, c.nz * c.dt, c.low_cut, c.high_cut, float(rel.min()), wav_nc,
                                        prob.wnorm)

    def nc_chain(which):
        m = [g.numpy().copy() for g in prob.mu]
        dd_ = [_read(os.path.join(d, f"out_swd{l}.bin"), shape).copy() for l in range(nt)]
        cc = [oracle.fill_param_corr(nxp, nyp, nzp, prob.corr_xy, prob.circ_corr_t, 0.0, 0.0,
                                     np.float32(prob.sigma0[a, b])) for a, b in pairs]
        ee = oracle.fill_err_corr(nxp, nyp, nzp, prob.corr_xy, 0.0, 0.0)
        for g in m + dd_:
            oracle.fft_r2c(g, *n3, False, which)
        for g in cc + [ee]:
            oracle.fft_r2c(g, *n3, True, which)
        nf_ = oracle.posterior(*n3, prob.A, etc_nc, c.wnc, prob.sigma0, ntabs, m, dd_, cc, ee, which=which)
        for g in m + dd_:
            oracle.fft_c2r(g, *n3, False, which)
        return m, dd_, cc, nf_

    ncm, ncd, ncc, nf_nc = nc_chain("oracle")
    assert int(open(os.path.join(d, "nc_fail.txt")).read()) == nf_nc
    e_nc = {"mean": max(rel_l2(_read(os.path.join(d, f"nc_mu{i}.bin"), shape)[:, :, :nxp], ncm[i][:, :, :nxp])
                        for i in range(3)),
            "residual": max(rel_l2(_read(os.path.join(d, f"nc_res{l}.bin"), shape)[:, :, :nxp], ncd[l][:, :, :nxp])
                            for l in range(nt)),
            "cov": max(rel_l2(np.fromfile(os.path.join(d, f"nc_cov{q}.bin"), dtype=np.float32).view(np.complex64),
                              ncc[q].view(np.complex64).reshape(-1)) for q in range(6))}
    res_bound = 5e-4
    if oracle.ref_available():
        _, rd, _, _ = nc_chain("ref")
        spread = max(rel_l2(rd[l][:, :, :nxp], ncd[l][:, :, :nxp]) for l in range(nt))
        res_bound = 2.0 * spread
        print(f"  reference float chain vs oracle chain: residual {spread:.3g}", flush=True)
    print(f"  non-constant thickness: errors vs oracle {e_nc}", flush=True)
    assert e_nc["mean"] < 5e-5 and e_nc["cov"] < 5e-5
    assert e_nc["residual"] < res_bound
    # SEG-Y (all five trace header formats) and SGRI files of the posterior
    # mean of alpha, byte for byte against the reference's NRLib writer on
    # the same grid and geometry
    if oracle.ref_segy_available():
        vals = open(os.path.join(d, "segy_geom.txt")).read().split()
        keys = ("xmin", "ymin", "dx", "dy", "angle", "lz", "dz", "zmin", "zmax", "il0", "xl0", "ilStepX", "ilStepY",
                "xlStepX", "xlStepY")
        geom = {k: float(v) for k, v in zip(keys, vals)}
        z0 = np.float32(float(vals[15]))
        ntr = c.nx * c.ny
        tb = np.fromfile(os.path.join(d, "segy_geom.bin"), dtype=np.float64)
        top, bot = tb[:ntr].reshape(c.ny, c.nx), tb[ntr:].reshape(c.ny, c.nx)
        mu0 = _read(os.path.join(d, "out_mu0.bin"), shape)
        for f in range(5):
            mine = open(os.path.join(d, f"mu0_thf{f}.segy"), "rb").read()
            oracle.write_segy_ref(os.path.join(d, f"ref_thf{f}"), mu0, c.nx, c.ny, c.nz, geom, top, bot, z0, f)
            assert mine == open(os.path.join(d, f"ref_thf{f}.segy"), "rb").read(), f
            assert len(mine) > 3600 + 240
        # the SGRI header names its binary file: the reference writes under the same name
        mine_h = open(os.path.join(d, "mu0.Sgrh")).read()
        mine_b = open(os.path.join(d, "mu0.Sgri"), "rb").read()
        oracle.write_sgri_ref(os.path.join(d, "mu0"), "Posterior alpha", mu0, c.nx, c.ny, c.nz, geom, top, bot)
        assert open(os.path.join(d, "mu0.Sgrh")).read() == mine_h
        assert open(os.path.join(d, "mu0.Sgri"), "rb").read() == mine_b
    assert open(os.path.join(d, "illcond.txt")).read().strip() == "1"
    # createPostCov00 / getOrigin / updatePriorVar read single traces of the
    # time-domain covariance grids (crava_hip_grid_read_trace)
    covt = [_read(os.path.join(d, f"out_covt{q}.bin"), shape) for q in range(6)]
    assert np.array_equal(np.fromfile(os.path.join(d, "postcov00.bin"), dtype=np.float32), covt[0][:c.nz, 0, 0])
    org = np.fromfile(os.path.join(d, "origins.bin"), dtype=np.float32)
    assert np.array_equal(org, np.array([g[0, 0, 0] for g in covt], np.float32))
    pv = np.fromfile(os.path.join(d, "priorvar.bin"), dtype=np.float32).reshape(3, 3)
    assert np.array_equal(pv, org[[[0, 3, 4], [3, 1, 5], [4, 5, 2]]])
    # a realisation with the reference's RandomGen stream equals the
    # reference's: the oracle's noise from the same generator state, its
    # per-point step, c2r and the posterior mean
    st = 23665
    ref_sim = []
    covh = [_read(os.path.join(d, f"out_cov{q}.bin"), shape) for q in range(6)]
    for i in range(3):
        z, st = oracle.fill_complex_noise(nxp, nyp, nzp, st)
        ref_sim.append(z)
    oracle.simulate_points([g.view(np.complex64).reshape(-1).copy() for g in covh],
                           [z.view(np.complex64).reshape(-1) for z in ref_sim])
    assert int(open(os.path.join(d, "simref_seed.txt")).read()) == st
    for i in range(3):
        oracle.fft_c2r(ref_sim[i], nxp, nyp, nzp, False)
        mui = _read(os.path.join(d, f"out_mu{i}.bin"), shape)
        want = ref_sim[i][:, :, :nxp] + mui[:, :, :nxp]
        got = _read(os.path.join(d, f"out_simref{i}.bin"), shape)[:, :, :nxp]
        assert rel_l2(got, want) < 2e-6, i
    # correctAlphaBetaRho with the host computeG (the same call on the same
    # inputs gives the mirror's G), then the realisation with local noise
    from crava_amd.core import local_noise_g
    dg = np.array([0.6, 0.7, 0.8])
    smd = (dg[:, None] * np.asarray(prob.sigma0, dtype=np.float64)) * dg[None, :]
    Gl = local_noise_g(etc, prob.sigma0, smd)
    tabs = [np.array([[1.0 + ((i * 7 + j * 3 + a * 5) % 11) / 5.0 for i in range(c.nx)] for j in range(c.ny)])
            for a in range(nt)]
    sn, tc = lnoise.trace_matrices(Gl, tabs, etc, prob.sigma0, smd)
    postm = [_read(os.path.join(d, f"out_mu{i}.bin"), shape) for i in range(3)]
    want = [g.copy() for g in postm]
    lnoise.apply_correction(want, [g.numpy() for g in prob.mu], sn, c.nx, c.ny, c.nz)
    for i in range(3):
        got = _read(os.path.join(d, f"out_lnc{i}.bin"), shape)
        w = want[i][:c.nz, :c.ny, :c.nx]
        g = got[:c.nz, :c.ny, :c.nx]
        assert np.max(np.abs(g.astype(np.float64) - w) / np.spacing(np.abs(w))) <= 1.0, i
        assert not np.array_equal(g, postm[i][:c.nz, :c.ny, :c.nx]), i
    draws = [r.copy() for r in ref_sim]
    lnoise.transform_realisation(draws, tc, c.nx, c.ny, c.nz)
    # the realisation adds the corrected posterior means (src/crava.cpp:1568)
    for i in range(3):
        corrected = _read(os.path.join(d, f"out_lnc{i}.bin"), shape)
        want = draws[i][:, :, :nxp] + corrected[:, :, :nxp]
        got = _read(os.path.join(d, f"out_simlnc{i}.bin"), shape)[:, :, :nxp]
        assert rel_l2(got, want) < 2e-6, i
    for i in range(3):
        sim = _read(os.path.join(d, f"out_sim{i}.bin"), shape)[:, :, :nxp]
        assert np.isfinite(sim).all()
        dev = sim - mu[i][:, :, :nxp]
        assert 0.0 < float(dev.std()) < 3.0 * math.sqrt(prob.sigma0[i, i])
