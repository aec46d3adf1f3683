"""ctypes binding of the CPU checker libraries.

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, and there only as the checker.  The product
never imports this package.

* liboracle.so       -- crava_oracle.c, the C restatement (always built);
* _ref/libcrava_ref.so -- the reference's own FFTW 2.1.2 + lib_matr.c, built
  from /root/reference sources by `make ref` (optional; absent on machines
  without the reference sources unless copied with the tree).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_SO = os.path.join(HERE, "liboracle.so")
REF_SO = os.path.join(HERE, "_ref", "libcrava_ref.so")
REF_RANDOM_SO = os.path.join(HERE, "_ref", "libcrava_ref_random.so")

_orc = None
_ref = None

FP = C.POINTER(C.c_float)
DP = C.POINTER(C.c_double)
VP = C.c_void_p


def build(ref: bool = True) -> None:
    """Compile liboracle.so (and _ref when the reference sources exist)."""
    subprocess.run(["make", "-s", "-C", HERE, "-B", "liboracle.so"], check=True)
    if ref and os.path.isdir("/root/reference/libs/fft"):
        subprocess.run(["make", "-s", "-C", HERE, "-B", "ref"], check=True)


def orc():
    global _orc
    if _orc is None:
        if not os.path.exists(ORACLE_SO):
            build(ref=False)
        L = C.CDLL(ORACLE_SO)
        I, F, D, LL, PP = C.c_int, C.c_float, C.c_double, C.c_longlong, C.POINTER(VP)
        L.orc_fft3d_r2c.argtypes = [VP, I, I, I, I]
        L.orc_fft3d_c2r.argtypes = [VP, I, I, I, I]
        L.orc_fft3d_r2c_f64.argtypes = [VP, I, I, I, I]
        L.orc_fft3d_c2r_f64.argtypes = [VP, I, I, I, I]
        L.orc_dft1d.argtypes = [VP, I, I]
        L.orc_perk_tables.argtypes = [I, I, I, F, D, F, F, F, VP, VP, VP, VP, VP, VP, VP]
        L.orc_posterior.argtypes = [I, I, I, I, VP, VP, F, VP, VP, VP, VP, VP, VP, PP, PP, PP, VP]
        L.orc_posterior_f64.argtypes = [I, I, I, I, VP, VP, D, VP, VP, VP, VP, VP, VP, PP, PP, PP, VP]
        L.orc_simulate_points.argtypes = [LL, PP, PP]
        L.orc_rnorm01.argtypes = [C.POINTER(C.c_uint32)]
        L.orc_unif01.argtypes = [C.POINTER(C.c_uint32)]
        L.orc_fill_complex_noise.argtypes = [VP, I, I, I, C.POINTER(C.c_uint32)]
        L.orc_chol_cpx.argtypes = [I, VP]
        L.orc_axeqb_mat_cpx.argtypes = [I, VP, VP, I]
        L.orc_facies.argtypes = [I, I, I, I, I, I, VP, VP, VP, I, VP, I, I, I, VP, VP, VP, F, PP]
        L.orc_fill_param_corr.argtypes = [VP, I, I, I, VP, VP, F, F, F]
        L.orc_fill_err_corr.argtypes = [VP, I, I, I, VP, F, F]
        L.orc_grid_op.argtypes = [I, VP, VP, LL, I, F]
        L.orc_data_variance.argtypes = [VP, I, I, I, I, I, I, C.POINTER(C.c_longlong)]
        L.orc_data_variance.restype = C.c_float
        L.orc_facies_geomodel.argtypes = [I, I, I, I, VP, PP]
        L.orc_facies_find_density.argtypes = [D, D, D, VP, I, I, VP, VP]
        L.orc_facies_find_density.restype = C.c_double
        L.orc_facies_densdim.argtypes = [I, I, I, I, I, I, VP, VP, VP, I, I, VP, I, I, I, VP, VP, VP, VP, F, PP]
        L.orc_perk_tables_nonconst.argtypes = [I, I, I, D, F, F, F, VP, VP, VP, VP, VP, VP, VP]
        L.orc_wavelet_resample.argtypes = [VP, I, I, F, F, I, VP, VP]
        L.orc_wavelet_energy.argtypes = [I, I, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]
        L.orc_merge_mean_4d.argtypes = [LL, PP, PP, PP]
        L.orc_merge_cov_4d.argtypes = [LL, PP, PP, PP, PP]
        L.orc_synthetic_seismic.argtypes = [I, I, I, I, I, I, I, VP, VP, F, PP, PP, PP, VP, VP, VP, PP]
        L.orc_divide_by_scale_wavelet.argtypes = [I, I, I, I, I, I, I, VP, VP, VP, VP, VP, VP, VP, PP, PP,
                                                  F, D, F, F, F, PP, I, I]
        L.orc_interpolate_seismic.argtypes = [I, I, I, I, I, I, F, VP, C.POINTER(C.c_int)]
        L.orc_dist_to_boundary.argtypes = [I, I, I]
        L.orc_dist_to_boundary.restype = C.c_float
        L.orc_fill_number.argtypes = [I, I, I]
        L.orc_posterior.restype = C.c_longlong
        L.orc_posterior_f64.restype = C.c_longlong
        L.orc_simulate_points.restype = C.c_longlong
        L.orc_rnorm01.restype = C.c_double
        L.orc_unif01.restype = C.c_double
        L.orc_chol_cpx.restype = C.c_int
        _orc = L
    return _orc


def ref_available() -> bool:
    return os.path.exists(REF_SO)


def ref():
    global _ref
    if _ref is None:
        if not ref_available():
            raise FileNotFoundError(REF_SO)
        L = C.CDLL(REF_SO)
        I, F, D, LL, PP = C.c_int, C.c_float, C.c_double, C.c_longlong, C.POINTER(VP)
        L.ref_fft3d_r2c.argtypes = [VP, I, I, I, I]
        L.ref_fft3d_c2r.argtypes = [VP, I, I, I, I]
        L.ref_fft1d_r2c.argtypes = [VP, I]
        L.ref_circ_corr_t.argtypes = [VP, I, I, I, VP]
        L.ref_circ_corr_t.restype = C.c_int
        L.ref_chol_cpx.argtypes = [I, VP]
        L.ref_axeqb_mat_cpx.argtypes = [I, VP, VP, I]
        L.ref_posterior.argtypes = [I, I, I, I, VP, VP, F, VP, VP, VP, VP, VP, VP, PP, PP, PP, VP]
        L.ref_simulate_points.argtypes = [LL, PP, PP]
        L.ref_synthetic_seismic.argtypes = [I, I, I, I, I, I, I, VP, VP, F, PP, PP, PP, VP, VP, VP, PP, I, I]
        L.ref_divide_by_scale_wavelet.argtypes = [I, I, I, I, I, I, I, VP, VP, VP, VP, VP, VP, VP, PP, PP,
                                                  F, D, F, F, F, PP, I, I]
        L.ref_posterior.restype = C.c_longlong
        L.ref_simulate_points.restype = C.c_longlong
        L.ref_wall_seconds.restype = C.c_double
        L.ref_chol_cpx.restype = C.c_int
        _ref = L
    return _ref


_ref_random = None


def ref_random_available() -> bool:
    return os.path.exists(REF_RANDOM_SO)


def _random_lib():
    global _ref_random
    if _ref_random is None:
        L = C.CDLL(REF_RANDOM_SO)
        L.ref_rnorm01_stream.argtypes = [C.c_uint32, C.c_longlong, VP]
        L.ref_unif01_stream.argtypes = [C.c_uint32, C.c_longlong, VP]
        L.ref_fill_complex_noise.argtypes = [VP, C.c_int, C.c_int, C.c_int, C.c_uint32, C.c_int, VP]
        _ref_random = L
    return _ref_random


def ref_rnorm01_stream(seed: int, n: int) -> np.ndarray:
    """n draws of the reference's own RandomGen::rnorm01 from `seed`."""
    out = np.zeros(n, np.float64)
    _random_lib().ref_rnorm01_stream(seed, n, out.ctypes.data)
    return out


def ref_unif01_stream(seed: int, n: int) -> np.ndarray:
    out = np.zeros(n, np.float64)
    _random_lib().ref_unif01_stream(seed, n, out.ctypes.data)
    return out


def ref_fill_complex_noise(nxp, nyp, nzp, seed: int, extra: int = 0):
    """FFTGrid::fillInComplexNoise driven by the reference's RandomGen; also
    the next `extra` rnorm01 draws of the stream."""
    g = np.zeros((nzp, nyp, 2 * (nxp // 2 + 1)), np.float32)
    after = np.zeros(max(extra, 1), np.float64)
    _random_lib().ref_fill_complex_noise(g.ctypes.data, nxp, nyp, nzp, seed, extra, after.ctypes.data)
    return g, after[:extra]


def _ptrs(arrs):
    a = (VP * len(arrs))()
    for i, x in enumerate(arrs):
        assert x.flags["C_CONTIGUOUS"]
        a[i] = x.ctypes.data
    return a


# ---------------------------------------------------------------- transforms
def fft_r2c(grid: np.ndarray, nxp, nyp, nzp, is_cov: bool, which: str = "oracle") -> np.ndarray:
    """In-place r2c on a float32 [nzp][nyp][rnxp] array; returns complex64 view."""
    assert grid.dtype == np.float32 and grid.flags["C_CONTIGUOUS"]
    if which == "ref":
        ref().ref_fft3d_r2c(grid.ctypes.data, nxp, nyp, nzp, int(is_cov))
    else:
        orc().orc_fft3d_r2c(grid.ctypes.data, nxp, nyp, nzp, int(is_cov))
    return grid.view(np.complex64)


def fft_c2r(grid: np.ndarray, nxp, nyp, nzp, is_cov: bool, which: str = "oracle") -> np.ndarray:
    assert grid.dtype == np.float32 and grid.flags["C_CONTIGUOUS"]
    if which == "ref":
        ref().ref_fft3d_c2r(grid.ctypes.data, nxp, nyp, nzp, int(is_cov))
    else:
        orc().orc_fft3d_c2r(grid.ctypes.data, nxp, nyp, nzp, int(is_cov))
    return grid


def fft_r2c_f64(grid: np.ndarray, nxp, nyp, nzp, is_cov: bool) -> np.ndarray:
    assert grid.dtype == np.float64 and grid.flags["C_CONTIGUOUS"]
    orc().orc_fft3d_r2c_f64(grid.ctypes.data, nxp, nyp, nzp, int(is_cov))
    return grid.view(np.complex128)


def fft_c2r_f64(grid: np.ndarray, nxp, nyp, nzp, is_cov: bool) -> np.ndarray:
    assert grid.dtype == np.float64 and grid.flags["C_CONTIGUOUS"]
    orc().orc_fft3d_c2r_f64(grid.ctypes.data, nxp, nyp, nzp, int(is_cov))
    return grid


# ----------------------------------------------------------------- tables
def perk_tables(ntheta, nz, nzp, lz, low_cut, high_cut, min_rel_thick, wavelets, wnorm):
    wav = np.ascontiguousarray(wavelets, dtype=np.float32)
    wn = np.ascontiguousarray(wnorm, dtype=np.float32)
    inv = np.zeros(nzp, dtype=np.uint8)
    out = [np.zeros((nzp, ntheta, 2), dtype=np.float32) for _ in range(4)]
    orc().orc_perk_tables(ntheta, nz, nzp, 0.0, lz, low_cut,
                          high_cut, min_rel_thick, wav.ctypes.data, wn.ctypes.data,
                          inv.ctypes.data, *[o.ctypes.data for o in out])
    return {"invert": inv, "kWD": out[0], "e1": out[1], "e2": out[2], "e3": out[3]}


def perk_tables_nonconst(ntheta, nz, nzp, lz, low_cut, high_cut, min_rel_thick, wavelets, wnorm):
    """Per-k tables of a simbox that is not of constant thickness."""
    wav = np.ascontiguousarray(wavelets, dtype=np.float32)
    wn = np.ascontiguousarray(wnorm, dtype=np.float32)
    inv = np.zeros(nzp, dtype=np.uint8)
    out = [np.zeros((nzp, ntheta, 2), dtype=np.float32) for _ in range(4)]
    orc().orc_perk_tables_nonconst(ntheta, nz, nzp, lz, low_cut, high_cut, min_rel_thick, wav.ctypes.data,
                                   wn.ctypes.data, inv.ctypes.data, *[o.ctypes.data for o in out])
    return {"invert": inv, "kWD": out[0], "e1": out[1], "e2": out[2], "e3": out[3]}


def wavelet_resample(w, center, dzw, dz, nzp):
    """Wavelet::resample: (nzp samples in FFT order, norm); None where a
    kriging matrix is not positive definite."""
    w = np.ascontiguousarray(w, dtype=np.float32)
    out = np.zeros(nzp, np.float32)
    norm = np.zeros(1, np.float32)
    rc = orc().orc_wavelet_resample(w.ctypes.data, w.size, center, dzw, dz, nzp, out.ctypes.data, norm.ctypes.data)
    return None if rc != 0 else (out, float(norm[0]))


def wavelet_energy(wavelets, A, prior_var0, corr_t, error_variance, emp_sn, match=None):
    """Crava::computeVariances: dict of per-angle gain, model_var, signal_var, theo_sn."""
    wav = np.ascontiguousarray(wavelets, dtype=np.float32)
    nt, nzp = wav.shape
    A = np.ascontiguousarray(A, dtype=np.float32)
    pv = np.ascontiguousarray(prior_var0, dtype=np.float64)
    ct = np.ascontiguousarray(corr_t, dtype=np.float32)
    ev = np.ascontiguousarray(error_variance, dtype=np.float32)
    sn = np.ascontiguousarray(emp_sn, dtype=np.float32)
    m = None if match is None else np.ascontiguousarray(match, dtype=np.uint8)
    out = [np.zeros(nt, np.float32) for _ in range(4)]
    orc().orc_wavelet_energy(nt, nzp, wav.ctypes.data, A.ctypes.data, pv.ctypes.data, ct.ctypes.data, ev.ctypes.data,
                             sn.ctypes.data, None if m is None else m.ctypes.data, *[o.ctypes.data for o in out])
    return dict(zip(("gain", "model_var", "signal_var", "theo_sn"), out))


def prior_var0_robustify(prior_var0):
    """SeismicParametersHolder::setCorrelationParameters' lift of Sigma0
    (src/seismicparametersholder.cpp:114-135) with LAPACK's DGEEV (scipy's
    geev, right eigenvectors in columns), row k of the eigenvector matrix as
    the reference indexes it: returns (3x3 float64, changed)."""
    import scipy.linalg as sl
    a = np.array(prior_var0, dtype=np.float64).reshape(3, 3)
    w, vr = sl.eig(a.copy(), left=False, right=True)
    w = w.real
    vr = vr.real
    mx = w[0]
    for i in range(1, 3):
        if mx < w[i]:
            mx = w[i]
    out = a.copy()
    changed = False
    for k in range(3):
        if w[k] < mx * 0.001:
            changed = True
            for i in range(3):
                for j in range(3):
                    out[i, j] += vr[k, i] * vr[k, j] * (mx * 0.0011 - w[k])
    return out, changed


# -------------------------------------------------------------- posterior
def posterior(nxp, nyp, nzp, A, etc, wnc, sigma0, tabs, mu, data, cov, errcorr, which: str = "oracle") -> int:
    """Float per-point loop on complex64 grids (in place).  Returns #failed."""
    nt = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float32)
    etc = np.ascontiguousarray(etc, dtype=np.float64)
    s0 = np.ascontiguousarray(sigma0, dtype=np.float64)
    args = (nxp, nyp, nzp, nt, A.ctypes.data, etc.ctypes.data, wnc, s0.ctypes.data,
            tabs["invert"].ctypes.data, tabs["kWD"].ctypes.data, tabs["e1"].ctypes.data, tabs["e2"].ctypes.data,
            tabs["e3"].ctypes.data, _ptrs(mu), _ptrs(data), _ptrs(cov), errcorr.ctypes.data)
    if which == "ref":
        return int(ref().ref_posterior(*args))
    return int(orc().orc_posterior(*args))


def posterior_f64(nxp, nyp, nzp, A, etc, wnc, sigma0, tabs, mu, data, cov, errcorr) -> int:
    """FP64 per-point posterior on complex128 grids (in place)."""
    nt = A.shape[0]
    A = np.ascontiguousarray(A, dtype=np.float32)
    etc = np.ascontiguousarray(etc, dtype=np.float64)
    s0 = np.ascontiguousarray(sigma0, dtype=np.float64)
    return int(orc().orc_posterior_f64(nxp, nyp, nzp, nt, A.ctypes.data, etc.ctypes.data, wnc,
                                       s0.ctypes.data, tabs["invert"].ctypes.data, tabs["kWD"].ctypes.data,
                                       tabs["e1"].ctypes.data, tabs["e2"].ctypes.data, tabs["e3"].ctypes.data,
                                       _ptrs(mu), _ptrs(data), _ptrs(cov), errcorr.ctypes.data))


def circ_corr_t_ref(prior: np.ndarray, min_int_fq: int, nzp: int) -> tuple[np.ndarray, int]:
    """computeCircCorrT on the reference's FFTW 1-D transforms (oracle/_ref):
    (nzp lags, 1 where the reference would exit)."""
    p = np.ascontiguousarray(prior, dtype=np.float32)
    out = np.zeros(nzp, np.float32)
    bad = ref().ref_circ_corr_t(p.ctypes.data, int(p.size), int(min_int_fq), int(nzp), out.ctypes.data)
    return out, int(bad)


def chol_cpx(mat: np.ndarray, which: str = "oracle") -> int:
    """lib_matrCholCpx on a complex64 (n,n) array, in place."""
    assert mat.dtype == np.complex64 and mat.flags["C_CONTIGUOUS"]
    n = mat.shape[0]
    if which == "ref":
        return int(ref().ref_chol_cpx(n, mat.ctypes.data))
    return int(orc().orc_chol_cpx(n, mat.ctypes.data))


# ------------------------------------------------------------- simulation
def rnorm01_stream(seed: int, n: int) -> tuple[np.ndarray, int]:
    s = C.c_uint32(seed)
    out = np.array([orc().orc_rnorm01(C.byref(s)) for _ in range(n)])
    return out, s.value


def fill_complex_noise(nxp, nyp, nzp, seed: int) -> tuple[np.ndarray, int]:
    cn = nxp // 2 + 1
    g = np.zeros((nzp, nyp, 2 * cn), dtype=np.float32)
    s = C.c_uint32(seed)
    orc().orc_fill_complex_noise(g.ctypes.data, nxp, nyp, nzp, C.byref(s))
    return g, s.value


def simulate_points(cov, seed, which: str = "oracle") -> int:
    npts = cov[0].size
    if which == "ref":
        return int(ref().ref_simulate_points(npts, _ptrs(cov), _ptrs(seed)))
    return int(orc().orc_simulate_points(npts, _ptrs(cov), _ptrs(seed)))


# ----------------------------------------------------------------- facies
def facies(nx, ny, nz, nxp, nyp, nzp, alpha, beta, rho, tables, tn, vol, vn, prior, p_undef):
    nfac = len(prior)
    sm = 2 * (nx // 2 + 1)
    out = [np.zeros((nz, ny, sm), dtype=np.float32) for _ in range(nfac + 1)]
    tab = np.ascontiguousarray(tables, dtype=np.float32)
    v = np.ascontiguousarray(vol, dtype=np.float64)
    n = np.ascontiguousarray(vn, dtype=np.int32)
    pr = np.ascontiguousarray(prior, dtype=np.float32)
    orc().orc_facies(nx, ny, nz, nxp, nyp, nzp, alpha.ctypes.data, beta.ctypes.data, rho.ctypes.data, nfac,
                     tab.ctypes.data, tn[0], tn[1], tn[2], v.ctypes.data, n.ctypes.data, pr.ctypes.data,
                     p_undef, _ptrs(out))
    return out


def facies_densdim(nx, ny, nz, nxp, nyp, nzp, alpha, beta, rho, tables, tn, vols, vns, noise, prior, p_undef):
    """Facies probabilities with 2^nang density sets weighted by the angle
    weights of the noise-scale tables noise [nang][ny][nx] (float64);
    tables [ndens][nfac][tn3][tn2][rn1], vols [ndens][6], vns [ndens][3]."""
    nfac = len(prior)
    noise = np.ascontiguousarray(noise, dtype=np.float64)
    nang = 0 if noise.size == 0 else noise.shape[0]
    sm = 2 * (nx // 2 + 1)
    out = [np.zeros((nz, ny, sm), dtype=np.float32) for _ in range(nfac + 1)]
    tab = np.ascontiguousarray(tables, dtype=np.float32)
    v = np.ascontiguousarray(vols, dtype=np.float64)
    n = np.ascontiguousarray(vns, dtype=np.int32)
    pr = np.ascontiguousarray(prior, dtype=np.float32)
    nz_ = noise if nang else np.zeros(1)
    orc().orc_facies_densdim(nx, ny, nz, nxp, nyp, nzp, alpha.ctypes.data, beta.ctypes.data, rho.ctypes.data, nfac,
                             nang, tab.ctypes.data, tn[0], tn[1], tn[2], v.ctypes.data, n.ctypes.data,
                             nz_.ctypes.data, pr.ctypes.data, p_undef, _ptrs(out))
    return out


def facies_find_density(vp, vs, rho, table, tn, vol, vn) -> float:
    """One PosteriorElasticPDF3D::FindDensity lookup; table [tn3][tn2][2*(tn1//2+1)]."""
    tab = np.ascontiguousarray(table, dtype=np.float32)
    v = np.ascontiguousarray(vol, dtype=np.float64)
    n = np.ascontiguousarray(vn, dtype=np.int32)
    return float(orc().orc_facies_find_density(float(vp), float(vs), float(rho), tab.ctypes.data, tn[0], tn[1],
                                               v.ctypes.data, n.ctypes.data))


# ------------------------------------------------------------ grid fills / operators
def fill_param_corr(nxp, nyp, nzp, corr_xy, circ_t, grad_i, grad_j, scale) -> np.ndarray:
    rnxp = 2 * (nxp // 2 + 1)
    g = np.zeros((nzp, nyp, rnxp), dtype=np.float32)
    cxy = np.ascontiguousarray(corr_xy, dtype=np.float32)
    ct = np.ascontiguousarray(circ_t, dtype=np.float32)
    orc().orc_fill_param_corr(g.ctypes.data, nxp, nyp, nzp, cxy.ctypes.data, ct.ctypes.data, grad_i, grad_j, scale)
    return g


def fill_err_corr(nxp, nyp, nzp, corr_xy, grad_i, grad_j) -> np.ndarray:
    rnxp = 2 * (nxp // 2 + 1)
    g = np.zeros((nzp, nyp, rnxp), dtype=np.float32)
    cxy = np.ascontiguousarray(corr_xy, dtype=np.float32)
    orc().orc_fill_err_corr(g.ctypes.data, nxp, nyp, nzp, cxy.ctypes.data, grad_i, grad_j)
    return g


def grid_op(op, a: np.ndarray, b, transformed: bool, scalar: float = 0.0) -> np.ndarray:
    """In place on the float32 array `a` (all of it: rsize floats or csize complex)."""
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    n = a.size // 2 if transformed else a.size
    orc().orc_grid_op(op, a.ctypes.data, None if b is None else b.ctypes.data, n, int(transformed), scalar)
    return a


def data_variance(grid, nx, ny, nz, nxp, nyp, nzp):
    n = C.c_longlong(0)
    v = orc().orc_data_variance(grid.ctypes.data, nx, ny, nz, nxp, nyp, nzp, C.byref(n))
    return float(np.float32(v)), n.value


def facies_geomodel(nx, ny, nz, prior, prob):
    pr = np.ascontiguousarray(prior, dtype=np.float32)
    orc().orc_facies_geomodel(nx, ny, nz, len(prior), pr.ctypes.data, _ptrs(prob))
    return prob


# ------------------------------------------------------------ forward modelling
def _opt_tables(tabs, nt, nx, ny):
    """[nt] optional per-trace tables -> (keep-alive list, pointer array or None)."""
    if tabs is None:
        return [], None
    keep = [None if t is None else np.ascontiguousarray(t, dtype=np.float32).reshape(ny, nx) for t in tabs]
    arr = (VP * nt)()
    for l, t in enumerate(keep):
        arr[l] = None if t is None else t.ctypes.data
    return keep, arr


def synthetic_seismic(nx, ny, nz, nxp, nyp, nzp, A, wavelets, dz, alpha, beta, rho, shift=None, gain=None,
                      stretch=None, which: str = "oracle", i_range=(0, 0)):
    """Crava::computeSyntSeismic on real padded float grids; returns the
    ntheta synthetic grids.  shift/gain/stretch: None or a list of ntheta
    entries, each None or [ny][nx].  which="ref" runs the reference's FFTW
    (traces with i in i_range only, (0, 0) = all)."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    nt = A.shape[0]
    wav = np.ascontiguousarray(wavelets, dtype=np.float32)
    grids = [np.ascontiguousarray(g, dtype=np.float32) for g in (alpha, beta, rho)]
    out = [np.zeros_like(grids[0]) for _ in range(nt)]
    keep = [_opt_tables(t, nt, nx, ny) for t in (shift, gain, stretch)]
    args = (nx, ny, nz, nxp, nyp, nzp, nt, A.ctypes.data, wav.ctypes.data, dz, keep[0][1], keep[1][1], keep[2][1],
            grids[0].ctypes.data, grids[1].ctypes.data, grids[2].ctypes.data, _ptrs(out))
    if which == "ref":
        ref().ref_synthetic_seismic(*args, i_range[0], i_range[1])
    else:
        orc().orc_synthetic_seismic(*args)
    return out


# ------------------------------------------------ SEG-Y / SGRI (reference NRLib)
REF_SEGY_SO = os.path.join(HERE, "_ref", "libcrava_ref_segy.so")
_ref_segy = None


def ref_segy_available() -> bool:
    return os.path.exists(REF_SEGY_SO)


def _segy_lib():
    global _ref_segy
    if _ref_segy is None:
        L = C.CDLL(REF_SEGY_SO)
        I, F, D, S = C.c_int, C.c_float, C.c_double, C.c_char_p
        L.ref_write_segy.argtypes = [S, VP, I, I, I, I, I] + [D] * 15 + [VP, VP, F, I]
        L.ref_write_sgri.argtypes = [S, S, VP, I, I, I, I, I] + [D] * 9 + [VP, VP]
        L.ref_write_storm.argtypes = [S, VP, I, I, I, I, I, I] + [D] * 6 + [S, S, I, I, I]
        _ref_segy = L
    return _ref_segy


_GEOM_KEYS = ("xmin", "ymin", "dx", "dy", "angle", "lz", "dz", "zmin", "zmax",
              "il0", "xl0", "ilStepX", "ilStepY", "xlStepX", "xlStepY")


def write_segy_ref(path, grid, nx, ny, nz, geom, top, bot, z0, thf=0):
    """FFTGrid::writeSegyFile through the reference's NRLib SegY: writes
    path + ".segy".  grid: real padded float [nzp][nyp][rnxp]; geom: dict of
    the simbox geometry (xmin ... xlStepY); top/bot: [ny][nx] float64."""
    g = np.ascontiguousarray(grid, dtype=np.float32)
    t = np.ascontiguousarray(top, dtype=np.float64)
    b = np.ascontiguousarray(bot, dtype=np.float64)
    rc = _segy_lib().ref_write_segy(path.encode(), g.ctypes.data, nx, ny, nz, g.shape[2], g.shape[1],
                                    *[float(geom[k]) for k in _GEOM_KEYS], t.ctypes.data, b.ctypes.data, z0, thf)
    assert rc == 0


def write_sgri_ref(path, label, grid, nx, ny, nz, geom, top, bot):
    """FFTGrid::writeSgriFile with NRLib's float writer: path + ".Sgrh" / ".Sgri"."""
    g = np.ascontiguousarray(grid, dtype=np.float32)
    t = np.ascontiguousarray(top, dtype=np.float64)
    b = np.ascontiguousarray(bot, dtype=np.float64)
    rc = _segy_lib().ref_write_sgri(path.encode(), label.encode(), g.ctypes.data, nx, ny, nz, g.shape[2], g.shape[1],
                                    *[float(geom[k]) for k in _GEOM_KEYS[:9]], t.ctypes.data, b.ctypes.data)
    assert rc == 0


def write_storm_ref(path, grid, nx, ny, nz, cubetype, head, ascii=False, flat=False, scientific=False):
    """FFTGrid::writeStormFile with Simbox::getStormHeader on NRLib's ToString
    and WriteBinaryFloat: path + ".storm" (binary) or path + ".txt" (ascii).
    grid: real padded float [nzp][nyp][rnxp]; (nx, ny, nz) the cube written
    (the padded sizes for a padded file); head: dict xmin, lx, ymin, ly, lz,
    angle (radians), top, bot (surface names)."""
    g = np.ascontiguousarray(grid, dtype=np.float32)
    rc = _segy_lib().ref_write_storm(path.encode(), g.ctypes.data, nx, ny, nz, g.shape[2], g.shape[1], cubetype,
                                     *[float(head[k]) for k in ("xmin", "lx", "ymin", "ly", "lz", "angle")],
                                     head.get("top", "").encode(), head.get("bot", "").encode(), int(flat),
                                     int(ascii), int(scientific))
    assert rc == 0


# ----------------------------------------------- non-constant-thickness path
def divide_by_scale_wavelet(nx, ny, nz, nxp, nyp, nzp, grids, wavelets, wnorm, A, var0, corr_t, err_var,
                            rel_thick, dz, lz, low_cut, high_cut, wnc, shift=None, gain=None,
                            which: str = "oracle", i_range=(0, 0)):
    """Crava::divideDataByScaleWavelet on ntheta real padded float grids;
    returns the divided copies.  rel_thick: [nyp][nxp] float64; shift/gain:
    None or ntheta entries, each None or [ny][nx].  which="ref" runs the
    reference's FFTW 2.1.2 with the global wavelet's per-trace round trip.
    Traces with i in i_range only ((0, 0) = all); the others are left as
    given."""
    A = np.ascontiguousarray(A, dtype=np.float32)
    nt = A.shape[0]
    out = [np.array(g, dtype=np.float32, order="C", copy=True) for g in grids]
    wav = np.ascontiguousarray(wavelets, dtype=np.float32)
    wn = np.ascontiguousarray(wnorm, dtype=np.float32)
    v0 = np.ascontiguousarray(var0, dtype=np.float64)
    ct = np.ascontiguousarray(corr_t, dtype=np.float32)
    ev = np.ascontiguousarray(err_var, dtype=np.float32)
    rt = np.ascontiguousarray(rel_thick, dtype=np.float64)
    assert wav.shape == (nt, nzp) and ct.shape == (nzp,) and rt.shape == (nyp, nxp)
    keep = [_opt_tables(t, nt, nx, ny) for t in (shift, gain)]
    args = (nx, ny, nz, nxp, nyp, nzp, nt, wav.ctypes.data, wn.ctypes.data, A.ctypes.data, v0.ctypes.data,
            ct.ctypes.data, ev.ctypes.data, rt.ctypes.data, keep[0][1], keep[1][1], dz, lz, low_cut, high_cut,
            wnc, _ptrs(out))
    if which == "ref":
        ref().ref_divide_by_scale_wavelet(*args, i_range[0], i_range[1])
    else:
        orc().orc_divide_by_scale_wavelet(*args, i_range[0], i_range[1])
    return out


def interpolate_seismic(nx, ny, nz, nxp, nyp, nzp, threshold, grid):
    """FFTGrid::interpolateSeismic on a real padded float grid; returns
    (new grid, traces flagged, traces with zero energy)."""
    g = np.array(grid, dtype=np.float32, order="C", copy=True)
    nz0 = C.c_int(0)
    n = orc().orc_interpolate_seismic(nx, ny, nz, nxp, nyp, nzp, threshold, g.ctypes.data, C.byref(nz0))
    return g, n, nz0.value


def dist_to_boundary(i, n, np_):
    return orc().orc_dist_to_boundary(i, n, np_)


def fill_number(i, n, np_):
    return orc().orc_fill_number(i, n, np_)


# ------------------------------------------------------------ time-lapse merge
def merge_mean_4d(ms, md):
    """State4D::merge mean loop on lists of 3 complex64 arrays; returns 3 arrays."""
    ms = [np.ascontiguousarray(x, dtype=np.complex64) for x in ms]
    md = [np.ascontiguousarray(x, dtype=np.complex64) for x in md]
    out = [np.empty_like(ms[0]) for _ in range(3)]
    orc().orc_merge_mean_4d(ms[0].size, _ptrs(ms), _ptrs(md), _ptrs(out))
    return out


def merge_cov_4d(ss, dd, sd):
    """State4D::mergeCov on complex64 arrays: ss[6], dd[6] (aa ab ar bb br rr),
    sd[9] (sd[3a+b]); returns the 6 merged grids."""
    cv = lambda L_: [np.ascontiguousarray(x, dtype=np.complex64) for x in L_]
    ss, dd, sd = cv(ss), cv(dd), cv(sd)
    out = [np.empty_like(ss[0]) for _ in range(6)]
    orc().orc_merge_cov_4d(ss[0].size, _ptrs(ss), _ptrs(dd), _ptrs(sd), _ptrs(out))
    return out
