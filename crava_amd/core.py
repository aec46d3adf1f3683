"""torch-facing wrappers over the C ABI (device memory and streams only).

Every call goes to libcrava_hip.so; tensors must live on the GPU.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from ._lib import check, lib, ptr_array


def _stream_ptr(stream=None) -> int:
    s = stream if stream is not None else torch.cuda.current_stream()
    return int(s.cuda_stream)


def _dev_ptr(t: torch.Tensor) -> int:
    if not t.is_cuda:
        raise L.CravaHipError("tensor must be on the GPU")
    if not t.is_contiguous():
        raise L.CravaHipError("tensor must be contiguous")
    return int(t.data_ptr())


def precision_of(t: torch.Tensor) -> int:
    if t.dtype == torch.float64:
        return L.F64
    if t.dtype == torch.float32:
        return L.F32
    raise L.CravaHipError(f"unsupported grid dtype {t.dtype}")


class Context:
    """A crava_hip_ctx bound to a torch stream."""

    def __init__(self, stream=None):
        self.h = C.c_void_p()
        check(lib().crava_hip_ctx_create(C.byref(self.h), _stream_ptr(stream)), "crava_hip_ctx_create")

    def close(self):
        if self.h:
            lib().crava_hip_ctx_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream):
        check(lib().crava_hip_ctx_set_stream(self.h, _stream_ptr(stream)), "crava_hip_ctx_set_stream")

    def synchronize(self):
        check(lib().crava_hip_synchronize(self.h), "crava_hip_synchronize")

    def r2c(self, grid: torch.Tensor, nxp: int, nyp: int, nzp: int, cubetype: int, precision: int | None = None):
        """precision: None -> from the dtype; L.MIXED -> float grid, double arithmetic."""
        p = precision_of(grid) if precision is None else precision
        check(lib().crava_hip_fft_r2c(self.h, _dev_ptr(grid), nxp, nyp, nzp, cubetype, p), "crava_hip_fft_r2c")

    def fft_pass(self, grid: torch.Tensor, nxp, nyp, nzp, axis: int, inverse: bool, scale: float = 1.0):
        check(lib().crava_hip_fft_pass(self.h, _dev_ptr(grid), nxp, nyp, nzp, axis, int(inverse),
                                       precision_of(grid), scale), "crava_hip_fft_pass")

    def fft_pass_many(self, grids, nxp, nyp, nzp, axis: int, inverse: bool, scale: float = 1.0,
                      precision: int | None = None):
        p = precision_of(grids[0]) if precision is None else precision
        check(lib().crava_hip_fft_pass_many(self.h, ptr_array([_dev_ptr(g) for g in grids]), len(grids), nxp, nyp,
                                            nzp, axis, int(inverse), p, scale), "crava_hip_fft_pass_many")

    def c2r(self, grid: torch.Tensor, nxp: int, nyp: int, nzp: int, cubetype: int, precision: int | None = None):
        p = precision_of(grid) if precision is None else precision
        check(lib().crava_hip_fft_c2r(self.h, _dev_ptr(grid), nxp, nyp, nzp, cubetype, p), "crava_hip_fft_c2r")

    def r2c_many(self, grids, nxp: int, nyp: int, nzp: int, cubetype: int, precision: int | None = None):
        """One transform launch per pass for a list of same-size grids."""
        p = precision_of(grids[0]) if precision is None else precision
        check(lib().crava_hip_fft_r2c_many(self.h, ptr_array([_dev_ptr(g) for g in grids]), len(grids), nxp, nyp,
                                           nzp, cubetype, p), "crava_hip_fft_r2c_many")

    def c2r_many(self, grids, nxp: int, nyp: int, nzp: int, cubetype: int, precision: int | None = None):
        p = precision_of(grids[0]) if precision is None else precision
        check(lib().crava_hip_fft_c2r_many(self.h, ptr_array([_dev_ptr(g) for g in grids]), len(grids), nxp, nyp,
                                           nzp, cubetype, p), "crava_hip_fft_c2r_many")


def fill_param_corr(ctx: Context, grids, scales, nxp, nyp, nzp, corr_xy: torch.Tensor, circ_corr_t: torch.Tensor,
                    grad_i: float = 0.0, grad_j: float = 0.0):
    """crava_hip_fill_param_corr: the prior correlation times scales[g] into every grid."""
    sc = (C.c_float * len(grids))(*[float(x) for x in scales])
    check(lib().crava_hip_fill_param_corr(ctx.h, ptr_array([_dev_ptr(g) for g in grids]), len(grids), sc, nxp, nyp,
                                          nzp, _dev_ptr(corr_xy), _dev_ptr(circ_corr_t), grad_i, grad_j),
          "crava_hip_fill_param_corr")


def fill_err_corr(ctx: Context, grid, nxp, nyp, nzp, corr_xy, grad_i: float = 0.0, grad_j: float = 0.0):
    check(lib().crava_hip_fill_err_corr(ctx.h, _dev_ptr(grid), nxp, nyp, nzp, _dev_ptr(corr_xy), grad_i, grad_j),
          "crava_hip_fill_err_corr")


def err_corr_plane(ctx: Context, plane, nxp, nyp, nzp, corr_xy, precision: int):
    """crava_hip_err_corr_plane: the error-correlation spectrum of zero
    correlation gradients as one Fourier plane, in place in `plane` (nyp rows
    of 2 (nxp/2 + 1) floats, or doubles for F64)."""
    check(lib().crava_hip_err_corr_plane(ctx.h, _dev_ptr(plane), nxp, nyp, nzp, _dev_ptr(corr_xy), precision),
          "crava_hip_err_corr_plane")


def corr_spectrum_separable(ctx: Context, grid, nxp, nyp, nzp, corr_xy, circ_t, precision: int):
    """crava_hip_corr_spectrum_separable: the transformed correlation grid of
    a separable prior with zero gradients, from the transforms of corrXY and
    circT."""
    check(lib().crava_hip_corr_spectrum_separable(ctx.h, _dev_ptr(grid), nxp, nyp, nzp, _dev_ptr(corr_xy),
                                                  _dev_ptr(circ_t), precision), "crava_hip_corr_spectrum_separable")


def grid_op(ctx: Context, op: int, grid, other, nxp, nyp, nzp, transformed: bool, scalar: float = 0.0):
    check(lib().crava_hip_grid_op(ctx.h, op, _dev_ptr(grid), None if other is None else _dev_ptr(other), nxp, nyp,
                                  nzp, int(transformed), scalar), "crava_hip_grid_op")


def data_variance(ctx: Context, grid, nx, ny, nz, nxp, nyp, nzp):
    """Returns (float variance, ndata)."""
    v, n = C.c_float(0), C.c_longlong(0)
    check(lib().crava_hip_data_variance(ctx.h, _dev_ptr(grid), nx, ny, nz, nxp, nyp, nzp, C.byref(v), C.byref(n)),
          "crava_hip_data_variance")
    return v.value, n.value


def facies_geomodel(ctx: Context, nx, ny, nz, prior, prob):
    pr = (C.c_float * len(prior))(*[float(x) for x in prior])
    check(lib().crava_hip_facies_geomodel(ctx.h, nx, ny, nz, len(prior), pr, ptr_array([_dev_ptr(g) for g in prob])),
          "crava_hip_facies_geomodel")


def perk_tables(ntheta, nz, nzp, lz, low_cut, high_cut, min_rel_thick, wavelets: np.ndarray, wnorm: np.ndarray):
    """Host per-k vectors (crava_hip_perk_tables)."""
    wav = np.ascontiguousarray(wavelets, dtype=np.float32)
    wn = np.ascontiguousarray(wnorm, dtype=np.float32)
    inv = np.zeros(nzp, dtype=np.uint8)
    out = [np.zeros((nzp, ntheta, 2), dtype=np.float32) for _ in range(4)]
    check(lib().crava_hip_perk_tables(ntheta, nz, nzp, float(lz), float(low_cut), float(high_cut),
                                      float(min_rel_thick), wav.ctypes.data, wn.ctypes.data, inv.ctypes.data,
                                      *[o.ctypes.data for o in out]), "crava_hip_perk_tables")
    return {"invert": inv, "kWD": out[0], "e1": out[1], "e2": out[2], "e3": out[3]}


def perk_tables_nonconst(ntheta, nz, nzp, lz, low_cut, high_cut, min_rel_thick, wavelets: np.ndarray,
                         wnorm: np.ndarray):
    """Host per-k vectors of a simbox that is not of constant thickness
    (crava_hip_perk_tables_nonconst): K = D A on data already divided by the
    wavelet."""
    wav = np.ascontiguousarray(wavelets, dtype=np.float32)
    wn = np.ascontiguousarray(wnorm, dtype=np.float32)
    inv = np.zeros(nzp, dtype=np.uint8)
    out = [np.zeros((nzp, ntheta, 2), dtype=np.float32) for _ in range(4)]
    check(lib().crava_hip_perk_tables_nonconst(ntheta, nz, nzp, float(lz), float(low_cut), float(high_cut),
                                               float(min_rel_thick), wav.ctypes.data, wn.ctypes.data,
                                               inv.ctypes.data, *[o.ctypes.data for o in out]),
          "crava_hip_perk_tables_nonconst")
    return {"invert": inv, "kWD": out[0], "e1": out[1], "e2": out[2], "e3": out[3]}


def wavelet_resample(wavelet: np.ndarray, center: int, dz_wavelet: float, dz: float, nzp: int):
    """Wavelet::resample (crava_hip_wavelet_resample): (nzp samples in FFT
    order, norm)."""
    w = np.ascontiguousarray(wavelet, dtype=np.float32)
    out = np.zeros(nzp, np.float32)
    norm = C.c_float(0.0)
    check(lib().crava_hip_wavelet_resample(w.ctypes.data, w.size, int(center), float(dz_wavelet), float(dz), nzp,
                                           out.ctypes.data, C.byref(norm)), "crava_hip_wavelet_resample")
    return out, norm.value


def wavelet_energy(wavelets, A, prior_var0, corr_t, error_variance, emp_sn, match=None):
    """Crava::computeVariances (crava_hip_wavelet_energy): dict of per-angle
    gain, model_var, signal_var and theo_sn."""
    wav = np.ascontiguousarray(wavelets, dtype=np.float32)
    nt, nzp = wav.shape
    A = np.ascontiguousarray(A, dtype=np.float32)
    pv = np.ascontiguousarray(prior_var0, dtype=np.float64)
    ct = np.ascontiguousarray(corr_t, dtype=np.float32)
    ev = np.ascontiguousarray(error_variance, dtype=np.float32)
    sn = np.ascontiguousarray(emp_sn, dtype=np.float32)
    m = None if match is None else np.ascontiguousarray(match, dtype=np.uint8)
    out = [np.zeros(nt, np.float32) for _ in range(4)]
    check(lib().crava_hip_wavelet_energy(nt, nzp, wav.ctypes.data, A.ctypes.data, pv.ctypes.data, ct.ctypes.data,
                                         ev.ctypes.data, sn.ctypes.data, None if m is None else m.ctypes.data,
                                         *[o.ctypes.data for o in out]), "crava_hip_wavelet_energy")
    return dict(zip(("gain", "model_var", "signal_var", "theo_sn"), out))


def prior_var0_robustify(prior_var0):
    """setCorrelationParameters' lift of an ill-conditioned Sigma0
    (crava_hip_prior_var0_robustify): (3x3 float64, changed)."""
    a = np.ascontiguousarray(prior_var0, dtype=np.float64).reshape(3, 3)
    out = np.zeros((3, 3), np.float64)
    ch = C.c_int(0)
    check(lib().crava_hip_prior_var0_robustify(a.ctypes.data, out.ctypes.data, C.byref(ch)),
          "crava_hip_prior_var0_robustify")
    return out, bool(ch.value)


def circ_corr_t(prior_corr_t, min_int_fq: int, nzp: int):
    """SeismicParametersHolder::computeCircCorrT (crava_hip_circ_corr_t):
    float32 array of nzp lags."""
    import numpy as np
    p = np.ascontiguousarray(prior_corr_t, dtype=np.float32)
    out = np.empty(nzp, np.float32)
    check(lib().crava_hip_circ_corr_t(p.ctypes.data, int(p.size), int(min_int_fq), int(nzp), out.ctypes.data),
          "crava_hip_circ_corr_t")
    return out


def padding_size(n: int, pad_fac: float) -> int:
    """ModelGeneral::setPaddingSize (crava_hip_padding_size)."""
    out = C.c_int(0)
    check(lib().crava_hip_padding_size(int(n), float(pad_fac), C.byref(out)), "crava_hip_padding_size")
    return out.value


def estimate_xy_padding(nx, ny, dx, dy, lx, ly, x_pad_fac, y_pad_fac, estimate, range1, range2, angle):
    """ModelGeneral::estimateXYPaddingSizes: (nxpad, nypad, true x factor, true y factor)."""
    a, b = C.c_int(0), C.c_int(0)
    fa, fb = C.c_double(0.0), C.c_double(0.0)
    check(lib().crava_hip_estimate_xy_padding(nx, ny, dx, dy, lx, ly, x_pad_fac, y_pad_fac, int(bool(estimate)),
                                              range1, range2, angle, C.byref(a), C.byref(b), C.byref(fa),
                                              C.byref(fb)), "crava_hip_estimate_xy_padding")
    return a.value, b.value, fa.value, fb.value


def estimate_z_padding(nz, lz, min_rel_thick, z_pad_fac, estimate, default_wavelet_length):
    """ModelGeneral::estimateZPaddingSize: (nzpad, true z factor)."""
    a, fa = C.c_int(0), C.c_double(0.0)
    check(lib().crava_hip_estimate_z_padding(nz, lz, min_rel_thick, z_pad_fac, int(bool(estimate)),
                                             default_wavelet_length, C.byref(a), C.byref(fa)),
          "crava_hip_estimate_z_padding")
    return a.value, fa.value


def kde_table(ctx: Context, d1, d2, d3, sigma, n1: int, n2: int, n3: int, limits, table: torch.Tensor):
    """A facies density table from samples (crava_hip_kde_table) into the
    device grid `table` [n3][n2][2(n1/2+1)] float32."""
    a = [np.ascontiguousarray(x, dtype=np.float64) for x in (d1, d2, d3)]
    sg = np.ascontiguousarray(sigma, dtype=np.float64).reshape(9)
    lim = np.ascontiguousarray(limits, dtype=np.float64)
    check(lib().crava_hip_kde_table(ctx.h, len(a[0]), a[0].ctypes.data, a[1].ctypes.data, a[2].ctypes.data,
                                    sg.ctypes.data, n1, n2, n3, lim.ctypes.data, _dev_ptr(table)),
          "crava_hip_kde_table")


def facies_densities(ctx: Context, alpha, beta, rho, facies_log, nfac: int, ndens: int = 1, sample_transform=None,
                     sigma_e=None, no_vs: bool = False, n=(100, 100, 50)):
    """MakePosteriorElasticPDF3D's density sets (crava_hip_facies_densities):
    (tables tensor [ndens][nfac][n3][n2][2(n1/2+1)] on the GPU, volumes
    [ndens][6], vn [ndens][3])."""
    n1, n2, n3 = n
    rn = 2 * (n1 // 2 + 1)
    tabs = torch.zeros((ndens, nfac, n3, n2, rn), dtype=torch.float32, device="cuda")
    ptrs = ptr_array([int(tabs[d, l].data_ptr()) for d in range(ndens) for l in range(nfac)])
    a, b, r = (np.ascontiguousarray(x, dtype=np.float32) for x in (alpha, beta, rho))
    fl = np.ascontiguousarray(facies_log, dtype=np.int32)
    st = None if sample_transform is None else np.ascontiguousarray(sample_transform, dtype=np.float64)
    se = None if sigma_e is None else np.ascontiguousarray(sigma_e, dtype=np.float64)
    vols = np.zeros((ndens, 6), np.float64)
    vn = np.zeros((ndens, 3), np.int32)
    check(lib().crava_hip_facies_densities(ctx.h, len(a), a.ctypes.data, b.ctypes.data, r.ctypes.data, fl.ctypes.data,
                                           nfac, ndens, None if st is None else st.ctypes.data,
                                           None if se is None else se.ctypes.data, int(no_vs), n1, n2, n3, ptrs,
                                           vols.ctypes.data, vn.ctypes.data), "crava_hip_facies_densities")
    return tabs, vols, vn


def local_noise_sigmanew(G, err_theta_cov, prior_var0, post_var0, scale):
    """newPosteriorCovPointwise's sigmanew for one scale vector
    (crava_hip_local_noise_sigmanew): 3x3 float64."""
    g = np.ascontiguousarray(G, dtype=np.float64)
    e = np.ascontiguousarray(err_theta_cov, dtype=np.float64)
    sm = np.ascontiguousarray(prior_var0, dtype=np.float64)
    smd = np.ascontiguousarray(post_var0, dtype=np.float64)
    sc = np.ascontiguousarray(scale, dtype=np.float64)
    out = np.zeros((3, 3), np.float64)
    check(lib().crava_hip_local_noise_sigmanew(g.shape[0], g.ctypes.data, e.ctypes.data, sm.ctypes.data,
                                               smd.ctypes.data, sc.ctypes.data, out.ctypes.data),
          "crava_hip_local_noise_sigmanew")
    return out


def facies_local_noise(ctx: Context, nx, ny, nz, nxp, nyp, nzp, alpha, beta, rho, tables: torch.Tensor, tn, vols, vn,
                       noise_scale, prior, p_undefined, argmax=None, counts=None):
    """Facies probabilities with the local-noise density dimension
    (crava_hip_facies_local_noise): tables [2^nang][nfac][tn3][tn2][rn1] on
    the GPU (tn = (tn1, tn2, tn3)), noise_scale a list of nang host [ny][nx]
    float64 tables.  Returns the nfac + 1 probability grids
    [nz][ny][2(nx/2+1)]."""
    nfac = tables.shape[1]
    tn1, tn2, tn3 = tn
    nang = len(noise_scale)
    out = [torch.zeros((nz, ny, 2 * (nx // 2 + 1)), dtype=torch.float32, device="cuda") for _ in range(nfac + 1)]
    ns = [np.ascontiguousarray(t, dtype=np.float64) for t in noise_scale]
    ns_arr = (C.c_void_p * max(1, nang))(*[t.ctypes.data for t in ns])
    v = np.ascontiguousarray(vols, dtype=np.float64)
    n = np.ascontiguousarray(vn, dtype=np.int32)
    pr = np.ascontiguousarray(prior, dtype=np.float32)
    check(lib().crava_hip_facies_local_noise(ctx.h, nx, ny, nz, nxp, nyp, nzp, _dev_ptr(alpha), _dev_ptr(beta),
                                             _dev_ptr(rho), nfac, nang, _dev_ptr(tables), tn1, tn2, tn3,
                                             v.ctypes.data, n.ctypes.data, ns_arr, pr.ctypes.data, p_undefined,
                                             ptr_array([_dev_ptr(o) for o in out]),
                                             None if argmax is None else _dev_ptr(argmax),
                                             None if counts is None else _dev_ptr(counts)),
          "crava_hip_facies_local_noise")
    return out


def problem_tables(prob):
    cfg = prob.cfg
    return perk_tables(prob.ntheta, cfg.nz, cfg.nzp, prob.lz, cfg.low_cut, cfg.high_cut, 1.0,
                       prob.wavelets, prob.wnorm)


class Posterior:
    """crava_hip_post: per-k tables of one problem resident on the GPU."""

    def __init__(self, ctx: Context, params, tables):
        self.ctx = ctx
        self.params = params
        self._tabs = {k: np.ascontiguousarray(v) for k, v in tables.items()}
        self.h = C.c_void_p()
        t = self._tabs
        check(lib().crava_hip_post_create(ctx.h, C.byref(params), t["invert"].ctypes.data, t["kWD"].ctypes.data,
                                          t["e1"].ctypes.data, t["e2"].ctypes.data, t["e3"].ctypes.data,
                                          C.byref(self.h)), "crava_hip_post_create")

    def close(self):
        if self.h:
            lib().crava_hip_post_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_errcorr_plane(self, plane: bool):
        """crava_hip_post_set_errcorr_plane: later runs take the error
        correlation as one Fourier plane (err_corr_plane) or as a grid."""
        check(lib().crava_hip_post_set_errcorr_plane(self.h, 1 if plane else 0), "crava_hip_post_set_errcorr_plane")

    def run(self, precision: int, mu, data, cov, errcorr, counts: bool = False, corr=None):
        """crava_hip_post_run; with `corr` (the transformed correlation grid of
        a separable prior) crava_hip_post_run_separable, `cov` then receiving
        the posterior covariance."""
        mu_p = ptr_array([_dev_ptr(t) for t in mu])
        d_p = ptr_array([_dev_ptr(t) for t in data])
        c_p = ptr_array([_dev_ptr(t) for t in cov])
        nf, n64 = C.c_longlong(0), C.c_longlong(0)
        cnt = (C.byref(nf) if counts else None, C.byref(n64) if counts else None)
        if corr is None:
            check(lib().crava_hip_post_run(self.h, precision, mu_p, d_p, c_p, _dev_ptr(errcorr), *cnt),
                  "crava_hip_post_run")
        else:
            check(lib().crava_hip_post_run_separable(self.h, precision, mu_p, d_p, _dev_ptr(corr), _dev_ptr(errcorr),
                                                     c_p, *cnt), "crava_hip_post_run_separable")
        return (nf.value, n64.value) if counts else None


def autotune_fft(ctx: Context, scratch: torch.Tensor, nxp: int, nyp: int, nzp: int, precision: int | None = None):
    """Run the device autotuner for one transform size; returns (variants, ms)."""
    v = (C.c_int * 3)()
    ms = (C.c_float * 3)()
    p = precision_of(scratch) if precision is None else precision
    check(lib().crava_hip_autotune_fft(ctx.h, _dev_ptr(scratch), nxp, nyp, nzp, p, v, ms),
          "crava_hip_autotune_fft")
    return list(v), list(ms)


def set_fft_variants(ctx: Context, nxp: int, nyp: int, nzp: int, precision: int, variants):
    v = (C.c_int * 3)(*variants)
    check(lib().crava_hip_fft_set_variants(ctx.h, nxp, nyp, nzp, precision, v), "crava_hip_fft_set_variants")


class Batch:
    """crava_hip_batch over volumes given as (Posterior, mu[3], data[nt], cov[6], errcorr[, corr])
    of time-domain GPU grids (corr: a separable prior's correlation grid, cov
    then outputs only).  The grids are registered: keep them alive."""

    def __init__(self, ctx: Context, precision: int, volumes):
        self.ctx = ctx
        self.volumes = list(volumes)
        arr = (L.BatchVolume * len(self.volumes))()
        for i, vol in enumerate(self.volumes):
            post, mu, data, cov, ec = vol[:5]
            corr = vol[5] if len(vol) > 5 else None
            v = arr[i]
            v.corr = _dev_ptr(corr) if corr is not None else None
            v.post = post.h.value
            for c in range(3):
                v.mu[c] = _dev_ptr(mu[c])
            for l, g in enumerate(data):
                v.data[l] = _dev_ptr(g)
            for q in range(6):
                v.cov[q] = _dev_ptr(cov[q])
            v.errcorr = _dev_ptr(ec)
        self.h = C.c_void_p()
        check(lib().crava_hip_batch_create(ctx.h, precision, len(self.volumes), arr, C.byref(self.h)),
              "crava_hip_batch_create")

    def close(self):
        if self.h:
            lib().crava_hip_batch_destroy(self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # posterior launch forms (crava_hip_batch_set_launch)
    FORM_SINGLE, FORM_PER_CLASS, FORM_PER_VOLUME = 0, 1, 2

    def autotune(self):
        """Returns (points_per_block, blocks, form, posterior ms)."""
        p, b, c, ms = C.c_int(0), C.c_int(0), C.c_int(0), C.c_float(0)
        check(lib().crava_hip_batch_autotune(self.h, C.byref(p), C.byref(b), C.byref(c), C.byref(ms)),
              "crava_hip_batch_autotune")
        return p.value, b.value, c.value, ms.value

    def set_launch(self, points_per_block: int, blocks: int, form: int = 1):
        check(lib().crava_hip_batch_set_launch(self.h, points_per_block, blocks, int(form)),
              "crava_hip_batch_set_launch")

    def run(self, graph: bool = True):
        check(lib().crava_hip_batch_run(self.h, int(graph)), "crava_hip_batch_run")

    def counts(self):
        n = len(self.volumes)
        nf, n64 = (C.c_longlong * n)(), (C.c_longlong * n)()
        check(lib().crava_hip_batch_counts(self.h, nf, n64), "crava_hip_batch_counts")
        return list(nf), list(n64)

    def phase_times(self):
        """One eager pass timed per phase: (forward transforms, posterior,
        inverse transforms) in ms."""
        ms = (C.c_float * 3)()
        check(lib().crava_hip_batch_phase_times(self.h, ms), "crava_hip_batch_phase_times")
        return tuple(ms)


def synthetic_seismic(ctx: Context, nx, ny, nz, nxp, nyp, nzp, A, wavelets, dz, alpha, beta, rho, out,
                      shift=None, gain=None, stretch=None):
    """Forward modelling (crava_hip_synthetic_seismic) into the float grids
    `out` (one per angle).  A [ntheta][3], wavelets [ntheta][nzp] and the
    optional per-angle [ny][nx] tables (lists with None entries allowed) are
    host arrays."""
    import numpy as np
    A = np.ascontiguousarray(A, dtype=np.float32)
    nt = A.shape[0]
    wav = np.ascontiguousarray(wavelets, dtype=np.float32)
    keep = []

    def table(t):
        if t is None:
            return None
        arr = (C.c_void_p * nt)()
        for l, x in enumerate(t):
            if x is not None:
                x = np.ascontiguousarray(x, dtype=np.float32)
                keep.append(x)
                arr[l] = x.ctypes.data
        keep.append(arr)
        return C.cast(arr, C.c_void_p)

    p = L.SyntParams(nx, ny, nz, nxp, nyp, nzp, nt, dz, A.ctypes.data, wav.ctypes.data, table(shift), table(gain),
                   table(stretch))
    check(lib().crava_hip_synthetic_seismic(ctx.h, C.byref(p), _dev_ptr(alpha), _dev_ptr(beta), _dev_ptr(rho),
                                            ptr_array([_dev_ptr(g) for g in out])), "crava_hip_synthetic_seismic")


def reflection_matrix(theta_rad, vsvp, seismic_type=None):
    """ModelAVODynamic::setupDefaultReflectionMatrix (crava_hip_reflection_matrix):
    [ntheta][3] float32 for angles in radians; seismic_type None (all PP) or
    a list of 0 (PP) / 1 (PS)."""
    import numpy as np
    th = np.ascontiguousarray(theta_rad, dtype=np.float32)
    A = np.zeros((th.size, 3), dtype=np.float32)
    st = None if seismic_type is None else np.ascontiguousarray(seismic_type, dtype=np.int32)
    check(lib().crava_hip_reflection_matrix(th.size, th.ctypes.data, None if st is None else st.ctypes.data,
                                            float(vsvp), A.ctypes.data), "crava_hip_reflection_matrix")
    return A


def divide_by_scale_wavelet(ctx: Context, nx, ny, nz, nxp, nyp, nzp, grids, wavelets, wnorm, A, var0, corr_t,
                            err_var, rel_thick, dz, lz, low_cut, high_cut, wnc, shift=None, gain=None):
    """Wavelet division of the non-constant-thickness path
    (crava_hip_divide_by_scale_wavelet) in place on the float grids `grids`
    (one per angle).  wavelets [ntheta][nzp], wnorm [ntheta], A [ntheta][3],
    var0 [3][3], corr_t [nzp], err_var [ntheta], rel_thick [nyp][nxp] and the
    optional per-angle [ny][nx] shift / gain tables are host arrays."""
    import numpy as np
    A = np.ascontiguousarray(A, dtype=np.float32)
    nt = A.shape[0]
    keep = [np.ascontiguousarray(wavelets, dtype=np.float32), np.ascontiguousarray(wnorm, dtype=np.float32), A,
            np.ascontiguousarray(var0, dtype=np.float64), np.ascontiguousarray(corr_t, dtype=np.float32),
            np.ascontiguousarray(err_var, dtype=np.float32), np.ascontiguousarray(rel_thick, dtype=np.float64)]

    def table(t):
        if t is None:
            return None
        arr = (C.c_void_p * nt)()
        for l, x in enumerate(t):
            if x is not None:
                x = np.ascontiguousarray(x, dtype=np.float32)
                keep.append(x)
                arr[l] = x.ctypes.data
        keep.append(arr)
        return C.cast(arr, C.c_void_p)

    p = L.SwdivParams(nx, ny, nz, nxp, nyp, nzp, nt, *[k.ctypes.data for k in keep[:7]], table(shift), table(gain),
                      dz, lz, low_cut, high_cut, wnc)
    check(lib().crava_hip_divide_by_scale_wavelet(ctx.h, C.byref(p), ptr_array([_dev_ptr(g) for g in grids])),
          "crava_hip_divide_by_scale_wavelet")


def interpolate_seismic(ctx: Context, nx, ny, nz, nxp, nyp, nzp, threshold, grid):
    """FFTGrid::interpolateSeismic (crava_hip_interpolate_seismic) in place on
    one float grid; returns (traces flagged, traces with zero energy)."""
    nf, n0 = C.c_int(0), C.c_int(0)
    check(lib().crava_hip_interpolate_seismic(ctx.h, nx, ny, nz, nxp, nyp, nzp, threshold, _dev_ptr(grid),
                                              C.byref(nf), C.byref(n0)), "crava_hip_interpolate_seismic")
    return nf.value, n0.value


def local_noise_g(err_theta_cov, prior_var0, post_var0):
    """Crava::computeG on the host (crava_hip_local_noise_g): G [ntheta][3],
    eigenvectors signed with their largest component positive."""
    import numpy as np
    e = np.ascontiguousarray(err_theta_cov, dtype=np.float64)
    sm = np.ascontiguousarray(prior_var0, dtype=np.float64)
    smd = np.ascontiguousarray(post_var0, dtype=np.float64)
    g = np.zeros((e.shape[0], 3))
    check(lib().crava_hip_local_noise_g(e.shape[0], e.ctypes.data, sm.ctypes.data, smd.ctypes.data, g.ctypes.data),
          "crava_hip_local_noise_g")
    return g


def local_noise_correct(ctx: Context, nx, ny, nz, nxp, nyp, nzp, G, err_theta_cov, prior_var0, post_var0,
                        noise_scales, post, mean2, trace_cov=None):
    """Crava::correctAlphaBetaRho's trace loop (crava_hip_local_noise_correct)
    in place on the posterior float grids post[3]; noise_scales: ntheta host
    arrays [ny][nx]; trace_cov: optional device float64 tensor of ny*nx*9
    that receives the realisation factors."""
    import numpy as np
    G = np.ascontiguousarray(G, dtype=np.float64)
    nt = G.shape[0]
    keep = [G, np.ascontiguousarray(err_theta_cov, dtype=np.float64),
            np.ascontiguousarray(prior_var0, dtype=np.float64), np.ascontiguousarray(post_var0, dtype=np.float64)]
    tabs = [np.ascontiguousarray(t, dtype=np.float64) for t in noise_scales]
    arr = (C.c_void_p * nt)(*[t.ctypes.data for t in tabs])
    p = L.LocalNoiseParams(nx, ny, nz, nxp, nyp, nzp, nt, *[k.ctypes.data for k in keep], C.cast(arr, C.c_void_p))
    check(lib().crava_hip_local_noise_correct(ctx.h, C.byref(p), ptr_array([_dev_ptr(g) for g in post]),
                                              ptr_array([_dev_ptr(g) for g in mean2]),
                                              None if trace_cov is None else _dev_ptr(trace_cov)),
          "crava_hip_local_noise_correct")


def local_noise_realisation(ctx: Context, nx, ny, nz, nxp, nyp, nzp, trace_cov, seed):
    """The local-noise block of Crava::simulate
    (crava_hip_local_noise_realisation) in place on the float grids seed[3]."""
    check(lib().crava_hip_local_noise_realisation(ctx.h, nx, ny, nz, nxp, nyp, nzp, _dev_ptr(trace_cov),
                                                  ptr_array([_dev_ptr(g) for g in seed])),
          "crava_hip_local_noise_realisation")


def merge_mean_4d(ctx: Context, nxp, nyp, nzp, mu_static, mu_dynamic, mu):
    """crava_hip_merge_mean_4d on transformed float grids (lists of 3)."""
    P = lambda gs: ptr_array([_dev_ptr(g) for g in gs])
    check(lib().crava_hip_merge_mean_4d(ctx.h, nxp, nyp, nzp, P(mu_static), P(mu_dynamic), P(mu)),
          "crava_hip_merge_mean_4d")


def merge_cov_4d(ctx: Context, nxp, nyp, nzp, ss, dd, sd, sigma):
    """crava_hip_merge_cov_4d on transformed float grids: ss[6], dd[6], sd[9] -> sigma[6]."""
    P = lambda gs: ptr_array([_dev_ptr(g) for g in gs])
    check(lib().crava_hip_merge_cov_4d(ctx.h, nxp, nyp, nzp, P(ss), P(dd), P(sd), P(sigma)),
          "crava_hip_merge_cov_4d")
