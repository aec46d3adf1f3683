"""ctypes binding of libcrava_hip.so (the C ABI in include/crava_hip.h).

The library is loaded from this package directory.  If it is missing the
import fails loudly: there is no CPU fallback for any entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libcrava_hip.so")

OK = 0
E_INVALID = -1
E_OOM = -2
E_HIP = -3
E_STATE = -4

DATA, PARAMETER, COVARIANCE = 0, 1, 2
OP_ADD, OP_SUBTRACT, OP_MULTIPLY, OP_CHANGE_SIGN, OP_CONJUGATE, OP_MULTIPLY_SCALAR, OP_ADD_SCALAR, \
    OP_FILL_CONSTANT, OP_COPY = range(9)
F32, MIXED, F64 = 0, 1, 2
F32_DEFLATE = 3   # transforms only: float arithmetic with mean deflation

_lib = None


class CravaHipError(RuntimeError):
    pass


class PostParams(C.Structure):
    _fields_ = [
        ("nxp", C.c_int), ("nyp", C.c_int), ("nzp", C.c_int), ("ntheta", C.c_int),
        ("wnc", C.c_float),
        ("sigma0", C.c_double * 9),
        ("A", C.c_float * 24),
        ("err_theta_cov", C.c_double * 64),
        ("mixed_threshold", C.c_double),
    ]


VP = C.c_void_p
I = C.c_int
LL = C.c_longlong


class BatchVolume(C.Structure):
    _fields_ = [
        ("post", VP),
        ("mu", VP * 3),
        ("data", VP * 8),
        ("cov", VP * 6),
        ("errcorr", VP),
        ("corr", VP),
    ]

class SyntParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("nxp", C.c_int), ("nyp", C.c_int), ("nzp", C.c_int),
        ("ntheta", C.c_int),
        ("dz", C.c_float),
        ("A", VP),
        ("wavelets", VP),
        ("shift", VP),
        ("gain", VP),
        ("stretch", VP),
    ]


class SwdivParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("nxp", C.c_int), ("nyp", C.c_int), ("nzp", C.c_int),
        ("ntheta", C.c_int),
        ("wavelets", VP),
        ("wavelet_norm", VP),
        ("A", VP),
        ("prior_var0", VP),
        ("corr_t", VP),
        ("err_var", VP),
        ("rel_thick", VP),
        ("shift", VP),
        ("gain", VP),
        ("dz", C.c_float),
        ("lz", C.c_double),
        ("low_cut", C.c_float),
        ("high_cut", C.c_float),
        ("wnc", C.c_float),
    ]


class LocalNoiseParams(C.Structure):
    _fields_ = [
        ("nx", C.c_int), ("ny", C.c_int), ("nz", C.c_int),
        ("nxp", C.c_int), ("nyp", C.c_int), ("nzp", C.c_int),
        ("ntheta", C.c_int),
        ("G", VP),
        ("errThetaCov", VP),
        ("priorVar0", VP),
        ("postVar0", VP),
        ("noiseScale", VP),
    ]


# name -> (restype, argtypes); every function declared in include/crava_hip.h
SIGNATURES = {
    "crava_hip_version": (I, []),
    "crava_hip_last_error": (C.c_char_p, []),
    "crava_hip_ctx_create": (I, [C.POINTER(VP), VP]),
    "crava_hip_ctx_destroy": (I, [VP]),
    "crava_hip_ctx_set_stream": (I, [VP, VP]),
    "crava_hip_synchronize": (I, [VP]),
    "crava_hip_grid_alloc": (I, [VP, I, I, I, I, C.POINTER(VP)]),
    "crava_hip_grid_free": (I, [VP]),
    "crava_hip_grid_upload": (I, [VP, VP, VP, C.c_size_t]),
    "crava_hip_grid_download": (I, [VP, VP, VP, C.c_size_t]),
    "crava_hip_fft_r2c": (I, [VP, VP, I, I, I, I, I]),
    "crava_hip_fft_c2r": (I, [VP, VP, I, I, I, I, I]),
    "crava_hip_fft_pass": (I, [VP, VP, I, I, I, I, I, I, C.c_double]),
    "crava_hip_fft_pass_many": (I, [VP, C.POINTER(VP), I, I, I, I, I, I, I, C.c_double]),
    "crava_hip_perk_tables": (I, [I, I, I, C.c_double, C.c_float, C.c_float, C.c_float, VP, VP, VP, VP, VP, VP, VP]),
    "crava_hip_post_create": (I, [VP, C.POINTER(PostParams), VP, VP, VP, VP, VP, C.POINTER(VP)]),
    "crava_hip_post_destroy": (I, [VP]),
    "crava_hip_post_run": (I, [VP, I, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), VP,
                               C.POINTER(LL), C.POINTER(LL)]),
    "crava_hip_post_run_static": (I, [VP, I, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), VP,
                                      C.POINTER(LL), C.POINTER(LL)]),
    "crava_hip_post_run_separable": (I, [VP, I, C.POINTER(VP), C.POINTER(VP), VP, VP, C.POINTER(VP),
                                         C.POINTER(LL), C.POINTER(LL)]),
    "crava_hip_simulate": (I, [VP, I, I, I, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), C.POINTER(LL)]),
    "crava_hip_simulate_points": (I, [VP, LL, C.POINTER(VP), C.POINTER(VP), C.POINTER(LL)]),
    "crava_hip_fill_complex_noise": (I, [VP, VP, I, I, I, C.c_uint64, C.c_uint64]),
    "crava_hip_facies": (I, [VP, I, I, I, I, I, I, VP, VP, VP, I, VP, I, I, I, VP, VP, VP, C.c_float,
                             C.POINTER(VP), VP, VP]),
    "crava_hip_nll_grad": (I, [VP, I, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), VP, VP, VP, VP,
                               C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "crava_hip_fft_r2c_many": (I, [VP, C.POINTER(VP), I, I, I, I, I, I]),
    "crava_hip_fft_c2r_many": (I, [VP, C.POINTER(VP), I, I, I, I, I, I]),
    "crava_hip_fill_param_corr": (I, [VP, C.POINTER(VP), I, C.POINTER(C.c_float), I, I, I, VP, VP, C.c_float,
                                      C.c_float]),
    "crava_hip_fill_err_corr": (I, [VP, VP, I, I, I, VP, C.c_float, C.c_float]),
    "crava_hip_err_corr_plane": (I, [VP, VP, I, I, I, VP, I]),
    "crava_hip_corr_spectrum_separable": (I, [VP, VP, I, I, I, VP, VP, I]),
    "crava_hip_post_set_errcorr_plane": (I, [VP, I]),
    "crava_hip_grid_op": (I, [VP, I, VP, VP, I, I, I, I, C.c_float]),
    "crava_hip_data_variance": (I, [VP, VP, I, I, I, I, I, I, C.POINTER(C.c_float), C.POINTER(LL)]),
    "crava_hip_facies_geomodel": (I, [VP, I, I, I, I, C.POINTER(C.c_float), C.POINTER(VP)]),
    "crava_hip_batch_create": (I, [VP, I, I, C.POINTER(BatchVolume), C.POINTER(VP)]),
    "crava_hip_batch_autotune": (I, [VP, C.POINTER(I), C.POINTER(I), C.POINTER(I), C.POINTER(C.c_float)]),
    "crava_hip_batch_set_launch": (I, [VP, I, I, I]),
    "crava_hip_batch_run": (I, [VP, I]),
    "crava_hip_batch_counts": (I, [VP, C.POINTER(LL), C.POINTER(LL)]),
    "crava_hip_batch_phase_times": (I, [VP, C.POINTER(C.c_float)]),
    "crava_hip_batch_destroy": (I, [VP]),
    "crava_hip_autotune_fft": (I, [VP, VP, I, I, I, I, C.POINTER(I), C.POINTER(C.c_float)]),
    "crava_hip_fft_set_variants": (I, [VP, I, I, I, I, C.POINTER(I)]),
    "crava_hip_synthetic_seismic": (I, [VP, C.POINTER(SyntParams), VP, VP, VP, C.POINTER(VP)]),
    "crava_hip_grid_read_trace": (I, [VP, VP, I, I, I, I, I, I, I, VP]),
    "crava_hip_complex_noise_reference": (I, [I, I, I, C.POINTER(C.c_uint32), VP]),
    "crava_hip_reflection_matrix": (I, [I, VP, VP, C.c_double, VP]),
    "crava_hip_perk_tables_nonconst": (I, [I, I, I, C.c_double, C.c_float, C.c_float, C.c_float, VP, VP, VP, VP, VP,
                                           VP, VP]),
    "crava_hip_wavelet_resample": (I, [VP, I, I, C.c_float, C.c_float, I, VP, C.POINTER(C.c_float)]),
    "crava_hip_wavelet_energy": (I, [I, I, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP, VP]),
    "crava_hip_prior_var0_robustify": (I, [VP, VP, C.POINTER(I)]),
    "crava_hip_facies_local_noise": (I, [VP, I, I, I, I, I, I, VP, VP, VP, I, I, VP, I, I, I, VP, VP, VP, VP,
                                         C.c_float, C.POINTER(VP), VP, VP]),
    "crava_hip_kde_table": (I, [VP, LL, VP, VP, VP, VP, I, I, I, VP, VP]),
    "crava_hip_facies_densities": (I, [VP, I, VP, VP, VP, VP, I, I, VP, VP, I, I, I, I, C.POINTER(VP), VP, VP]),
    "crava_hip_local_noise_sigmanew": (I, [I, VP, VP, VP, VP, VP, VP]),
    "crava_hip_padding_size": (I, [I, C.c_double, C.POINTER(I)]),
    "crava_hip_circ_corr_t": (I, [VP, I, I, I, VP]),
    "crava_hip_estimate_xy_padding": (I, [I, I, C.c_double, C.c_double, C.c_double, C.c_double, C.c_double,
                                          C.c_double, I, C.c_float, C.c_float, C.c_float, C.POINTER(I), C.POINTER(I),
                                          C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "crava_hip_estimate_z_padding": (I, [I, C.c_double, C.c_double, C.c_double, I, C.c_float, C.POINTER(I),
                                         C.POINTER(C.c_double)]),
    "crava_hip_divide_by_scale_wavelet": (I, [VP, C.POINTER(SwdivParams), C.POINTER(VP)]),
    "crava_hip_interpolate_seismic": (I, [VP, I, I, I, I, I, I, C.c_float, VP, C.POINTER(I), C.POINTER(I)]),
    "crava_hip_local_noise_g": (I, [I, VP, VP, VP, VP]),
    "crava_hip_local_noise_correct": (I, [VP, C.POINTER(LocalNoiseParams), C.POINTER(VP), C.POINTER(VP), VP]),
    "crava_hip_local_noise_realisation": (I, [VP, I, I, I, I, I, I, VP, C.POINTER(VP)]),
    "crava_hip_merge_mean_4d": (I, [VP, I, I, I, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP)]),
    "crava_hip_merge_cov_4d": (I, [VP, I, I, I, C.POINTER(VP), C.POINTER(VP), C.POINTER(VP), C.POINTER(VP)]),
}


def lib():
    """Load libcrava_hip.so once; raise if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise CravaHipError(
                f"{LIB_PATH} is missing; build it with `python -m crava_amd.build` "
                "(there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc: int, what: str = "") -> None:
    if rc != OK:
        msg = lib().crava_hip_last_error()
        raise CravaHipError(f"{what} failed with {rc}: {msg.decode() if msg else ''}")


def ptr_array(ptrs):
    arr = (VP * max(1, len(ptrs)))()
    for i, p in enumerate(ptrs):
        arr[i] = p
    return arr
