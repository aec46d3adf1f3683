#!/usr/bin/env python3
"""Posterior Fourier points/s of the CRAVA Fourier-domain inversion core on one MI355X.

Workload (default): 512x512x320 FFT grid (interior 448x448x256), 6 angle
stacks at 4-40 deg, per-angle S/N 1e4, 3e3, 1e3, 300, 100, 20, seeded
synthetic float32 inputs generated on the GPU.  One step is the whole
posterior path of Crava::computePostMeanResidAndFFTCov with its transforms,
for the separable prior Sigma0 x corr that SeismicParametersHolder::
initializeCorrelations builds (the six covariance grids are one correlation
grid times Sigma0_ab, so one correlation transform gives all six spectra):

  11 forward r2c (6 data, 3 prior means, the correlation, errCorr)
  + the per-point posterior over all P = 257*512*320 Fourier points
  + 15 inverse c2r (3 posterior means, 6 posterior covariances, 6 residuals)

The general form (six covariance grids transformed and read: 16 forward
transforms, as a time-lapse merged prior needs) is timed too and reported
under "general_prior".

run twice: the MIXED form (float grids; covariance transforms forward with
double arithmetic inside each pass, mean transforms in float with mean
deflation, the rest in float; FP32 square-root solve, FP64 solve for the
waves holding a point with noise/signal < threshold) and the plain FP64 form (double grids, double
transforms, double solve).  Each form's step runs two ways: as a one-volume
crava_hip_batch replayed as a hipGraph (the data/mean and covariance
transform sets on two streams) and as separate launches on one stream;
both times are reported and the faster is the form's time.  value = P * steps / time of the MIXED form.
The batching workload (48 uneven volumes as one replayed hipGraph) and the
differentiation workload (likelihood adjoint at 288x288x216, 5 angles, and
50 L-BFGS iterations) are measured after it and reported under "batching"
and "differentiation".  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0

# Launch geometry of the specialised transform kernels' variants, as in
# fft_static.hip: (pencils or row pairs per workgroup, threads, form) with
# form "false" (LDS-staged), "true" (direct global I/O in the first and last
# pass), "pipe" (direct I/O in persistent workgroups, k_sfft_pipe) or "cols"
# (direct I/O on pencil-fastest passes, k_sfft_cols).
STRIDED_VARIANTS = ([(tx, th, dio) for dio in ("false", "true") for tx, th in ((16, 256), (16, 512), (8, 256), (32, 512))]
                    + [(16, 256, "pipe"), (16, 512, "pipe"), (16, 512, "cols"), (16, 256, "cols")])
ROW_VARIANTS = [(tx, th, dio) for dio in ("false", "true") for tx, th in ((8, 256), (16, 512), (4, 256))]


def kernel_identity(key, variants, nxp, nyp, nzp, ngrid, nt):
    """(kernel name as profiled, total work-items per launch) of a bench kernel key."""
    cn = nxp // 2 + 1
    vx, vy, vz = variants
    if key == "fft_x_r2c":
        if vx < 0:
            return "crava::k_fft_rows_r2c<float, float, float>", None
        tx, th, dio = ROW_VARIANTS[vx]
        rows = nyp * nzp
        return (f"crava::k_srows_r2c<{nxp}, {tx}, {th}, {dio}, float, float, float>",
                -(-rows // (2 * tx)) * ngrid * th)
    if key in ("fft_y", "fft_z"):
        n, v = (nyp, vy) if key == "fft_y" else (nzp, vz)
        ncols, nbatch = (cn, nzp) if key == "fft_y" else (cn * nyp, 1)
        if v < 0:
            return "crava::k_fft_strided<-1, float, float, float>", None
        tx, th, dio = STRIDED_VARIANTS[v]
        if dio == "pipe":
            # persistent: the grid follows the device's occupancy
            return f"crava::k_sfft_pipe<{n}, {tx}, {th}, -1, float, float, float>", None
        if dio == "cols":
            return (f"crava::k_sfft_cols<{n}, {tx}, {th}, -1, float, float, float>",
                    -(-ncols // tx) * nbatch * ngrid * th)
        return (f"crava::k_sfft_strided<{n}, {tx}, {th}, -1, {dio}, float, float, float>",
                -(-ncols // tx) * nbatch * ngrid * th)
    mode = {"posterior_mixed": ("float", 1), "posterior_f32": ("float", 0), "posterior_f64": ("double", 0)}[key]
    return f"crava::k_posterior<{mode[0]}, {nt}, {mode[1]}>", -(-(cn * nyp) // 256) * 256 * nzp


def profiled_traffic(name, grid):
    """HBM bytes per launch from the committed counter summaries (TRAFFIC lines
    of profiles/hbm_traffic_*.txt), or None.  A launch shape whose other I/O
    form (LDS-staged or direct first/last pass, the same grid accesses) was
    the one profiled falls back to that form's figure; the pencil-fastest
    strided form to the direct-I/O strided form of its shape."""
    import glob

    def lookup(nm):
        best = None
        for f in glob.glob(os.path.join(ROOT, "profiles", "hbm_traffic_*.txt")):
            for line in open(f):
                if not line.startswith("TRAFFIC "):
                    continue
                body = line[len("TRAFFIC "):].rsplit(" ", 2)
                if len(body) != 3 or body[0] != nm:
                    continue
                g = int(body[1].split("=")[1])
                if grid is None or g == grid:
                    best = float(body[2])
        return best

    found = lookup(name)
    if found is None and "k_sfft_cols<" in name:
        # the same global accesses as the direct-I/O strided form of the shape
        head, args = name.split("<", 1)
        a = [x.strip() for x in args.rstrip(">").split(",")]
        found = lookup("crava::k_sfft_strided<" + ", ".join(a[:4] + ["true"] + a[4:]) + ">")
    if found is None and ("k_sfft_strided<" in name or "k_srows_" in name):
        head, args = name.split("<", 1)
        a = [x.strip() for x in args.rstrip(">").split(",")]
        i = 4 if "k_sfft_strided<" in name else 3     # the direct-I/O flag
        a[i] = "false" if a[i] == "true" else "true"
        found = lookup(head + "<" + ", ".join(a) + ">")
    if found is None and "k_sfft_strided<" in name:
        # the pencil-fastest form of the same shape (no direct-I/O flag)
        head, args = name.split("<", 1)
        a = [x.strip() for x in args.rstrip(">").split(",")]
        found = lookup("crava::k_sfft_cols<" + ", ".join(a[:4] + a[5:]) + ">")
    return found


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--config", default="precision")
    ap.add_argument("--threshold", type=float, default=1e-3, help="noise/signal below which MIXED uses FP64")
    ap.add_argument("--cpu-seconds", type=float, default=20.0, help="budget of the CPU baseline sample")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--errcorr-grid", action="store_true",
                    help="fill and transform the whole error-correlation grid every step (default: its spectrum "
                         "as one Fourier plane, exact for the workload's zero correlation gradients)")
    ap.add_argument("--kernel-reps", type=int, default=5)
    ap.add_argument("--sweep", action="store_true", help="report accuracy over a range of FP64 thresholds")
    ap.add_argument("--batch-volumes", type=int, default=48, help="volumes of the batching workload (0: skip)")
    ap.add_argument("--batch-reps", type=int, default=5)
    ap.add_argument("--forward-reps", type=int, default=5, help="timed calls of the forward modelling "
                    "workload (0: skip)")
    ap.add_argument("--diff-steps", type=int, default=50, help="L-BFGS iterations of the differentiation workload "
                    "(0: skip it)")
    ap.add_argument("--only", choices=["forward_modelling", "nonconstant_thickness", "time_lapse_merge", "local_noise",
                                       "batching"],
                    help="run one secondary workload alone (for profiling) and print its result")
    return ap.parse_args()


class Pipeline:
    """Working grids + posterior handle for one precision form."""

    def __init__(self, ctx, prob, params, tables, fp64: bool, separable: bool = True, ec_plane: bool = True):
        import numpy as np
        import torch
        from crava_amd import _lib as L
        from crava_amd.core import Posterior
        self.ctx, self.prob, self.fp64 = ctx, prob, fp64
        dt = torch.float64 if fp64 else torch.float32
        self.mu = [g.to(dt).clone() for g in prob.mu]
        self.data = [g.to(dt).clone() for g in prob.data]
        self.cov = [g.to(dt).clone() for g in prob.cov]
        self.cov0 = [g.clone() for g in self.cov]
        # separable prior (initializeCorrelations): the one correlation grid
        self.separable = separable
        self.corr0 = prob.corr.to(dt).clone()
        self.corr = self.corr0.clone()
        self.ec0 = prob.errcorr.to(dt).clone()
        self.ec = self.ec0.clone()
        # zero correlation gradients: fillInErrCorr writes corrXY into plane
        # k = 0 only, so the grid's transform is one Fourier plane repeated
        # along kz (crava_hip_err_corr_plane); the posterior reads that plane
        self.ec_plane = ec_plane
        self.ecp = torch.empty((prob.cfg.nyp, 2 * (prob.cfg.nxp // 2 + 1)), dtype=dt, device=self.ec.device)
        self.post = Posterior(ctx, params, tables)
        # the prior covariance fill (SeismicParametersHolder::initializeCorrelations)
        pairs = [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]
        self.cov_scales = [float(np.float32(prob.sigma0[a, b])) for a, b in pairs]
        self.corr_xy = torch.from_numpy(np.ascontiguousarray(prob.corr_xy, dtype=np.float32)).cuda()
        self.circ_t = torch.from_numpy(np.ascontiguousarray(prob.circ_corr_t, dtype=np.float32)).cuda()
        self.precision = L.F64 if fp64 else L.MIXED
        # transform precision of the data grids, of the means and of the covariances
        self.data_precision = L.F64 if fp64 else L.F32
        self.mean_precision = L.F64 if fp64 else L.F32_DEFLATE
        self.cov_precision = L.F64 if fp64 else L.MIXED
        self.L = L
        self.batch = None
        if ec_plane:
            # valid before the first step: use_graph's autotuner runs the
            # posterior on it
            from crava_amd.core import err_corr_plane
            c = prob.cfg
            err_corr_plane(ctx, self.ecp, c.nxp, c.nyp, c.nzp, self.corr_xy, L.F64 if fp64 else L.MIXED)

    def use_graph(self):
        """Run the step as a one-volume batch replayed as a hipGraph: the
        data/mean and covariance transform sets on two streams (the mixed
        form's double-arithmetic covariance transforms overlap the float data
        transforms), the posterior, the inverse sets."""
        from crava_amd.core import Batch
        self.post.set_errcorr_plane(self.ec_plane)
        vol = (self.post, self.mu, self.data, self.cov, self.ecp if self.ec_plane else self.ec) + \
            ((self.corr,) if self.separable else ())
        self.batch = Batch(self.ctx, self.precision, [vol])
        return self.batch.autotune()

    def step(self, counts=False):
        """One pass of the path; each group of same-kind grids is one launch per transform pass."""
        L, c = self.L, self.prob.cfg
        n = (c.nxp, c.nyp, c.nzp)
        # fill the prior (the six covariance grids, or the one correlation grid
        # of the separable form) and the error correlation afresh
        # (initializeCorrelations, one write per grid): every step then solves
        # the same problem (the FP64 point count depends on them; the previous
        # step's means and residuals as inputs do not change the work).  The
        # FP64 form copies its double grids back.
        prior = [self.corr] if self.separable else self.cov
        self.post.set_errcorr_plane(self.ec_plane)
        if self.ec_plane:
            from crava_amd.core import err_corr_plane
            err_corr_plane(self.ctx, self.ecp, *n, self.corr_xy, L.F64 if self.fp64 else L.MIXED)
        ec = self.ecp if self.ec_plane else self.ec
        # zero gradients, separable prior, one stream: the correlation
        # spectrum straight from the transforms of corrXY and circT
        # (crava_hip_corr_spectrum_separable) instead of the fill and the
        # 3-D transform (the batch transforms its own correlation grid)
        factors = self.ec_plane and self.separable and self.batch is None
        if factors:
            from crava_amd.core import corr_spectrum_separable
            corr_spectrum_separable(self.ctx, self.corr, *n, self.corr_xy, self.circ_t,
                                    L.F64 if self.fp64 else L.MIXED)
        if self.fp64:
            prior0 = [self.corr0] if self.separable else self.cov0
            pairs = list(zip(prior, prior0)) if not factors else []
            if not self.ec_plane:
                pairs.append((self.ec, self.ec0))
            for d, s_ in pairs:
                d.copy_(s_)
        else:
            from crava_amd.core import fill_err_corr, fill_param_corr
            if not factors:
                fill_param_corr(self.ctx, prior, [1.0] if self.separable else self.cov_scales, *n, self.corr_xy,
                                self.circ_t)
            if not self.ec_plane:
                fill_err_corr(self.ctx, self.ec, *n, self.corr_xy)
        if self.batch is not None:
            self.batch.run(graph=True)
            if counts:
                nf, n64 = self.batch.counts()
                return nf[0], n64[0]
            return None
        # data and prior means share the 1/sqrt(N) scaling (DATA / PARAMETER);
        # the mixed form transforms the covariances forward with double
        # arithmetic inside each pass (float storage: the small prior
        # eigenvalues need it), the means in float with mean deflation (the
        # mean update needs the accuracy of their variation, not of their
        # level) and the rest in float
        self.ctx.r2c_many(self.data, *n, L.DATA, self.data_precision)
        if not self.ec_plane:
            self.ctx.r2c(self.ec, *n, L.COVARIANCE, self.data_precision)
        self.ctx.r2c_many(self.mu, *n, L.DATA, self.mean_precision)
        if not factors:
            self.ctx.r2c_many(prior, *n, L.COVARIANCE, self.cov_precision)
        r = self.post.run(self.precision, self.mu, self.data, self.cov, ec, counts=counts,
                          corr=self.corr if self.separable else None)
        self.ctx.c2r_many(self.mu, *n, L.DATA, self.mean_precision)
        self.ctx.c2r_many(self.data, *n, L.DATA, self.data_precision)
        self.ctx.c2r_many(self.cov, *n, L.COVARIANCE, self.data_precision)
        return r

    def reset(self):
        """Reload the inputs, the error correlation included, with the
        library's grid copy (the plain copy the HBM counters are calibrated
        against)."""
        from crava_amd.core import grid_op
        c = self.prob.cfg
        for dst, src in zip(self.mu + self.data + self.cov + [self.ec, self.corr],
                            self.prob.mu + self.prob.data + self.prob.cov + [self.prob.errcorr, self.prob.corr]):
            if self.fp64:
                dst.copy_(src)
            else:
                grid_op(self.ctx, self.L.OP_COPY, dst, src, c.nxp, c.nyp, c.nzp, False)


def rel_err(a, b, nxp, ref_base=None):
    import torch
    a = a[..., :nxp].double()
    b = b[..., :nxp].double()
    den = (b - ref_base[..., :nxp].double()) if ref_base is not None else b
    return float(torch.linalg.vector_norm(a - b) / torch.linalg.vector_norm(den))


def timed_kernel(fn, reps, setup=None):
    """Average duration (ms) of fn() over reps launches, HIP events on torch's
    stream; setup() (e.g. restoring in-place inputs) runs before each launch,
    outside the events."""
    import torch
    if setup:
        setup()
    fn()
    torch.cuda.synchronize()
    ev = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(reps)]
    for s, e in ev:
        if setup:
            setup()
        s.record()
        fn()
        e.record()
    torch.cuda.synchronize()
    return sum(s.elapsed_time(e) for s, e in ev) / reps


def cpu_info():
    """CPU model (from /proc/cpuinfo), logical CPUs and the CPUs this process may use."""
    model = platform.processor() or platform.machine()
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                model = line.split(":", 1)[1].strip()
                break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def cpu_end_to_end_small():
    """The reference CPU path end to end on one core at the CPU-runnable
    configuration (64x64x128 cells, FFT grid 80x80x160, 3 angles): 16
    forward transforms, the per-point loop over all P points, 15 inverse
    transforms, timed as one wall-clock interval."""
    import oracle
    from crava_amd import synth
    from crava_amd.core import problem_tables
    kind = "reference" if oracle.ref_available() else "port"
    which = "ref" if kind == "reference" else "oracle"
    prob = synth.make_problem(synth.config("plain_cpu"), "cpu")
    cfg = prob.cfg
    n = (cfg.nxp, cfg.nyp, cfg.nzp)
    tabs = problem_tables(prob)
    mu = [g.numpy().copy() for g in prob.mu]
    d = [g.numpy().copy() for g in prob.data]
    cv = [g.numpy().copy() for g in prob.cov]
    ec = prob.errcorr.numpy().copy()
    t0 = time.perf_counter()
    for g in d + mu:
        oracle.fft_r2c(g, *n, False, which)
    for g in cv + [ec]:
        oracle.fft_r2c(g, *n, True, which)
    oracle.posterior(*n, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, tabs, mu, d, cv, ec, which)
    for g in mu + d:
        oracle.fft_c2r(g, *n, False, which)
    for g in cv:
        oracle.fft_c2r(g, *n, True, which)
    dt = time.perf_counter() - t0
    P = (cfg.nxp // 2 + 1) * cfg.nyp * cfg.nzp
    return {"workload": f"posterior path {cfg.nxp}x{cfg.nyp}x{cfg.nzp} FFT grid, ntheta={prob.ntheta}",
            "seconds": dt, "value": P / dt, "unit": "Fourier pts/s", "cores": 1, "kind": kind,
            "sample": "the whole path, timed end to end (not composed)"}


def cpu_baseline(prob, tabs, budget_s):
    """Reference CPU path (FFTW 2.1.2 float + lib_matr.c built from the
    reference sources) on one core, on a bounded sample: one full-size 3-D
    r2c, one c2r, and the per-point loop over a few in-band and out-of-band
    k-planes; the step time is composed from the measured per-unit costs
    (labelled "composed").  The same path is also timed end to end at the
    CPU-runnable configuration."""
    import numpy as np
    import oracle
    cfg = prob.cfg
    nxp, nyp, nzp, nt = cfg.nxp, cfg.nyp, cfg.nzp, prob.ntheta
    kind = "reference" if oracle.ref_available() else "port"
    which = "ref" if kind == "reference" else "oracle"
    cn = nxp // 2 + 1
    P = cn * nyp * nzp
    g = prob.data[0].cpu().numpy().copy()
    t0 = time.perf_counter()
    oracle.fft_r2c(g, nxp, nyp, nzp, False, which)
    t_r2c = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.fft_c2r(g, nxp, nyp, nzp, False, which)
    t_c2r = time.perf_counter() - t0
    # loop sample: planes of Fourier-domain inputs (random but representative magnitudes)
    inv = tabs["invert"]
    kin = [k for k in range(nzp) if inv[k]]
    kout = [k for k in range(nzp) if not inv[k]]
    rng = np.random.default_rng(0)

    def plane_inputs(ks):
        m = len(ks)
        shp = (m, nyp, 2 * cn)
        mk = lambda s: (rng.standard_normal(shp) * s).astype(np.float32)
        mu = [mk(1e-2) for _ in range(3)]
        d = [mk(1e-2) for _ in range(nt)]
        base = np.abs(rng.standard_normal(shp)).astype(np.float32) * 50 + 1
        cv = [(base * np.float32(prob.sigma0[a, b])).astype(np.float32)
              for a, b in [(0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)]]
        ec = np.abs(rng.standard_normal(shp)).astype(np.float32) * 10 + 1
        sub = {k: np.ascontiguousarray(v[ks]) for k, v in tabs.items()}
        return mu, d, cv, ec, sub

    per_pt = {}
    remaining = max(2.0, budget_s - 2 * (t_r2c + t_c2r))
    for label, ks in (("in", kin), ("out", kout)):
        if not ks:
            per_pt[label] = 0.0
            continue
        m = max(1, min(len(ks), 4))
        mu, d, cv, ec, sub = plane_inputs(ks[:m])
        t0 = time.perf_counter()
        oracle.posterior(nxp, nyp, m, prob.A, prob.err_theta_cov, cfg.wnc, prob.sigma0, sub, mu, d, cv, ec, which)
        dtm = time.perf_counter() - t0
        per_pt[label] = dtm / (m * cn * nyp)
        remaining -= dtm
    n_in = len(kin) * cn * nyp
    n_out = len(kout) * cn * nyp
    t_step = 16 * t_r2c + 15 * t_c2r + per_pt["in"] * n_in + per_pt["out"] * n_out
    return {
        "value": P / t_step,
        "unit": "Fourier pts/s",
        "cores": 1,
        "kind": kind,
        "sample": (f"composed: 1 full-grid r2c ({t_r2c:.2f}s) + 1 c2r ({t_c2r:.2f}s) of {nxp}x{nyp}x{nzp} and the "
                   f"per-point loop over up to 4 in-band + 4 out-of-band k-planes ({per_pt['in']*1e6:.2f}/"
                   f"{per_pt['out']*1e6:.3f} us per point); step = 16 r2c + 15 c2r + loop over all P points"),
        "composed": True,
        "seconds_per_step": t_step,
        "cpu": cpu_info(),
        "end_to_end_small": cpu_end_to_end_small(),
    }


def batching(ctx, volumes: int, reps: int, seed: int = 3, ec_plane: bool = True):
    """The batching workload: `volumes` inversion volumes of uneven size
    (nx, ny in 32..192, nz in 64..256, ntheta 3..8, own wavelet, ranges and
    S/N), each with its separable prior, MIXED form, autotuned, replayed as
    one hipGraph per batch.  ec_plane: each volume's error-correlation
    spectrum is one Fourier plane (zero correlation gradients), prepared
    once before the timed replays (read only); else its grid is reloaded
    and transformed in every replay."""
    import numpy as np
    import torch
    from crava_amd import _lib as L
    from crava_amd import synth
    from crava_amd.core import Batch, Posterior, err_corr_plane, problem_tables
    t0 = time.time()
    rng = np.random.default_rng(seed)
    vols, src, posts = [], [], []
    P = P_in = 0
    nts = []
    fwd_bytes = inv_bytes = post_bytes = 0
    for i in range(volumes):
        prob = synth.make_problem(synth.random_config(rng, i, seed=seed), "cuda")
        c = prob.cfg
        tabs = problem_tables(prob)
        posts.append(Posterior(ctx, synth.post_params(prob, 1e-3), tabs))
        cn = c.nxp // 2 + 1
        if ec_plane:
            ecp = torch.empty((c.nyp, 2 * cn), dtype=torch.float32, device="cuda")
            cxy = torch.from_numpy(np.ascontiguousarray(prob.corr_xy, dtype=np.float32)).cuda()
            err_corr_plane(ctx, ecp, c.nxp, c.nyp, c.nzp, cxy, L.MIXED)
            posts[-1].set_errcorr_plane(True)
            pristine = prob.mu + prob.data + prob.cov + [prob.corr]
        else:
            pristine = prob.mu + prob.data + prob.cov + [prob.errcorr, prob.corr]
        work = [g.clone() for g in pristine]
        nt = prob.ntheta
        nts.append(nt)
        src.append((pristine, work))
        # separable prior (initializeCorrelations): one correlation transform per volume
        ec = ecp if ec_plane else work[9 + nt]
        vols.append((posts[-1], work[:3], work[3:3 + nt], work[3 + nt:9 + nt], ec, work[-1]))
        pts = cn * c.nyp * c.nzp
        n_in = int(tabs["invert"].sum()) * cn * c.nyp
        P += pts
        P_in += n_in
        # bytes of the three passes of one grid's transform: x reads the real
        # grid and writes the spectrum, y and z read and write the spectrum
        passes = 4 * c.nxp * c.nyp * c.nzp + 8 * pts + 4 * 8 * pts
        fwd_bytes += (nt + (4 if ec_plane else 5)) * passes   # data, (error correlation,) means; correlation
        inv_bytes += (nt + 9) * passes        # residuals, covariances; means
        post_bytes += n_in * 8 * ((13 if ec_plane else 14) + 2 * nt) + (pts - n_in) * 56
        del prob
    torch.cuda.synchronize()
    log(f"batching: {volumes} volumes generated in {time.time()-t0:.1f}s, {P} Fourier points")
    b = Batch(ctx, L.MIXED, vols)
    t0 = time.time()
    ppb, blocks, form, post_ms = b.autotune()
    log(f"batching autotune {time.time()-t0:.1f}s: {ppb} points/block, {blocks} workgroups, "
        f"form={form}, posterior {post_ms:.3f} ms")

    def reload():
        for pristine, work in src:
            for d, s_ in zip(work, pristine):
                d.copy_(s_)

    def timed(graph):
        reload()
        b.run(graph=graph)          # warm-up (and graph capture)
        torch.cuda.synchronize()
        tot = 0.0
        for _ in range(reps):
            reload()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            b.run(graph=graph)
            e1.record()
            torch.cuda.synchronize()
            tot += e0.elapsed_time(e1)
        return tot / reps

    ms_graph = timed(True)
    ms_eager = timed(False)
    nf, n64 = b.counts()
    phases = []
    for _ in range(reps):
        reload()
        phases.append(b.phase_times())
    ph = [float(np.median([p_[i] for p_ in phases])) for i in range(3)]
    b.close()
    for p in posts:
        p.close()
    out = {"volumes": volumes, "fourier_points": P, "in_band_points": P_in,
           "ntheta_histogram": {str(k): nts.count(k) for k in sorted(set(nts))},
           "ms_per_batch_graph": ms_graph, "ms_per_batch_eager": ms_eager,
           "value": P / (ms_graph * 1e-3), "unit": "Fourier pts/s (whole path: transforms + posterior)",
           "posterior_launch": {"points_per_block": ppb, "workgroups": blocks,
                                "form": ["one persistent launch", "one persistent launch per ntheta class",
                                         "one launch per volume"][form],
                                "ms": post_ms, "pts_per_s": P / (post_ms * 1e-3)},
           "failed_factorisations": int(sum(nf)), "fp64_points": int(sum(n64)), "threshold": 1e-3,
           "errcorr": "one Fourier plane per volume (zero correlation gradients), prepared before the timed "
                      "replays" if ec_plane else "grid reloaded outside and transformed inside every replay"}

    def roof(ms, nbytes, what):
        gbs = nbytes / (ms * 1e-3) / 1e9
        return {"what": what, "ms": ms, "bytes": nbytes, "GBps": gbs, "frac": gbs / PEAK_HBM_GBS}

    # one eager pass timed per phase (HIP events on the launch stream);
    # bytes: each transform pass reads and writes its grids, the posterior
    # its per-point inputs and outputs
    out["phase_rooflines"] = {
        "forward_transforms": roof(ph[0], fwd_bytes, f"r2c of (ntheta + {4 if ec_plane else 5}) grids per volume, "
                                                     "3 passes each"),
        "posterior": roof(ph[1], post_bytes, f"separable MIXED posterior: 8 ({13 if ec_plane else 14} + 2 ntheta) B "
                                             "in band, 56 B out"),
        "inverse_transforms": roof(ph[2], inv_bytes, "c2r of (ntheta + 9) grids per volume, 3 passes each"),
    }
    for k, v in out["phase_rooflines"].items():
        log(f"batching phase {k}: {v['ms']:.2f} ms, {v['GBps']:.0f} GB/s ({v['frac']:.2f} of peak)")
    log(f"batching: graph {ms_graph:.2f} ms, eager {ms_eager:.2f} ms per batch, {out['value']:.4g} pts/s")
    return out


def differentiation(ctx, steps: int = 50):
    """The differentiation workload: 256x256x192 cells (FFT grid 288x288x216),
    5 angles.  Seismic data drawn from the likelihood model at known
    parameters; the hand-written adjoint (crava_hip_nll_grad, FP64, Cholesky
    recomputed per point) is timed, checked against FP64 central
    differences (relative h = 1e-6) on every parameter, and drives L-BFGS
    from perturbed wavelet scales and noise variances (Sigma0 held)."""
    import numpy as np
    import torch
    from crava_amd import _lib as L
    from crava_amd import synth
    from crava_amd.autograd import nll_grad
    from crava_amd.core import Posterior, problem_tables
    from crava_amd.estimate import fit
    t0 = time.time()
    prob = synth.make_problem(synth.config("differentiation"), "cuda")
    c = prob.cfg
    n3 = (c.nxp, c.nyp, c.nzp)
    nt = prob.ntheta
    cn = c.nxp // 2 + 1
    tabs = problem_tables(prob)
    mu = [g.double() for g in prob.mu]
    cv = [g.double() for g in prob.cov]
    ec = prob.errcorr.double()
    ctx.r2c_many(mu, *n3, L.DATA)
    ctx.r2c_many(cv + [ec], *n3, L.COVARIANCE)
    torch.cuda.synchronize()
    cplx = lambda t: t.cpu().numpy().view(np.complex128)
    L0 = np.linalg.cholesky(prob.sigma0)
    truth_s = np.ones(nt)
    truth_v = np.diag(prob.err_theta_cov).copy()
    truth_l = np.array([L0[0, 0], L0[1, 0], L0[1, 1], L0[2, 0], L0[2, 1], L0[2, 2]])
    d = synth.model_data(prob, tabs, [cplx(g) for g in mu], cplx(cv[0]), cplx(ec), truth_s, truth_v, truth_l, seed=3)
    data = [torch.from_numpy(np.ascontiguousarray(x).view(np.float64)).cuda() for x in d]
    del d
    post = Posterior(ctx, synth.post_params(prob), tabs)
    problem = (post, L.F64, mu, data, cv, ec)
    n_in = int(tabs["invert"].sum()) * cn * c.nyp
    log(f"differentiation: {c.nxp}x{c.nyp}x{c.nzp}, {n_in} in-band points, set up in {time.time()-t0:.1f}s")
    nll_grad(post, L.F64, mu, data, cv, ec, truth_s, truth_v, truth_l)
    reps = 5
    t1 = time.perf_counter()
    for _ in range(reps):
        f0, g0 = nll_grad(post, L.F64, mu, data, cv, ec, truth_s, truth_v, truth_l)
    ms = (time.perf_counter() - t1) / reps * 1e3
    # FP64 central differences (relative step 1e-6) on every parameter, at
    # the L-BFGS start point.  At the generating parameters the gradient is a
    # sum of zero-mean terms that nearly cancel, so a difference quotient
    # there resolves it poorly; away from them it is well conditioned.
    s0 = truth_s * np.array([1.15, 0.85, 1.1, 0.9, 1.05, 0.95, 1.12, 0.88])[:nt]
    v0 = truth_v * 2.0
    start = [s0, v0, truth_l]
    f_s, g_s = nll_grad(post, L.F64, mu, data, cv, ec, *start)
    names = [f"scale{i}" for i in range(nt)] + [f"noise_var{i}" for i in range(nt)] + \
            ["chol0_L00", "chol0_L10", "chol0_L11", "chol0_L20", "chol0_L21", "chol0_L22"]
    flat = np.concatenate(start)
    fd = {}
    for j, name in enumerate(names):
        h = 1e-6 * abs(flat[j])
        vals = []
        for sgn in (1, -1):
            x = flat.copy()
            x[j] += sgn * h
            vals.append(nll_grad(post, L.F64, mu, data, cv, ec, x[:nt], x[nt:2 * nt], x[2 * nt:])[0])
        num = (vals[0] - vals[1]) / (2 * h)
        # the difference quotient resolves the derivative to about
        # eps |f| / (h |f'|): the likelihood sums millions of points
        fd[name] = {"adjoint": g_s[j], "central_difference": num, "h": h,
                    "rel_err": abs(g_s[j] - num) / max(abs(num), 1e-300),
                    "fd_resolution": 2.2e-16 * abs(f_s) / (h * max(abs(num), 1e-300))}
    # L-BFGS from the same start, Sigma0 held at the truth
    free = [1.0] * (2 * nt) + [0.0] * 6
    t1 = time.perf_counter()
    s, v, _, trace = fit(problem, nt, s0, v0, truth_l, steps=steps, free=free)
    t_fit = time.perf_counter() - t1
    post.close()
    out = {"grid": list(n3), "ntheta": nt, "in_band_points": n_in, "ms_per_nll_and_gradient": ms,
           "pts_per_s": n_in / (ms * 1e-3),
           "gradient_vs_fd": {"point": "L-BFGS start", "max_rel_err": max(x["rel_err"] for x in fd.values()),
                              "per_parameter": fd},
           "lbfgs": {"iterations": steps, "evaluations": len(trace), "seconds": t_fit,
                     "nll_start": trace[0][0], "nll_end": trace[-1][0],
                     "max_rel_err_scales": float(np.abs(s.numpy() / truth_s - 1).max()),
                     "max_rel_err_noise_var": float(np.abs(v.numpy() / truth_v - 1).max()),
                     "start_rel_err_scales": float(np.abs(s0 / truth_s - 1).max()), "start_rel_err_noise_var": 1.0}}
    log(f"differentiation: {ms:.2f} ms per NLL+gradient; FD max rel err {max(x['rel_err'] for x in fd.values()):.3g}; "
        f"L-BFGS {t_fit:.1f}s, scales err {out['lbfgs']['max_rel_err_scales']:.3g}, "
        f"noise err {out['lbfgs']['max_rel_err_noise_var']:.3g}")
    return out


def forward_modelling(ctx, prob, reps: int, cpu: bool):
    """Forward modelling (crava_hip_synthetic_seismic, Crava::computeSyntSeismic)
    of the posterior workload's volume: synthetic seismic for every angle
    from the elastic parameter grids (the prior means here), timed per call
    (host wall clock, tables and uploads included).  Algorithmic bytes: one
    read of the three parameter grids and one write per angle.  The
    reference's own FFTW path (oracle/_ref) is timed on one core on the
    first columns of traces and scaled to the volume."""
    import numpy as np
    import torch
    from crava_amd.core import synthetic_seismic
    c = prob.cfg
    nt = prob.ntheta
    n6 = (c.nx, c.ny, c.nz, c.nxp, c.nyp, c.nzp)
    A = np.asarray(prob.A, dtype=np.float32)
    wav = np.asarray(prob.wavelets, dtype=np.float32)
    g = [m.float().contiguous() for m in prob.mu]
    out = [torch.empty_like(g[0]) for _ in range(nt)]
    synthetic_seismic(ctx, *n6, A, wav, c.dt, *g, out)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        synthetic_seismic(ctx, *n6, A, wav, c.dt, *g, out)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) / reps * 1e3
    rsize = g[0].numel()
    alg = (12 + 4 * nt) * rsize
    traces = c.nx * c.ny
    res = {"volume": [c.nx, c.ny, c.nz], "grid": [c.nxp, c.nyp, c.nzp], "ntheta": nt, "ms_per_call": ms,
           "traces_per_s": traces * nt / (ms * 1e-3), "alg_bytes": alg,
           "alg_GBps": alg / (ms * 1e-3) / 1e9, "frac_of_peak": alg / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS,
           "cpu_reference": None}
    log(f"forward modelling: {ms:.2f} ms per call ({nt} angles), {res['alg_GBps']:.0f} GB/s on the minimal bytes")
    if cpu:
        try:
            import oracle
            if oracle.ref_available():
                h = [m.cpu().numpy() for m in g]
                t1 = time.perf_counter()
                oracle.synthetic_seismic(*n6, A, wav, c.dt, *h, which="ref", i_range=(0, 1))
                t_one = time.perf_counter() - t1
                cols = 8
                t1 = time.perf_counter()
                oracle.synthetic_seismic(*n6, A, wav, c.dt, *h, which="ref", i_range=(0, 1 + cols))
                t_more = time.perf_counter() - t1
                per_col = max(0.0, t_more - t_one) / cols
                total = t_one + per_col * (c.nx - 1)
                res["cpu_reference"] = {"seconds_per_call": total, "cores": 1, "kind": "reference",
                                        "sample": f"impedance of every cell + traces of the first 1 and "
                                                  f"{1 + cols} x-columns ({t_one:.1f}s, {t_more:.1f}s), "
                                                  f"scaled to {c.nx} columns",
                                        "gpu_speedup": total / (ms * 1e-3)}
                log(f"forward modelling, reference CPU path: {total:.1f} s per call (1 core)")
        except Exception as e:  # the baseline is reported, never the target
            log(f"forward modelling cpu baseline failed: {e}")
    del out, g
    return res


def nonconstant_thickness(ctx, reps: int, cpu: bool):
    """Data preparation of the non-constant-thickness path on the posterior
    workload's grid (448x448x256 volume in 512x512x320, 6 angles): the
    wavelet division of every padded trace (crava_hip_divide_by_scale_wavelet,
    Crava::divideDataByScaleWavelet) with relThick U(0.6, 1), local shifts on
    three angles and gains on five, then interpolateSeismic of one angle with
    2000 empty and 500 weak traces (crava_hip_interpolate_seismic).  HIP
    events around each call on the launch stream (table set-up, uploads and
    the schedule's host work included).  Division algorithmic bytes: one
    read and one write of every grid.  The reference's FFTW path
    (oracle/_ref) is timed on one core on the first x-columns and scaled."""
    import numpy as np
    import torch
    from crava_amd.core import divide_by_scale_wavelet, interpolate_seismic
    from crava_amd.synth import isq_grid, swdiv_inputs
    n = (448, 448, 256, 512, 512, 320)
    nt = 6
    args, kw = swdiv_inputs(*n, nt, seed=21)
    src = [torch.from_numpy(g).cuda() for g in args[0]]
    work = [g.clone() for g in src]

    def restore():
        for w, g in zip(work, src):
            w.copy_(g)

    ms = timed_kernel(lambda: divide_by_scale_wavelet(ctx, *n, work, *args[1:], **kw), reps, setup=restore)
    rsize = src[0].numel()
    alg = 8 * rsize * nt
    res = {"volume": list(n[:3]), "grid": list(n[3:]), "ntheta": nt,
           "division": {"ms_per_call": ms, "traces_per_s": n[3] * n[4] * nt / (ms * 1e-3), "alg_bytes": alg,
                        "alg_GBps": alg / (ms * 1e-3) / 1e9, "frac_of_peak": alg / (ms * 1e-3) / 1e9 / PEAK_HBM_GBS},
           "cpu_reference": None}
    log(f"wavelet division: {ms:.2f} ms per call ({nt} angles), {res['division']['alg_GBps']:.0f} GB/s")
    del work, src
    g0 = torch.from_numpy(isq_grid(*n, seed=5, n_bad=2000, n_weak=500)).cuda()
    gi = g0.clone()
    t_isq = timed_kernel(lambda: interpolate_seismic(ctx, *n, 0.1, gi), reps, setup=lambda: gi.copy_(g0))
    res["interpolation"] = {"ms_per_call": t_isq, "traces_flagged": 2500}
    log(f"interpolateSeismic: {t_isq:.2f} ms per call")
    del g0, gi
    if cpu:
        try:
            import oracle
            if oracle.ref_available():
                t1 = time.perf_counter()
                oracle.divide_by_scale_wavelet(*n, *args, **kw, which="ref", i_range=(0, 1))
                t_one = time.perf_counter() - t1
                cols = 8
                t1 = time.perf_counter()
                oracle.divide_by_scale_wavelet(*n, *args, **kw, which="ref", i_range=(0, 1 + cols))
                t_more = time.perf_counter() - t1
                per_col = max(0.0, t_more - t_one) / cols
                total = t_one + per_col * (n[3] - 1)
                res["cpu_reference"] = {"seconds_per_call": total, "cores": 1, "kind": "reference",
                                        "sample": f"traces of the first 1 and {1 + cols} x-columns ({t_one:.1f}s, "
                                                  f"{t_more:.1f}s, grid copies included), scaled to {n[3]} columns",
                                        "gpu_speedup": total / (ms * 1e-3)}
                log(f"wavelet division, reference CPU path: {total:.1f} s per call (1 core)")
        except Exception as e:  # the baseline is reported, never the target
            log(f"wavelet division cpu baseline failed: {e}")
    torch.cuda.empty_cache()
    return res


def local_noise(ctx, reps: int, cpu: bool):
    """Local-noise correction on the posterior workload's grid (448x448x256
    volume in 512x512x320, 6 angles, noise scales U(1, 3) per trace and
    angle): crava_hip_local_noise_correct (Crava::correctAlphaBetaRho: the
    per-trace matrices of newPosteriorCovPointwise and the correction of the
    three posterior grids, scale-table upload included) and
    crava_hip_local_noise_realisation (the local-noise block of
    Crava::simulate).  HIP events on the launch stream.  Algorithmic bytes of
    the correction's grid pass: 6 float reads and 3 float writes per volume
    point; of the realisation: 3 reads and 3 writes.  The oracle's numpy
    restatement (a port: Crava does not build here) is timed on one core on
    the first traces and scaled."""
    import torch
    from crava_amd import synth
    from crava_amd.core import local_noise_correct, local_noise_g, local_noise_realisation
    n = (448, 448, 256, 512, 512, 320)
    nx, ny, nz, nxp, nyp, nzp = n
    nt = 6
    prob = synth.local_noise_problem(nt, nx, ny, seed=3)
    G = local_noise_g(prob["E"], prob["Sm"], prob["Smd"])
    shape = (nzp, nyp, 2 * (nxp // 2 + 1))
    gen = torch.Generator(device="cuda").manual_seed(9)
    mean = [8.0 + 0.05 * torch.randn(shape, device="cuda", generator=gen) for _ in range(3)]
    src = [m + 0.03 * torch.randn(shape, device="cuda", generator=gen) for m in mean]
    post = [g.clone() for g in src]
    tc = torch.empty(nx * ny * 9, dtype=torch.float64, device="cuda")

    def restore():
        for w, g in zip(post, src):
            w.copy_(g)

    ms = timed_kernel(lambda: local_noise_correct(ctx, *n, G, prob["E"], prob["Sm"], prob["Smd"], prob["scales"],
                                                  post, mean, tc), reps, setup=restore)
    pts = nx * ny * nz
    alg = 36 * pts
    ms_r = timed_kernel(lambda: local_noise_realisation(ctx, *n, tc, post), reps)
    alg_r = 24 * pts
    res = {"volume": [nx, ny, nz], "grid": [nxp, nyp, nzp], "ntheta": nt,
           "correction": {"ms_per_call": ms, "traces_per_s": nx * ny / (ms * 1e-3), "grid_alg_bytes": alg,
                          "grid_alg_GBps_over_call": alg / (ms * 1e-3) / 1e9},
           "realisation": {"ms_per_call": ms_r, "alg_bytes": alg_r, "alg_GBps": alg_r / (ms_r * 1e-3) / 1e9,
                           "frac_of_peak": alg_r / (ms_r * 1e-3) / 1e9 / PEAK_HBM_GBS},
           "cpu_port": None}
    log(f"local-noise correction: {ms:.2f} ms per call; realisation transform {ms_r:.2f} ms "
        f"({res['realisation']['alg_GBps']:.0f} GB/s)")
    del post, src, mean
    if cpu:
        try:
            import numpy as np
            from oracle import local_noise as ln
            ntr = 64
            sub = [t[:1, :ntr].copy() for t in prob["scales"]]
            sub = [np.concatenate([s, np.full((1, 1), t.min())], axis=1) for s, t in zip(sub, prob["scales"])]
            t1 = time.perf_counter()
            sn, tcs = ln.trace_matrices(G, sub, prob["E"], prob["Sm"], prob["Smd"])
            t_tr = (time.perf_counter() - t1) / (ntr + 1)
            pg = [np.full((nz, 1, ntr + 2), 8.0, np.float32) for _ in range(3)]
            mg = [np.full((nz, 1, ntr + 2), 7.9, np.float32) for _ in range(3)]
            t1 = time.perf_counter()
            ln.apply_correction(pg, mg, sn, ntr + 1, 1, nz)
            t_ap = (time.perf_counter() - t1) / (ntr + 1)
            total = (t_tr + t_ap) * nx * ny
            res["cpu_port"] = {"seconds_per_call": total, "cores": 1, "kind": "port",
                               "sample": f"{ntr + 1} traces of one row (matrices {t_tr * 1e3:.2f} ms, correction "
                                         f"{t_ap * 1e3:.2f} ms per trace), scaled to {nx * ny} traces",
                               "gpu_speedup": total / (ms * 1e-3)}
            log(f"local-noise correction, oracle port on 1 core: {total:.1f} s per call")
        except Exception as e:  # the baseline is reported, never the target
            log(f"local-noise cpu port failed: {e}")
    torch.cuda.empty_cache()
    return res


def time_lapse_merge(ctx, reps: int, cpu: bool):
    """Time-lapse merge (crava_hip_merge_cov_4d / _mean_4d, State4D::mergeCov /
    merge) on 512x512x320 transformed grids: 6 + 6 + 9 covariance parts into
    6 grids, 3 + 3 means into 3.  HIP events on the launch stream; the
    oracle's restatement (a port: the reference's State4D is C++ that does
    not build here) timed on one core on the first k-planes."""
    import numpy as np
    import torch
    from crava_amd.core import merge_cov_4d, merge_mean_4d
    nxp, nyp, nzp = 512, 512, 320
    cn = nxp // 2 + 1
    shape = (nzp, nyp, 2 * cn)
    gen = torch.Generator(device="cuda").manual_seed(5)
    parts = [torch.randn(shape, device="cuda", generator=gen) for _ in range(21)]
    out = [torch.empty(shape, device="cuda") for _ in range(6)]
    ms = [torch.randn(shape, device="cuda", generator=gen) for _ in range(6)]
    mo = [torch.empty(shape, device="cuda") for _ in range(3)]
    t_cov = timed_kernel(lambda: merge_cov_4d(ctx, nxp, nyp, nzp, parts[:6], parts[6:12], parts[12:], out), reps)
    t_mu = timed_kernel(lambda: merge_mean_4d(ctx, nxp, nyp, nzp, ms[:3], ms[3:], mo), reps)
    gb = 8 * cn * nyp * nzp
    res = {"grid": [nxp, nyp, nzp],
           "cov": {"ms": t_cov, "alg_bytes": 27 * gb, "GBps": 27 * gb / (t_cov * 1e-3) / 1e9},
           "mean": {"ms": t_mu, "alg_bytes": 9 * gb, "GBps": 9 * gb / (t_mu * 1e-3) / 1e9},
           "cpu_port": None}
    res["cov"]["frac_of_peak"] = res["cov"]["GBps"] / PEAK_HBM_GBS
    res["mean"]["frac_of_peak"] = res["mean"]["GBps"] / PEAK_HBM_GBS
    log(f"time-lapse merge: covariance {t_cov:.3f} ms ({res['cov']['GBps']:.0f} GB/s), "
        f"means {t_mu:.3f} ms ({res['mean']['GBps']:.0f} GB/s)")
    if cpu:
        try:
            import oracle
            kk = 8
            h = [p[:kk].cpu().numpy().view(np.complex64) for p in parts]
            t1 = time.perf_counter()
            oracle.merge_cov_4d(h[:6], h[6:12], h[12:])
            dt = time.perf_counter() - t1
            res["cpu_port"] = {"seconds_per_call": dt * nzp / kk, "cores": 1, "kind": "port",
                               "sample": f"the oracle's mergeCov over the first {kk} of {nzp} k-planes "
                                         f"({dt:.2f} s), scaled"}
        except Exception as e:
            log(f"time-lapse merge cpu baseline failed: {e}")
    del parts, out, ms, mo
    torch.cuda.empty_cache()
    return res


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU only; --gpus must be 1"}))
        sys.exit(2)
    import numpy as np
    import torch
    from crava_amd import _lib as L
    from crava_amd import synth
    from crava_amd.core import Context, problem_tables

    torch.cuda.init()
    dev = torch.cuda.current_device()
    if args.only:
        ctx = Context()
        reps = max(args.forward_reps, 1)
        fn = {"forward_modelling": lambda: forward_modelling(ctx, synth.make_problem(synth.config(args.config), "cuda"),
                                                             reps, not args.no_cpu),
              "nonconstant_thickness": lambda: nonconstant_thickness(ctx, reps, not args.no_cpu),
              "batching": lambda: batching(ctx, args.batch_volumes, args.batch_reps,
                                           ec_plane=not args.errcorr_grid),
              "time_lapse_merge": lambda: time_lapse_merge(ctx, reps, not args.no_cpu),
              "local_noise": lambda: local_noise(ctx, reps, not args.no_cpu)}[args.only]
        print(json.dumps({"workload": args.only, "result": fn()}))
        return
    cfg = synth.config(args.config)
    log(f"config {cfg.name}: grid {cfg.nxp}x{cfg.nyp}x{cfg.nzp}, ntheta={len(cfg.thetas_deg)}")
    t0 = time.time()
    prob = synth.make_problem(cfg, "cuda")
    torch.cuda.synchronize()
    log(f"generated inputs on the GPU in {time.time()-t0:.1f}s")
    tabs = problem_tables(prob)
    params = synth.post_params(prob, args.threshold)
    ctx = Context()
    nxp, nyp, nzp = cfg.nxp, cfg.nyp, cfg.nzp
    cn = nxp // 2 + 1
    P = cn * nyp * nzp
    n_in = int(tabs["invert"].sum()) * cn * nyp
    nt = prob.ntheta

    mixed = Pipeline(ctx, prob, params, tabs, fp64=False, ec_plane=not args.errcorr_grid)
    # autotune the transform launch shapes (float and double) on a working grid
    from crava_amd.core import autotune_fft
    tune = {"f32": autotune_fft(ctx, mixed.data[0], nxp, nyp, nzp),
            "f32_store_f64_math": autotune_fft(ctx, mixed.data[0], nxp, nyp, nzp, L.MIXED)}
    log(f"autotuned variants {tune}")
    mixed.reset()
    f64 = None
    if not args.no_fp64:
        f64 = Pipeline(ctx, prob, params, tabs, fp64=True, ec_plane=not args.errcorr_grid)
        tune["f64"] = autotune_fft(ctx, f64.data[0], nxp, nyp, nzp)
        log(f"autotuned f64 variants {tune['f64']}")
        f64.reset()

    def errors(m, ref):
        return {"mean_update": max(rel_err(a, b, nxp, p) for a, b, p in zip(m.mu, ref.mu, prob.mu)),
                "mean": max(rel_err(a, b, nxp) for a, b in zip(m.mu, ref.mu)),
                "cov": max(rel_err(a, b, nxp) for a, b in zip(m.cov, ref.cov)),
                "residual": max(rel_err(a, b, nxp) for a, b in zip(m.data, ref.data))}

    # accuracy: one step of each form from the same inputs
    acc, sweep, graph_tune = {}, [], {}
    if f64 is not None:
        graph_tune["f64"] = f64.use_graph()
        f64.reset()
        f64.step(counts=True)
        torch.cuda.synchronize()
        if args.sweep:
            from crava_amd.core import Posterior
            for thr in (1e-3, 1e-2, 3e-2, 1e-1, 3e-1, 1.0, 1e30):
                mixed.post.close()
                mixed.post = Posterior(ctx, synth.post_params(prob, thr), tabs)
                mixed.reset()
                nf_s, n64_s = mixed.step(counts=True)
                torch.cuda.synchronize()
                e = errors(mixed, f64)
                t_k = None
                sweep.append({"threshold": thr, "fp64_points": n64_s, "errors": e})
                log(f"threshold {thr:g}: {n64_s} FP64 points, errors {e}")
            mixed.post.close()
            mixed.post = Posterior(ctx, params, tabs)
        mixed.reset()
    graph_tune["mixed"] = mixed.use_graph()
    log(f"one-volume graph launch choices (points/block, workgroups, form, posterior ms): {graph_tune}")
    mixed.reset()
    nfail, n64 = mixed.step(counts=True)
    torch.cuda.synchronize()
    log(f"mixed step: {nfail} failed factorisations, {n64} of {n_in} in-band points in FP64")
    if f64 is not None:
        acc = errors(mixed, f64)
        log(f"rel err mixed vs FP64: {acc}")
        f64.reset()
    mixed.reset()

    def run(pipe, steps, warmup):
        for _ in range(warmup):
            pipe.step()
        torch.cuda.synchronize()
        t = time.perf_counter()
        for _ in range(steps):
            pipe.step()
        torch.cuda.synchronize()
        return (time.perf_counter() - t) / steps

    # the step as a one-volume batch replayed as a hipGraph (two transform
    # streams), then as separate launches on one stream: `value` is the faster
    ms_mixed_graph = run(mixed, args.steps, args.warmup) * 1e3
    log(f"mixed, one-volume graph: {ms_mixed_graph:.3f} ms/step")
    batch_keep, mixed.batch = mixed.batch, None
    ms_mixed_direct = run(mixed, args.steps, args.warmup) * 1e3
    log(f"mixed, one stream: {ms_mixed_direct:.3f} ms/step")
    if ms_mixed_graph < ms_mixed_direct:
        mixed.batch = batch_keep
    ms_mixed = min(ms_mixed_graph, ms_mixed_direct)
    log(f"mixed: {ms_mixed:.3f} ms/step, {P/ms_mixed*1e3:.4g} pts/s")
    ms_f64 = None
    if f64 is not None:
        ms_f64_graph = run(f64, max(1, args.steps), 1) * 1e3
        f64.batch = None
        ms_f64_direct = run(f64, max(1, args.steps), 1) * 1e3
        ms_f64 = min(ms_f64_graph, ms_f64_direct)
        log(f"fp64: {ms_f64:.3f} ms/step (graph {ms_f64_graph:.3f}, one stream {ms_f64_direct:.3f}), "
            f"{P/ms_f64*1e3:.4g} pts/s")

    # the general prior (six covariance grids transformed and read, as for a
    # time-lapse merged prior), one stream, both forms: time and accuracy
    general = None
    if f64 is not None:
        keep_m, keep_f = mixed.batch, f64.batch
        mixed.batch = f64.batch = None
        mixed.separable = f64.separable = False
        f64.reset()
        f64.step()
        mixed.reset()
        mixed.step()
        torch.cuda.synchronize()
        g_err = errors(mixed, f64)
        mixed.reset()
        f64.reset()
        g_mixed = run(mixed, args.steps, args.warmup) * 1e3
        g_f64 = run(f64, max(1, args.steps), 1) * 1e3
        general = {"ms_per_step_mixed": g_mixed, "ms_per_step_fp64": g_f64, "speedup_vs_fp64": g_f64 / g_mixed,
                   "value": P / (g_mixed * 1e-3), "rel_err_vs_fp64": g_err,
                   "transforms_per_step": 31 if args.errcorr_grid else 30}
        log(f"general prior: mixed {g_mixed:.3f} ms, fp64 {g_f64:.3f} ms ({g_f64 / g_mixed:.2f}x), errors {g_err}")
        mixed.separable = f64.separable = True
        mixed.batch, f64.batch = keep_m, keep_f
        mixed.reset()
        f64.reset()

    # per-kernel timing (HIP events on the launch stream), on the launches the
    # step makes: each forward pass over the set of nt data + 3 mean grids
    gset = mixed.data + mixed.mu
    ng = len(gset)
    n3 = (nxp, nyp, nzp)
    t_x = timed_kernel(lambda: ctx.fft_pass_many(gset, *n3, 0, False, 1e-3), args.kernel_reps)
    t_y = timed_kernel(lambda: ctx.fft_pass_many(gset, *n3, 1, False), args.kernel_reps)
    t_z = timed_kernel(lambda: ctx.fft_pass_many(gset, *n3, 2, False), args.kernel_reps)
    grid_c_bytes = 8 * cn * nyp * nzp * ng
    real_bytes = 4 * nxp * nyp * nzp * ng
    mixed.reset()
    # the posterior kernels below read the transformed error-correlation grid
    mixed.post.set_errcorr_plane(False)
    if f64 is not None:
        f64.post.set_errcorr_plane(False)
    ctx.r2c_many(mixed.data, *n3, L.DATA)
    ctx.r2c(mixed.ec, *n3, L.COVARIANCE)
    ctx.r2c_many(mixed.mu, *n3, L.DATA, L.F32_DEFLATE)
    ctx.r2c_many(mixed.cov, *n3, L.COVARIANCE, L.MIXED)

    def restorer(pipe):
        """The posterior works in place: every timed launch starts from the
        same Fourier-domain prior and data (a launch on its own outputs
        would see the posterior covariance and take fewer FP64 points)."""
        grids = pipe.mu + pipe.data + pipe.cov
        saved = [g.clone() for g in grids]

        def setup():
            for d, s_ in zip(grids, saved):
                d.copy_(s_)
        return setup

    setup32 = restorer(mixed)
    t_post = timed_kernel(lambda: mixed.post.run(L.MIXED, mixed.mu, mixed.data, mixed.cov, mixed.ec),
                          args.kernel_reps, setup32)
    # the separable form's posterior (its transformed correlation grid)
    ctx.r2c(mixed.corr, *n3, L.COVARIANCE, L.MIXED)
    t_post_sep = timed_kernel(lambda: mixed.post.run(L.MIXED, mixed.mu, mixed.data, mixed.cov, mixed.ec,
                                                     corr=mixed.corr), args.kernel_reps, setup32)
    t_post32 = timed_kernel(lambda: mixed.post.run(L.F32, mixed.mu, mixed.data, mixed.cov, mixed.ec),
                            args.kernel_reps, setup32)
    del setup32
    t_post64 = None
    if f64 is not None:
        ctx.r2c_many(f64.data + f64.mu, *n3, L.DATA)
        ctx.r2c_many(f64.cov + [f64.ec], *n3, L.COVARIANCE)
        setup64 = restorer(f64)
        t_post64 = timed_kernel(lambda: f64.post.run(L.F64, f64.mu, f64.data, f64.cov, f64.ec), args.kernel_reps,
                                setup64)
        del setup64
    post_bytes = n_in * 8 * (19 + 2 * nt) + (P - n_in) * 96
    # separable: reads mu 3, d nt, corr 1, errCorr 1; writes mu 3, Sigma 6, r nt (in band);
    # reads corr, writes Sigma 6 (out of band)
    sep_bytes = n_in * 8 * (14 + 2 * nt) + (P - n_in) * 56
    kernels = {
        "fft_x_r2c": {"ms": t_x, "alg_bytes": real_bytes + grid_c_bytes},
        "fft_y": {"ms": t_y, "alg_bytes": 2 * grid_c_bytes},
        "fft_z": {"ms": t_z, "alg_bytes": 2 * grid_c_bytes},
        "posterior_mixed": {"ms": t_post, "alg_bytes": post_bytes},
        "posterior_mixed_separable": {"ms": t_post_sep, "alg_bytes": sep_bytes},
        "posterior_f32": {"ms": t_post32, "alg_bytes": post_bytes},
    }
    if t_post64:
        kernels["posterior_f64"] = {"ms": t_post64, "alg_bytes": 2 * post_bytes}
    for k in kernels.values():
        k["GBps"] = k["alg_bytes"] / (k["ms"] * 1e-3) / 1e9
    # dominant kernel of the step by total time: 31 transforms x 3 passes vs one posterior
    fft_total = 31 / ng * (t_x + t_y + t_z)
    if fft_total >= t_post:
        dom_name = max(("fft_x_r2c", "fft_y", "fft_z"), key=lambda k: kernels[k]["ms"])
    else:
        dom_name = "posterior_mixed"
    dom = kernels[dom_name]
    kname, kgrid = kernel_identity(dom_name, tune["f32"][0], nxp, nyp, nzp, ng, nt)
    traffic = profiled_traffic(kname, kgrid)
    roofline = {"bound": "hbm", "kernel": dom_name, "kernel_name": kname, "achieved": round(dom["GBps"], 1),
                "peak": PEAK_HBM_GBS, "unit": "GB/s", "frac": round(dom["GBps"] / PEAK_HBM_GBS, 4),
                "alg_bytes_per_launch": dom["alg_bytes"], "traffic": traffic}
    batch = None
    if args.batch_volumes > 0:
        del mixed, f64
        torch.cuda.empty_cache()
        batch = batching(ctx, args.batch_volumes, args.batch_reps, ec_plane=not args.errcorr_grid)
    diff = None
    if args.diff_steps > 0:
        torch.cuda.empty_cache()
        diff = differentiation(ctx, args.diff_steps)
    tl = None
    if args.forward_reps > 0:
        torch.cuda.empty_cache()
        tl = time_lapse_merge(ctx, args.kernel_reps, not args.no_cpu)
    fwd = None
    if args.forward_reps > 0:
        torch.cuda.empty_cache()
        fwd = forward_modelling(ctx, prob, args.forward_reps, not args.no_cpu)
    nct = None
    if args.forward_reps > 0:
        torch.cuda.empty_cache()
        nct = nonconstant_thickness(ctx, args.forward_reps, not args.no_cpu)
    lnc = None
    if args.forward_reps > 0:
        torch.cuda.empty_cache()
        lnc = local_noise(ctx, args.forward_reps, not args.no_cpu)
    cpu = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(prob, tabs, args.cpu_seconds)
        except Exception as e:  # the baseline is reported, never the target
            log(f"cpu baseline failed: {e}")
    value = P / (ms_mixed * 1e-3)
    out = {
        "metric": "posterior Fourier pts/s (512x512x320, 6 angles): mixed vs FP64, rel err<=1e-5",
        "value": value,
        "unit": "Fourier pts/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_mixed,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64 (mixed)",
        "data": "synthetic (seeded, generated on the GPU)",
        "config": {"workload": f"posterior path {nxp}x{nyp}x{nzp} FFT grid, ntheta={nt}", "fourier_points": P,
                   "in_band_points": n_in, "transforms_per_step": 26 if args.errcorr_grid else (25 if ms_mixed_graph < ms_mixed_direct else 24),
                   "correlation_spectrum": "filled and transformed (one-volume graph)" if args.errcorr_grid or
                   ms_mixed_graph < ms_mixed_direct else "FFT2(corrXY) x FFT1(circT), no fill (one stream)",
                   "errcorr": "spectrum as one Fourier plane (zero correlation gradients: the fillInErrCorr grid "
                              "is corrXY in plane 0, its transform that plane's 2-D transform on every kz plane)"
                              if not args.errcorr_grid else "grid filled and transformed every step",
                   "prior": "separable: Sigma0 x one correlation grid, as initializeCorrelations builds it "
                            "(one correlation transform instead of six; the six-grid form is under general_prior)",
                   "parallelism": "single GPU"},
        "general_prior": general,
        "ms_per_step_graph": ms_mixed_graph,
        "ms_per_step_one_stream": ms_mixed_direct,
        "graph_launch": graph_tune,
        "fp64_value": (P / (ms_f64 * 1e-3)) if ms_f64 else None,
        "fp64_ms_per_step": ms_f64,
        "speedup_vs_fp64": (ms_f64 / ms_mixed) if ms_f64 else None,
        "rel_err_vs_fp64": acc,
        "threshold": args.threshold,
        "threshold_sweep": sweep or None,
        "fft_variants": tune,
        "fp64_points": n64,
        "posterior_kernel_pts_per_s": {"mixed": P / (t_post * 1e-3), "f32": P / (t_post32 * 1e-3),
                                       "f64": (P / (t_post64 * 1e-3)) if t_post64 else None},
        "kernels": {k: {kk: round(vv, 4) for kk, vv in v.items()} for k, v in kernels.items()},
        # three passes: x reads the real grid and writes the half spectrum, y and z read and write it
        "fft_3d_hbm_frac": (real_bytes + 5 * grid_c_bytes) / ((t_x + t_y + t_z) * 1e-3) / 1e9 / PEAK_HBM_GBS,
        # the same time against one read of the real grid + one write of the spectrum per transform
        "fft_3d_single_pass_equiv_GBps": (real_bytes + grid_c_bytes) / ((t_x + t_y + t_z) * 1e-3) / 1e9,
        "fft_set_grids": ng,
        "roofline": roofline,
        "cpu_baseline": cpu,
        "batching": batch,
        "differentiation": diff,
        "forward_modelling": fwd,
        "nonconstant_thickness": nct,
        "local_noise": lnc,
        "time_lapse_merge": tl,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
