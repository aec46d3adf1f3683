/*
 * crava_hip.h -- C ABI of the MI355X-native CRAVA Fourier-domain inversion core.
 *
 * Plain pointers and sizes only.  Device pointers are HIP device memory; the
 * stream is a hipStream_t passed as void*.  Every entry point returns
 * CRAVA_HIP_OK (0) or a negative error code and never exits the process;
 * crava_hip_last_error() gives the message of the calling thread's last error.
 *
 * Grid layout is the reference's FFTGrid layout (src/fftgrid.cpp:51-66):
 *   real    [nzp][nyp][rnxp] float, rnxp = 2*(nxp/2+1)
 *   complex [nzp][nyp][cnxp] {re,im}, cnxp = nxp/2+1, aliasing the same memory.
 * FP64 grids use the same layout with double.
 */
#ifndef CRAVA_HIP_H
#define CRAVA_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes (the reference's int-status convention, lib_matr.c / FFTGrid) */
#define CRAVA_HIP_OK 0
#define CRAVA_HIP_E_INVALID (-1)
#define CRAVA_HIP_E_OOM (-2)
#define CRAVA_HIP_E_HIP (-3)
#define CRAVA_HIP_E_STATE (-4)

/* FFTGrid::cubetype_ (src/fftgrid.h:167) decides the transform scaling */
#define CRAVA_HIP_DATA 0
#define CRAVA_HIP_PARAMETER 1
#define CRAVA_HIP_COVARIANCE 2

/* arithmetic of a posterior / transform call */
#define CRAVA_HIP_F32 0   /* float grids, float solve (the reference's arithmetic) */
#define CRAVA_HIP_MIXED 1 /* float grids; transforms with double arithmetic inside each pass;
                             FP32 square-root solve, FP64 below the noise/signal threshold */
#define CRAVA_HIP_F64 2   /* double grids, double transforms and solve */
/* transforms only: float grids and float arithmetic on each grid minus its
   first value (a value near its level); the level returns exactly through
   the DC bin (forward) or onto every value (inverse), with the scale
   factors in double.  For grids whose level dominates their variation (the
   prior means ln Vp, ln Vs, ln rho): the transform keeps the accuracy of the
   variation instead of the level's. */
#define CRAVA_HIP_F32_DEFLATE 3

typedef struct crava_hip_ctx crava_hip_ctx;
typedef struct crava_hip_post crava_hip_post;
typedef struct crava_hip_batch crava_hip_batch;

/* Library version (major*10000 + minor*100 + patch).  No reference counterpart. */
int crava_hip_version(void);

/* Message of the calling thread's last error ("" if none). */
const char *crava_hip_last_error(void);

/* Execution context: a stream, a cache of transform plans and small device
 * counters.  Replaces the implicit global state of FFTGrid/Crava
 * (src/fftgrid.cpp:2869-2876).  stream = NULL uses the null stream. */
int crava_hip_ctx_create(crava_hip_ctx **ctx, void *stream);
int crava_hip_ctx_destroy(crava_hip_ctx *ctx);
int crava_hip_ctx_set_stream(crava_hip_ctx *ctx, void *stream);
int crava_hip_synchronize(crava_hip_ctx *ctx);

/* Device storage of one grid in the FFTGrid layout (rsize = 2*(nxp/2+1)*nyp*nzp
 * floats, or doubles for CRAVA_HIP_F64), in place of FFTGrid::createRealGrid /
 * createComplexGrid (src/fftgrid.cpp:1283-1316, fftw_malloc).  Upload and
 * download are synchronous copies on the context's stream; with them a
 * host caller needs no HIP type (the FFTGrid cursor access of
 * src/fftgrid.cpp:1419-1653 runs on a host copy between them). */
int crava_hip_grid_alloc(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, int precision, void **grid);
int crava_hip_grid_free(void *grid);
int crava_hip_grid_upload(crava_hip_ctx *ctx, void *grid, const void *host, size_t bytes);
int crava_hip_grid_download(crava_hip_ctx *ctx, void *host, const void *grid, size_t bytes);
/* One trace of a real float grid, grid[k][j][i] for k0 <= k < k0 + nk
 * (i < 2(nxp/2+1)), copied to host memory (synchronous).  Serves the
 * reference's single-trace reads through FFTGrid::getRealValue
 * (SeismicParametersHolder::getOrigin, createPostCov00 and
 * extractParamCorrFromCovAlpha, src/seismicparametersholder.cpp:748-806)
 * without copying the whole grid. */
int crava_hip_grid_read_trace(crava_hip_ctx *ctx, const void *grid, int nxp, int nyp, int nzp, int i, int j, int k0,
                              int nk, float *out);

/* In-place 3-D real-to-half-complex transform of one device grid.
 * Replaces FFTGrid::fftInPlace (src/fftgrid.cpp:1804-1829): grids that are
 * not CRAVA_HIP_COVARIANCE are scaled by 1/sqrt(nxp*nyp*nzp); FFTW's forward
 * sign (-1).  precision: CRAVA_HIP_F32 (float grid, float arithmetic),
 * CRAVA_HIP_MIXED (float grid, double arithmetic inside each of the three
 * passes, whose outputs are rounded once to float; scale factors in double)
 * or CRAVA_HIP_F64 (double grid). */
int crava_hip_fft_r2c(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, int cubetype, int precision);

/* In-place 3-D half-complex-to-real transform.  Replaces
 * FFTGrid::invFFTInPlace (src/fftgrid.cpp:1832-1862): scaled by 1/N for
 * CRAVA_HIP_COVARIANCE, 1/sqrt(N) otherwise; Im of the self-conjugate x bins
 * is ignored as in FFTW's halfcomplex c2r.  precision as crava_hip_fft_r2c,
 * or CRAVA_HIP_F32_DEFLATE (both directions). */
int crava_hip_fft_c2r(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, int cubetype, int precision);

/* The same transforms over `ngrid` (1..16) device grids of one size and
 * cubetype, one kernel launch per pass for the whole set.  Replaces the
 * grid-by-grid loops of SeismicParametersHolder::invFFTAllGrids /
 * FFTAllGrids (src/seismicparametersholder.cpp:262-298, with the
 * covariance grids of FFTCovGrids / invFFTCovGrids) and the per-grid calls
 * of Crava (src/crava.cpp:177,188-190,1012,1181-1183,1201,1538-1545). */
int crava_hip_fft_r2c_many(crava_hip_ctx *ctx, void *const *grids, int ngrid, int nxp, int nyp, int nzp,
                           int cubetype, int precision);
int crava_hip_fft_c2r_many(crava_hip_ctx *ctx, void *const *grids, int ngrid, int nxp, int nyp, int nzp,
                           int cubetype, int precision);

/* One pass of a 3-D transform (axis 0 = x rows incl. the r2c/c2r split,
 * 1 = y, 2 = z), for profiling and per-pass timing; `scale` multiplies the
 * x-pass output.  No reference counterpart (FFTW does all passes inside
 * rfftwnd_one_*). */
int crava_hip_fft_pass(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, int axis, int inverse,
                       int precision, double scale);
/* The same pass over a set of 1..16 grids in one launch. */
int crava_hip_fft_pass_many(crava_hip_ctx *ctx, void *const *grids, int ngrid, int nxp, int nyp, int nzp, int axis,
                            int inverse, int precision, double scale);

/* Autotuner: time every launch variant of each pass of the nxp x nyp x nzp
 * transform (forward and inverse pass together) on `scratch` (a grid of that
 * size; its contents are destroyed) and keep the fastest in the context's
 * plan; ms[] receives the forward + inverse time per axis.  Variants: v >= 0 a
 * compile-time specialised kernel (lengths 64, 80, 128, 160, 216, 256, 288,
 * 320, 512): 0..11 for the strided y/z passes (4..7 loading the first radix
 * pass straight from HBM and storing the last straight to it, 8..9 the same
 * in persistent workgroups, 10..11 the same on pencil-fastest passes), 0..5
 * for the x (row) pass (3..5 with direct I/O); v = -1..-6 the runtime-length kernel
 * with radix order (large-first / small-first) x pencils per workgroup
 * (<=64, <=16, <=8), and v = -7..-12 the same with the first radix pass
 * loading straight from global memory.
 * variants[3] / ms[3] (host, may be NULL) receive the choice per axis and its
 * time.  No reference counterpart (FFTW_ESTIMATE plans are not measured). */
int crava_hip_autotune_fft(crava_hip_ctx *ctx, void *scratch, int nxp, int nyp, int nzp, int precision,
                           int *variants, float *ms);

/* Set the launch variant per axis directly (e.g. from a stored tuning). */
int crava_hip_fft_set_variants(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, int precision, const int *variants);

/* Per-k vectors of the constant-thickness posterior loop, host side.
 * Replaces the per-k preparation of Crava::computePostMeanResidAndFFTCov
 * (src/crava.cpp:938-1057 with fillkW/fillkWNorm :1337-1359,
 * Wavelet::getCAmp src/wavelet.cpp:309-323, Wavelet1D difference operators
 * src/wavelet1D.cpp:368-479).
 *   wavelets [ntheta][nzp] scaled wavelets in FFT order; wnorm[ntheta] their
 *   Wavelet::norm_.  Outputs: invert[nzp]; kWD, e1, e2, e3 [nzp][ntheta]
 *   complex float ({re,im} pairs). */
int crava_hip_perk_tables(int ntheta, int nz, int nzp, double lz, float low_cut, float high_cut,
                          float min_rel_thick, const float *wavelets, const float *wnorm,
                          unsigned char *invert, float *kWD, float *e1, float *e2, float *e3);

/* The per-k vectors for a simbox that is not of constant thickness, host
 * side (src/crava.cpp:1058-1083, fillInverseAbskWRobust :1361-1384; the
 * data were divided by the wavelets first, crava_hip_divide_by_scale_wavelet):
 *   kWD[k][l] = kD(k) (K = D A), e1[k][l] = 1/wnorm[l],
 *   e2[k][l] = kD3(k) / |e3_l| (kD3 the spectrum of the central o backward
 *   difference of the forward-difference operator, e3_l = central o backward
 *   o forward difference of wavelet l, Wavelet1D src/wavelet1D.cpp:420-460),
 *   e3[k][l] = 1 / sqrt(max(|W_l(k)|^2, |W_l(k, 0.999)|^2)) with
 *   Wavelet::getCAmp(k, scale) (src/wavelet.cpp:327-385), or
 *   wnorm[l] nzp^2 100 where both vanish.
 * Arguments as crava_hip_perk_tables (min_rel_thick = Simbox::getMinRelThick). */
int crava_hip_perk_tables_nonconst(int ntheta, int nz, int nzp, double lz, float low_cut, float high_cut,
                                   float min_rel_thick, const float *wavelets, const float *wnorm,
                                   unsigned char *invert, float *kWD, float *e1, float *e2, float *e3);

/* Wavelet::resample (src/wavelet.cpp:416-463), host side: nzp samples in FFT
 * order (z = dz k for k <= nzp/2, dz (k - nzp) after) from a wavelet of nw
 * samples at spacing dz_wavelet with its centre at sample `center`, each by
 * Wavelet::getWaveletValue's kriging of the six neighbouring samples with a
 * 20 Hz Ricker correlation (:1078-1122; samples outside the wavelet are 0);
 * *norm (may be NULL) receives Wavelet::norm_, the square root of the sum of
 * squares.  Wavelet::scale (:477-486) is a multiplication of `out` by the
 * gain of crava_hip_wavelet_energy.  CRAVA_HIP_E_INVALID where a kriging
 * matrix is not positive definite (NRLib::CholeskySolve throws there). */
int crava_hip_wavelet_resample(const float *wavelet, int nw, int center, float dz_wavelet, float dz, int nzp,
                               float *out, float *norm);

/* Padding sizes, host side.  crava_hip_padding_size: ModelGeneral::
 * setPaddingSize (src/modelgeneral.cpp:581-587), the smallest
 * 2^a 3^b 5^c 7^d >= ceil(n (1 + pad_fac)) as FFTGrid::
 * findClosestFactorableNumber finds it (src/fftgrid.cpp:2837-2866, loop
 * bounds from float logarithms).  crava_hip_estimate_xy_padding /
 * crava_hip_estimate_z_padding: ModelGeneral::estimateXYPaddingSizes /
 * estimateZPaddingSize (:1337-1425) -- with estimate != 0 the factors come
 * from half the lateral variogram range (range1, range2 in m, angle in
 * radians, at least one cell, at most 1) or from the default wavelet length
 * over lz * min_rel_thick (at most 1); the outputs are the padded sizes and
 * the true factors (npad - n) / n. */
int crava_hip_padding_size(int n, double pad_fac, int *npad);
int crava_hip_estimate_xy_padding(int nx, int ny, double dx, double dy, double lx, double ly, double x_pad_fac,
                                  double y_pad_fac, int estimate, float range1, float range2, float angle,
                                  int *nxpad, int *nypad, double *true_x_pad_fac, double *true_y_pad_fac);
int crava_hip_estimate_z_padding(int nz, double lz, double min_rel_thick, double z_pad_fac, int estimate,
                                 float default_wavelet_length, int *nzpad, double *true_z_pad_fac);

/* SeismicParametersHolder::computeCircCorrT with makeCircCorrTPosDef
 * (src/seismicparametersholder.cpp:431-499), host side: the prior temporal
 * correlation prior_corr_t[0..n) extended symmetrically to nzp lags, made
 * positive definite (|spectrum|, bins k <= min_int_fq zeroed -- the
 * reference's lowIntCut, src/modelgeneral.cpp:3961) and scaled to 1 at lag 0;
 * out: nzp floats.  The reference's float steps, each 1-D transform
 * evaluated in double and rounded once.  CRAVA_HIP_E_INVALID where the
 * reference asserts (prior_corr_t[0] == 0) or exits (c(0) <= 1e-5). */
int crava_hip_circ_corr_t(const float *prior_corr_t, int n, int min_int_fq, int nzp, float *out);

/* The robustification of the prior covariance in
 * SeismicParametersHolder::setCorrelationParameters
 * (src/seismicparametersholder.cpp:114-135), host side: with DGEEV's
 * eigenvalues lambda_k and right eigenvectors V (columns), every k with
 * lambda_k < 0.001 max(lambda) adds (0.0011 max - lambda_k) V(k,i) V(k,j)
 * to entry (i,j) -- row k of V, as the reference indexes it.  prior_var0
 * and out: row-major 3x3 (may alias); *changed (may be NULL): 0 when Sigma0
 * was kept, 1 when lifted.  CRAVA_HIP_E_INVALID for an ill-conditioned
 * Sigma0 whose DGEEV eigenvectors this library does not reproduce (a zero
 * off-diagonal entry other than a diagonal Sigma0, or rows DGEBAL would
 * scale). */
int crava_hip_prior_var0_robustify(const double *prior_var0, double *out, int *changed);

/* Crava::computeVariances' variances and energy-matching gain (src/crava.cpp:
 * 400-466, computeWDCorrMVar :1763-1775), host side.  wavelets [ntheta][nzp]
 * in FFT order; A [ntheta][3]; prior_var0 Sigma0 (row-major 3x3); corr_t
 * [nzp] the filtered temporal correlation (getPriorCorrTFiltered);
 * error_variance and emp_sn per angle (Crava::setupErrorCorrelation);
 * match[l] != 0 where the angle's energies are matched (NULL: none).
 * Outputs per angle: gain (1 when not matched; the caller scales the wavelet
 * by it), model variance, signal variance and the theoretical S/N. */
int crava_hip_wavelet_energy(int ntheta, int nzp, const float *wavelets, const float *A, const double *prior_var0,
                             const float *corr_t, const float *error_variance, const float *emp_sn,
                             const unsigned char *match, float *gain, float *model_var, float *signal_var,
                             float *theo_sn);

/* Reflection matrix of ModelAVODynamic::setupDefaultReflectionMatrix
 * (src/modelavodynamic.cpp:876-928), host side: row l holds the
 * coefficients of (ln Vp, ln Vs, ln rho) changes at incidence angle
 * theta[l] (radians) for the ratio vsvp = Vs/Vp, computed in double and
 * rounded to float.  seismic_type[l]: 0 PP (ModelSettings::STANDARDSEIS),
 * 1 PS (PSSEIS); NULL means every angle PP.  A: [ntheta][3]. */
int crava_hip_reflection_matrix(int ntheta, const float *theta, const int *seismic_type, double vsvp, float *A);

/* Posterior problem: everything that is not a grid. */
typedef struct {
  int nxp, nyp, nzp, ntheta;   /* ntheta in 1..8 */
  float wnc;                   /* white-noise component (ModelSettings, default 0.1) */
  double sigma0[9];            /* prior 3x3 covariance Sigma0, row-major (ln Vp, ln Vs, ln rho) */
  float A[8 * 3];              /* reflection matrix [ntheta][3] (ModelAVODynamic A_) */
  double err_theta_cov[8 * 8]; /* Sigma_e [ntheta][ntheta] (Crava::errThetaCov_) */
  double mixed_threshold;      /* CRAVA_HIP_MIXED: FP64 where tr(E)/tr(KSK^H) < this (for the
                                  whole 64-point wave such a point belongs to) */
} crava_hip_post_params;

/* Upload the per-k tables of one posterior problem.  The per-k vectors are
 * those of crava_hip_perk_tables (or the caller's own). */
int crava_hip_post_create(crava_hip_ctx *ctx, const crava_hip_post_params *params, const unsigned char *invert,
                          const float *kWD, const float *e1, const float *e2, const float *e3,
                          crava_hip_post **post);
int crava_hip_post_destroy(crava_hip_post *post);

/* Per-point posterior, in place, on Fourier-domain device grids.  Replaces
 * the loop of Crava::computePostMeanResidAndFFTCov (src/crava.cpp:1042-1160):
 *   mu[3]      prior mean (ln Vp, ln Vs, ln rho)  -> posterior mean
 *   data[nt]   seismic data                       -> residual d - K mu_post
 *   cov[6]     prior covariance grids, order (aa, bb, rr, ab, ar, br),
 *              unscaled transforms                -> posterior covariance
 *   errcorr    error-correlation grid (read only)
 * Out of band, or where the Cholesky factorisation fails, the prior passes
 * through (src/crava.cpp:1098,1109).  n_fail / n_fp64 (host, may be NULL):
 * if given, the call synchronises and reports the number of failed
 * factorisations / points solved in FP64. */
int crava_hip_post_run(crava_hip_post *post, int precision, void *const *mu, void *const *data, void *const *cov,
                       const void *errcorr, long long *n_fail, long long *n_fp64);

/* The same posterior for a separable prior covariance Sigma0 (x) corr, the
 * form SeismicParametersHolder::initializeCorrelations builds
 * (src/seismicparametersholder.cpp:197-221: the six covariance grids are
 * one correlation grid times priorVar0_(a,b)).  The prior comes from the
 * one Fourier-domain correlation grid `corr` (unscaled transform, read
 * only) as S(p) = |Re corr(p)| Sigma0, which equals the cleaned prior of the
 * six transformed grids (src/seismicparametersholder.cpp:373-426) in exact
 * arithmetic; cov_out[6] (order aa, bb, rr, ab, ar, br) receive the
 * posterior covariance (the cleaned prior out of band and where the
 * factorisation fails).  mu, data, errcorr, precision and the counts as
 * crava_hip_post_run.  The caller transforms one correlation grid instead
 * of six and the kernel reads one prior grid instead of six. */
int crava_hip_post_run_separable(crava_hip_post *post, int precision, void *const *mu, void *const *data,
                                 const void *corr, const void *errcorr, void *const *cov_out, long long *n_fail,
                                 long long *n_fp64);

/* The error-correlation spectrum as one Fourier plane.  With zero
 * correlation gradients (no correlation direction: ModelGeneral::
 * getCorrGradIJ gives 0, 0; src/crava.cpp:144-157) FFTGrid::fillInErrCorr
 * (src/fftgrid.cpp:1065-1116) writes priorCorrXY into plane k = 0 and zeros
 * into every other plane, so the grid's unscaled transform (COVARIANCE
 * fftInPlace, src/crava.cpp:1011) equals the 2-D transform of that plane on
 * every kz plane.  This fills that plane and transforms it: plane is a
 * device buffer of nyp rows of 2 (nxp/2 + 1) floats (doubles for
 * CRAVA_HIP_F64), which becomes nyp rows of nxp/2 + 1 complex values;
 * precision CRAVA_HIP_F32, CRAVA_HIP_MIXED (float storage, double
 * arithmetic) or CRAVA_HIP_F64.  With nonzero gradients the grid is not
 * constant along kz: fill and transform the whole grid instead. */
int crava_hip_err_corr_plane(crava_hip_ctx *ctx, void *plane, int nxp, int nyp, int nzp, const float *corr_xy,
                             int precision);

/* The transformed correlation grid of a separable prior with zero
 * correlation gradients, without filling it: fillInParamCorr
 * (src/fftgrid.cpp:1119-1164) then writes corrXY(i, j) circT(k) (sub_k = k
 * exactly), whose unscaled 3-D transform (COVARIANCE) is the 2-D transform
 * of corrXY times the 1-D transform of circT.  grid (device, the FFTGrid
 * layout, float or double for CRAVA_HIP_F64) receives that spectrum, as
 * crava_hip_fill_param_corr with scale 1 followed by crava_hip_fft_r2c
 * would give it up to the fill's rounding of each product; corr_xy
 * [nyp][nxp] and circ_t [nzp] are device arrays.  The 2-D transform uses
 * `precision` (CRAVA_HIP_MIXED: double arithmetic), the 1-D one and the
 * product double arithmetic, rounded once. */
int crava_hip_corr_spectrum_separable(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp,
                                      const float *corr_xy, const float *circ_t, int precision);

/* plane != 0: every later posterior run of `post` (crava_hip_post_run*,
 * crava_hip_nll_grad, batches created after this call) takes `errcorr` as the
 * one Fourier plane of crava_hip_err_corr_plane, read for every kz plane,
 * instead of a transformed grid; a batch does not transform such a volume's
 * error correlation.  plane == 0 restores grids.  A batch keeps the form its
 * volumes had when it was created (running it after a change is an error).
 * No reference counterpart: it saves the error-correlation grid's fill and
 * 3-D transform and the posterior's read of it. */
int crava_hip_post_set_errcorr_plane(crava_hip_post *post, int plane);

/* crava_hip_post_run that allocates nothing.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The mixed form's FP64 queue lives in a static device array
 * of the library (2^20 entries of one 64-point wave each: problems of up to
 * 67 108 864 in-band Fourier points; larger ones are refused with
 * CRAVA_HIP_E_INVALID); each GPU holds its own copy until the process ends,
 * and the call clears its count before use.  Calls that overlap on one GPU
 * (concurrent streams or threads) share that array and race.  Same
 * arguments, results and counts as crava_hip_post_run; without n_fail /
 * n_fp64 the call does not synchronise. */
int crava_hip_post_run_static(crava_hip_post *post, int precision, void *const *mu, void *const *data,
                              void *const *cov, const void *errcorr, long long *n_fail, long long *n_fp64);

/* Negative log marginal likelihood of the data and its gradient (no
 * reference counterpart; the model is the one the posterior loop assumes,
 * src/crava.cpp:1087-1130):
 *   NLL = 1/2 sum over in-band points of w [r^H M^-1 r + log det M] (+ const),
 *   r = d - K mu,  M = K S K^H + E,  w = 1 on self-conjugate x planes, else 2,
 * with parameters
 *   scales[nt]    per-angle wavelet scale s_l (K row l = s_l K0_l, K0 from the
 *                 per-k vectors given to crava_hip_post_create),
 *   noise_var[nt] per-angle noise variance v_l (Sigma_e,lm = sqrt(v_l v_m) rho_lm,
 *                 rho from the params' err_theta_cov),
 *   chol0[6]      lower Cholesky factor of Sigma0 (L00, L10, L11, L20, L21, L22);
 *                 S(p) = |c(p)| L0 L0^T with |c(p)| = |Re cov[0](p)| / params Sigma0_aa.
 * Grids are the Fourier-domain mu[3], data[nt], cov[6] (only cov[0] is read)
 * and errcorr, float (precision CRAVA_HIP_F32) or double (CRAVA_HIP_F64).
 * Outputs (host): *nll and grad[2*nt+6] = dNLL/d(scales, noise_var, chol0).
 * Arithmetic is FP64; Cholesky factors are recomputed per point.  Partial
 * sums are reduced in a fixed order, so repeated calls give identical results. */
int crava_hip_nll_grad(crava_hip_post *post, int precision, void *const *mu, void *const *data, void *const *cov,
                       const void *errcorr, const double *scales, const double *noise_var, const double *chol0,
                       double *nll, double *grad);

/* Fill a half-complex float grid with white complex Gaussian noise in place
 * of FFTGrid::fillInComplexNoise (src/fftgrid.cpp:1214-1263): CN(0,1) per
 * point, real N(0,1) at self-conjugate points, conjugate pairs on the
 * self-conjugate x planes.  Counter-based (Philox4x32-10) so the result does
 * not depend on launch shape; (seed, stream) select the realisation.  The
 * reference's sequential RandomGen stream is not reproduced on the device. */
int crava_hip_fill_complex_noise(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, uint64_t seed,
                                 uint64_t stream);

/* The reference's own noise, on the host: FFTGrid::fillInComplexNoise
 * (src/fftgrid.cpp:1214-1263) driven by RandomGen (libs/lib/random.cpp:
 * 38-119), the 32-bit LCG with Marsaglia-Bray normals, in the reference's
 * point order.  host_out: (nxp/2+1) x nyp x nzp {re, im} float pairs (the
 * half-complex grid, to upload with crava_hip_grid_upload); *seed is the
 * generator state (RandomGen::seed_) and is advanced as the reference
 * advances it, so consecutive calls continue its stream.  The stream is
 * sequential; use it where realisations must equal the reference's. */
int crava_hip_complex_noise_reference(int nxp, int nyp, int nzp, uint32_t *seed, float *host_out);

/* One posterior realisation, in place of the per-realisation body of
 * Crava::simulate (src/crava.cpp:1482-1573):
 *   seed[3]  in: complex noise grids (float); out: real realisation grids
 *   cov[6]   posterior covariance in the Fourier domain (read only)
 *   mean[3]  posterior mean, time domain (read only; NULL adds nothing)
 * Per point seed <- chol(Sigma_post) seed (0 where the factorisation fails,
 * :1503-1517), then c2r with PARAMETER scaling and + mean.  n_fail (host,
 * may be NULL) synchronises and reports failed factorisations. */
int crava_hip_simulate(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, void *const *cov, void *const *seed,
                       void *const *mean, long long *n_fail);

/* Only the per-point step of crava_hip_simulate (src/crava.cpp:1482-1521) on
 * npts contiguous half-complex points. */
int crava_hip_simulate_points(crava_hip_ctx *ctx, long long npts, void *const *cov, void *const *seed,
                              long long *n_fail);

/* Facies probabilities from per-facies posterior elastic PDFs, in place of
 * FaciesProb::CalculateFaciesProbFromPosteriorElasticPDF
 * (src/faciesprob.cpp:2198-2265, three elastic dimensions, no angle
 * weights) with PosteriorElasticPDF3D::FindDensity
 * (src/posteriorelasticpdf3d.cpp:511-604) on an unrotated volume.
 *   alpha/beta/rho  device, padded real grids [nzp][nyp][rnxp]
 *   tables          device, nfac tables laid out as FFTGrid(tn1,tn2,tn3)
 *   volume          host {x0, y0, dx, dy, ztop, zbot}; vn host {nx, ny, nz}
 *   prior           host [nfac]; p_undefined as the reference
 *   prob            host array of nfac+1 device grids [nz][ny][2*(nx/2+1)]
 *                   (last = undefined), RMISSING in the x padding
 *   argmax          device int [nz][ny][nx] most probable facies, or NULL
 *   counts          device uint64 [nfac] voxels per most probable facies, or NULL
 * Bit-exact with the reference's float/double arithmetic.  nfac <= 16. */
int crava_hip_facies(crava_hip_ctx *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp, const float *alpha,
                     const float *beta, const float *rho, int nfac, const float *tables, int tn1, int tn2, int tn3,
                     const double *volume, const int *vn, const float *prior, float p_undefined, float *const *prob,
                     int *argmax, unsigned long long *counts);

/* crava_hip_facies with the local-noise density dimension
 * (src/faciesprob.cpp:2160-2265, FindDensityFromPosteriorPDF :1277-1317):
 * 2^nang density sets, set d weighted per voxel by prod_j (bit j of d ?
 * t_j : 1 - t_j), t_j = (s_j - min s_j) / (max s_j - min s_j) (0 where s_j
 * is constant) with s_j the angle's noise-scale table at the voxel's trace.
 *   nang          0..8; nang = 0 is crava_hip_facies
 *   tables        device, (2^nang) x nfac tables, set d facies l at index d*nfac + l
 *   volumes, vn   host [2^nang][6] {x0, y0, dx, dy, ztop, zbot} and [2^nang][3]
 *                 {nx, ny, nz}: the lattice of each set (undefSum from set 0's)
 *   noise_scale   host, nang tables [ny][nx] (i fastest; RMISSING skipped in
 *                 the minimum and maximum)
 * The reference sizes its weight vectors before it counts the angles
 * (:2153-2157, a write past the end of empty vectors); this follows the
 * evident intent, one weight per angle.  Synchronises. */
int crava_hip_facies_local_noise(crava_hip_ctx *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp,
                                 const float *alpha, const float *beta, const float *rho, int nfac, int nang,
                                 const float *tables, int tn1, int tn2, int tn3, const double *volumes,
                                 const int *vn, const double *const *noise_scale, const float *prior,
                                 float p_undefined, float *const *prob, int *argmax, unsigned long long *counts);

/* A facies density table from samples, in place of the PosteriorElasticPDF3D
 * constructor (src/posteriorelasticpdf3d.cpp:12-138, SetupSmoothingGaussian3D
 * :360-413): the histogram of the npts points (d1, d2, d3: host double) on
 * the n1 x n2 x n3 lattice over limits {min1, max1, min2, max2, min3, max3}
 * (each point 1/npts; points outside dropped) circularly convolved with the
 * normalised Gaussian exp(-x^T sigma^-1 x / 2) (sigma: host, row-major 3x3)
 * at the wrapped lattice lags, by the float 3-D transforms with PARAMETER
 * scaling and a multiplication by sqrt(n1 n2 n3).  table: device float grid
 * [n3][n2][2(n1/2+1)] (FFTGrid(n1, n2, n3)), overwritten.  Uses the context
 * workspace; synchronises.  The reference's tables are 100 x 100 x 50
 * (100 x 1 x 50 without Vs). */
int crava_hip_kde_table(crava_hip_ctx *ctx, long long npts, const double *d1, const double *d2, const double *d3,
                        const double *sigma, int n1, int n2, int n3, const double *limits, void *table);

/* The density sets of FaciesProb::MakePosteriorElasticPDF3D
 * (src/faciesprob.cpp:926-1135) from well samples: for set d the samples
 * (for d > 0 transformed by sample_transform[d], row-major 3x3, e.g.
 * crava_hip_local_noise_sigmanew at the set's corner scales; NULL: none),
 * CalculateVariances, hopt = (4/7)^(1/7) nobs^(-1/7), the smoothing
 * covariance sigma_e[d] (row-major 3x3, upper triangle read; with no_vs its
 * alpha/rho 2x2 block in the top-left 2x2; NULL: zero) with each diagonal
 * raised to at least (hopt sd)^2, the lattice min - 5 sd .. max + 5 sd over
 * all samples, and one crava_hip_kde_table per facies (samples whose
 * facies_log is l; IMISSING = -99999 excluded).  Outputs: tables[d*nfac + l]
 * (device grids as crava_hip_kde_table), volumes [ndens][6] and vn
 * [ndens][3] for crava_hip_facies_local_noise. */
int crava_hip_facies_densities(crava_hip_ctx *ctx, int nsamples, const float *alpha, const float *beta,
                               const float *rho, const int *facies_log, int nfac, int ndens,
                               const double *sample_transform, const double *sigma_e, int no_vs, int n1, int n2,
                               int n3, void *const *tables, double *volumes, int *vn);

/* Crava::newPosteriorCovPointwise's sigmanew for one vector of noise scales
 * (src/crava.cpp:2214-2330), host side: sqrt(deltanew) (Sm - Smd)^(-1/2),
 * deltanew = Sm G^T (G Sm G^T + D E D)^+ G Sm, D = diag(sqrt(scale)).  G
 * [ntheta][3], errThetaCov [ntheta][ntheta], priorVar0 / postVar0 and
 * sigmanew 3x3, all row-major.  MakePosteriorElasticPDF3D's sample
 * transform of set d uses scale_j = max s_j / min s_j for bit j of d set,
 * else 1. */
int crava_hip_local_noise_sigmanew(int ntheta, const double *G, const double *errThetaCov, const double *priorVar0,
                                   const double *postVar0, const double *scale, double *sigmanew);

/* Prior correlation of the parameter covariance grids, in place of
 * FFTGrid::fillInParamCorr (src/fftgrid.cpp:1119-1164) followed by
 * multiplyByScalar(Sigma0_ab) as in
 * SeismicParametersHolder::initializeCorrelations
 * (src/seismicparametersholder.cpp:192-225): one pass writes every grid of
 * the set (1..16 real padded float grids, time domain):
 *   grid_g(i,j,k) = [(1-f) c(b) + f c(b+1 mod nzp)] * corr_xy[i + nxp*j] * scale[g]
 * with b + f = k + ci*grad_i + cj*grad_j (ci, cj the signed lateral lags),
 * RMISSING*scale[g] in the x padding.  corr_xy: device float [nyp*nxp];
 * circ_corr_t: device float [nzp] (computeCircCorrT); scale: host. */
int crava_hip_fill_param_corr(crava_hip_ctx *ctx, void *const *grids, int ngrid, const float *scale, int nxp, int nyp,
                              int nzp, const float *corr_xy, const float *circ_corr_t, float grad_i, float grad_j);

/* Error correlation grid, in place of FFTGrid::fillInErrCorr
 * (src/fftgrid.cpp:1065-1116): the lateral correlation at the (shifted) zero
 * time lag, linearly split between the two nearest samples; 0 elsewhere;
 * RMISSING in the x padding. */
int crava_hip_fill_err_corr(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, const float *corr_xy,
                            float grad_i, float grad_j);

/* Element-wise operators of FFTGrid (src/fftgrid.cpp:1877-1999) on a float
 * grid in the real domain (transformed = 0: rsize floats incl. padding) or
 * transformed (csize complex).  `other` is the second grid of the binary
 * operators; `scalar` the value of the scalar ones.  Scalar operators and
 * the fill are real-domain only and conjugate is transformed-only, as the
 * reference asserts (CRAVA_HIP_E_STATE otherwise). */
#define CRAVA_HIP_OP_ADD 0             /* add(grid)          */
#define CRAVA_HIP_OP_SUBTRACT 1        /* subtract(grid)     */
#define CRAVA_HIP_OP_MULTIPLY 2        /* multiply(grid)     */
#define CRAVA_HIP_OP_CHANGE_SIGN 3     /* changeSign()       */
#define CRAVA_HIP_OP_CONJUGATE 4       /* conjugate()        */
#define CRAVA_HIP_OP_MULTIPLY_SCALAR 5 /* multiplyByScalar() */
#define CRAVA_HIP_OP_ADD_SCALAR 6      /* addScalar()        */
#define CRAVA_HIP_OP_FILL_CONSTANT 7   /* fillInConstant()   */
#define CRAVA_HIP_OP_COPY 8            /* grid <- other, the FFTGrid(FFTGrid*) copy (src/fftgrid.cpp:69-100) */
int crava_hip_grid_op(crava_hip_ctx *ctx, int op, void *grid, const void *other, int nxp, int nyp, int nzp,
                      int transformed, float scalar);

/* Variance of one seismic stack, in place of Crava::computeDataVariance
 * (src/crava.cpp:322-363): mean of d^2 over the nonzero samples inside
 * nx x ny x nz, summed in the reference's order (double per row, rows, then
 * planes), so the float result is identical.  Synchronises. */
int crava_hip_data_variance(crava_hip_ctx *ctx, const void *grid, int nx, int ny, int nz, int nxp, int nyp, int nzp,
                            float *variance, long long *ndata);

/* Facies probabilities with constant prior facies probabilities, in place
 * of FaciesProb::calculateFaciesProbGeomodel (src/faciesprob.cpp:2286-2330):
 * prob[l] <- prior[l] * prob[nfac] + prob[l] inside nx, RMISSING in the x
 * padding; prob[] as in crava_hip_facies (nfac + 1 grids, last = undefined). */
int crava_hip_facies_geomodel(crava_hip_ctx *ctx, int nx, int ny, int nz, int nfac, const float *prior,
                              float *const *prob);

/* One inversion volume of a batch: its problem (from crava_hip_post_create,
 * which fixes nxp, nyp, nzp and ntheta) and its device grids in the FFTGrid
 * layout, float (CRAVA_HIP_F32 / CRAVA_HIP_MIXED) or double (CRAVA_HIP_F64). */
typedef struct {
  crava_hip_post *post;
  void *mu[3];      /* prior mean, time domain         -> posterior mean, time domain */
  void *data[8];    /* seismic data, time domain        -> residuals, time domain (first ntheta used) */
  void *cov[6];     /* prior covariance (aa,bb,rr,ab,ar,br), time domain -> posterior covariance, time domain */
  void *errcorr;    /* error correlation, time domain   -> its spectrum (consumed); or, if `post`
                       takes one plane (crava_hip_post_set_errcorr_plane), that plane's spectrum
                       (read only) */
  void *corr;       /* NULL, or a separable prior's correlation grid, time domain -> its spectrum
                       (consumed); cov[] are then outputs only (crava_hip_post_run_separable) */
} crava_hip_batch_volume;

/* Many volumes of uneven size and ntheta inverted as one sequence.  No
 * reference counterpart: the reference inverts one volume per Crava object
 * (main.cpp:279-308).  Per volume, as Crava::computePostMeanResidAndFFTCov
 * (src/crava.cpp:887-1204): forward transforms of data and mean (scaled as
 * DATA/PARAMETER) and of the covariance (or separable correlation) and
 * error-correlation grids (COVARIANCE), the posterior, then inverse
 * transforms of mean, residuals and covariances.  CRAVA_HIP_MIXED transforms
 * the covariance (or correlation) grids forward with double arithmetic in
 * each pass and the means with CRAVA_HIP_F32_DEFLATE.  All transforms of
 * one volume and kind are one launch per pass; the posterior of every volume
 * is one persistent launch.  The grids are registered here and must stay
 * allocated, like `ctx` and every `post`, until crava_hip_batch_destroy. */
int crava_hip_batch_create(crava_hip_ctx *ctx, int precision, int nvol, const crava_hip_batch_volume *vols,
                           crava_hip_batch **batch);

/* Time the transform variants of every grid shape of the batch and the
 * posterior launch and keep the fastest.  Posterior launch forms:
 *   0  one persistent launch over every volume (registers of the largest ntheta),
 *   1  one persistent launch per ntheta class, the classes running concurrently,
 *   2  one launch of the single-volume kernel per volume, on the volumes' streams;
 * forms 0 and 1 with 64/128/256 points per block and 0 (one workgroup per
 * tile) or 1024..8192 workgroups.  The registered grids' contents are saved
 * to a temporary device buffer and restored.  Outputs (host, may be NULL):
 * the chosen points per block, workgroups, form and posterior time. */
int crava_hip_batch_autotune(crava_hip_batch *batch, int *points_per_block, int *blocks, int *form,
                             float *post_ms);

/* Set the posterior launch directly (points_per_block 64, 128 or 256;
 * blocks 0 = one workgroup per tile; form 0, 1 or 2 as above). */
int crava_hip_batch_set_launch(crava_hip_batch *batch, int points_per_block, int blocks, int form);

/* Run the whole batch on the context's stream, asynchronously.  use_graph:
 * 0 enqueues the launches directly; 1 captures them into a hipGraph on first
 * use (and after autotune / set_launch) and replays it. */
int crava_hip_batch_run(crava_hip_batch *batch, int use_graph);

/* One eager pass of the batch with events between its phases, for the
 * per-phase rooflines: ms[0] the forward transforms of every volume, ms[1]
 * the posterior, ms[2] the inverse transforms.  The grids are transformed
 * and solved as by crava_hip_batch_run; synchronises.  No counterpart in
 * the reference (one volume per Crava there). */
int crava_hip_batch_phase_times(crava_hip_batch *batch, float *ms);

/* Per-volume counters of the last run (host arrays [nvol], may be NULL):
 * failed factorisations and points solved in FP64.  Synchronises. */
int crava_hip_batch_counts(crava_hip_batch *batch, long long *n_fail, long long *n_fp64);

int crava_hip_batch_destroy(crava_hip_batch *batch);

/* ---- Time-lapse (4D) merge ---------------------------------------------- */

/* Merge the static and dynamic parts of the prior into the current state's
 * means, in place of the mean loop of State4D::merge (src/state4d.cpp:90-146):
 * mu[l] = mu_static[l] + mu_dynamic[l], transformed grids (nxp/2+1 x nyp x
 * nzp complex float each).  mu may alias either input.  Float, bit-exact. */
int crava_hip_merge_mean_4d(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, const void *const *mu_static,
                            const void *const *mu_dynamic, void *const *mu);

/* The covariance merge of State4D::mergeCov (src/state4d.cpp:161-236), which
 * Crava::computePostMeanResidAndFFTCov calls in place of FFTCovGrids when 4D
 * is active (src/crava.cpp:992-1001): sigma = Sigma_ss + Sigma_dd +
 * Sigma_sd + Sigma_sd^H, per point, in the reference's float order.
 *   ss[6], dd[6]  upper triangles in the order aa, ab, ar, bb, br, rr
 *   sd[9]         static-dynamic blocks, sd[3a+b] = Cov(static a, dynamic b)
 *   sigma[6]      output, the order of ss (may alias ss or dd) */
int crava_hip_merge_cov_4d(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, const void *const *ss,
                           const void *const *dd, const void *const *sd, void *const *sigma);

/* ---- Forward modelling ------------------------------------------------- */

/* Synthetic seismic per angle from the elastic parameter grids, in place of
 * the trace loop of Crava::computeSyntSeismic (src/crava.cpp:1667-1759) with
 * Crava::computeSeismicImpedance (:1632-1663) and the local wavelet of
 * Wavelet1D::createLocalWavelet1D / shiftAndScale (src/wavelet1D.cpp:501-561)
 * read through Wavelet::getCAmp(k, scale) (src/wavelet.cpp:328-384).
 * Per trace (i < nx, j < ny): impedance A_l . (alpha, beta, rho) for k < nz,
 * the linear taper to nzp, forward difference, r2c, times conj(w), c2r, /nzp.
 * Padding traces receive the impedance, as the grid the reference writes
 * into holds it there.  Per-trace arrays are [ny][nx] (i fastest); a NULL
 * table or entry means shift 0, gain 1, stretch 1 for that angle.  All
 * arrays are host memory. */
typedef struct {
  int nx, ny, nz;                 /* volume */
  int nxp, nyp, nzp;              /* padded FFT grid */
  int ntheta;                     /* 1..8 */
  float dz;                       /* wavelet sampling interval, the unit of shift */
  const float *A;                 /* [ntheta][3] reflection matrix rows (Crava::A_) */
  const float *wavelets;          /* [ntheta][nzp] wavelets in FFT order (Wavelet::rAmp_) */
  const float *const *shift;      /* [ntheta] local time shifts (Wavelet::getLocalTimeshift) */
  const float *const *gain;       /* [ntheta] local gains (Wavelet::getLocalGainFactor) */
  const float *const *stretch;    /* [ntheta] sf = relThick * localStretch (:1706) */
} crava_hip_synt_params;

/* alpha, beta, rho: real padded float grids in the time domain (read only);
 * out[ntheta]: real padded float grids (written).  Asynchronous on the
 * context's stream once the tables are uploaded. */
int crava_hip_synthetic_seismic(crava_hip_ctx *ctx, const crava_hip_synt_params *p, const void *alpha,
                                const void *beta, const void *rho, void *const *out);

/* ---- Non-constant-thickness data preparation --------------------------- */

/* Divides the seismic of every padded trace by the local scale wavelet, in
 * place of the trace loop of Crava::divideDataByScaleWavelet
 * (src/crava.cpp:592-714) with Crava::computeAdjustmentFactor (:726-817),
 * for 1-D wavelets.  Per trace (i < nxp, j < nyp): scale by 1/sqrt(nzp),
 * taper k > nz (FFTGrid::getDistToBoundary), r2c, multiply by the adjustment
 * factor inside (lowCut, highCut) at the trace's frequency step
 * nz 1000 / (relThick lz nzp) and zero outside, c2r, scale by 1/sqrt(nzp).
 * The local wavelet of trace (i, j) is the angle's wavelet shifted and
 * scaled with the shift / gain at the mirrored volume trace (iInd, jInd) of
 * the reference and stretched by relThick(i, j).  All arrays are host
 * memory; a NULL shift / gain table or entry means shift 0 / gain 1.
 * Floating point: within rounding of the reference (the factor's
 * difference-operator spectrum is evaluated in closed form). */
typedef struct {
  int nx, ny, nz;                 /* volume */
  int nxp, nyp, nzp;              /* padded FFT grid */
  int ntheta;                     /* 1..8 */
  const float *wavelets;          /* [ntheta][nzp] global 1-D wavelets in FFT order (Wavelet::rAmp_) */
  const float *wavelet_norm;      /* [ntheta] Wavelet::getNorm() */
  const float *A;                 /* [ntheta][3] reflection matrix rows (Crava::A_) */
  const double *prior_var0;       /* [3][3] prior covariance (Crava::priorVar0_) */
  const float *corr_t;            /* [nzp] SeismicParametersHolder::getPriorCorrTFiltered(nz, nzp) */
  const float *err_var;           /* [ntheta] errThetaCov_[l][l] as float */
  const double *rel_thick;        /* [nyp][nxp] Simbox::getRelThick(i, j) of every padded trace, > 0 */
  const float *const *shift;      /* [ntheta] local time shifts [ny][nx] (Wavelet::getLocalTimeshift) */
  const float *const *gain;       /* [ntheta] local gains [ny][nx] (Wavelet::getLocalGainFactor) */
  float dz;                       /* wavelet sampling interval (Wavelet::dz_, ms), > 0 */
  double lz;                      /* Simbox::getlz(), > 0 */
  float low_cut, high_cut;        /* Crava::lowCut_, highCut_ (Hz) */
  float wnc;                      /* Crava::wnc_, white-noise component */
} crava_hip_swdiv_params;

/* grids[ntheta]: real padded float grids [nzp][nyp][2(nxp/2+1)] in the time
 * domain, divided in place.  Asynchronous on the context's stream once the
 * tables are uploaded. */
int crava_hip_divide_by_scale_wavelet(crava_hip_ctx *ctx, const crava_hip_swdiv_params *p, void *const *grids);

/* FFTGrid::interpolateSeismic (src/fftgrid.cpp:2564-2655) on one real padded
 * float grid, in place: volume traces whose energy over k < nz is at most
 * energy_threshold times the mean are rebuilt from their neighbours in the
 * reference's order (interpolateTrace, :2658-2696), then every padded trace
 * outside the box of traces above the threshold is its clamped fill trace
 * times the 3-D taper (extrapolateSeismic, :2699-2733).  Bit-exact.
 * n_flagged / n_zero (optional) receive the number of traces at or below the
 * threshold and of those with zero energy.  Synchronises the context's
 * stream (the schedule is derived from the energies on the host).  Fails
 * with CRAVA_HIP_E_INVALID when no trace is above the threshold (the
 * reference then extrapolates from traces outside the volume). */
int crava_hip_interpolate_seismic(crava_hip_ctx *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp,
                                  float energy_threshold, void *grid, int *n_flagged, int *n_zero);

/* ---- local-noise correction (src/crava.cpp:2130-2470, :1547-1566) ---- */

/* Inputs of Crava::correctAlphaBetaRho.  All matrices are host, row-major,
 * double. */
typedef struct crava_hip_local_noise {
  int nx, ny, nz;              /* volume */
  int nxp, nyp, nzp;           /* padded grid */
  int ntheta;                  /* 1..8 */
  const double *G;             /* [ntheta][3], Crava::computeG */
  const double *errThetaCov;   /* [ntheta][ntheta], errThetaCov_ */
  const double *priorVar0;     /* [3][3], priorVar0_ */
  const double *postVar0;      /* [3][3], postVar0_ */
  /* ntheta host tables [ny][nx]: getLocalNoiseScale(angle)(i, j) at j*nx + i */
  const double *const *noiseScale;
} crava_hip_local_noise;

/* Crava::computeG (src/crava.cpp:2130-2211) on the host.  The reference
 * takes the eigenvectors of Sm^(-1/2) (Sm - Smd)+ Sm^(-1/2) from LAPACK's
 * dgeev, whose signs G inherits (with noise scales that vary between angles
 * the correction depends on them).  Here they come from a restatement of
 * dgeev's path for a dense 3x3 matrix (balancing, Householder reduction,
 * Francis double-shift QR, back substitution); a matrix with a zero
 * off-diagonal entry falls back to Jacobi rotations with each eigenvector's
 * largest component positive.  Where the QR iteration's last reflections
 * depend on last-bit rounding the signs can still differ from a given
 * LAPACK build, so a caller that needs the reference's result exactly
 * passes the G of its own computeG to crava_hip_local_noise_correct.  Host
 * only; no context needed. */
int crava_hip_local_noise_g(int ntheta, const double *errThetaCov, const double *priorVar0, const double *postVar0,
                            double *G);

/* Crava::correctAlphaBetaRho's trace loop (src/crava.cpp:2340-2470):
 * post[3] (alpha, beta, rho: real padded float grids in the time domain)
 * are corrected in place towards mean2[3] (the prior means) by each
 * trace's sigmanew of newPosteriorCovPointwise (:2213-2312).  The per-trace
 * matrices are computed on the device (matrix functions by Jacobi
 * rotations; their sums differ from BLAS's in rounding, so the corrected
 * values agree to float rounding rather than bit for bit).  traceCov
 * (device, [ny][nx][9] row-major, optional) receives (*sigmamdnew_)(i,j),
 * the realisation factor that crava_hip_local_noise_realisation applies.
 * Synchronises the context's stream. */
int crava_hip_local_noise_correct(crava_hip_ctx *ctx, const crava_hip_local_noise *p, void *const *post,
                                  const void *const *mean2, double *traceCov);

/* The local-noise block of Crava::simulate (src/crava.cpp:1547-1566): seed[3]
 * (real padded float grids, the inverse-transformed draws, before the
 * posterior means are added) become traceCov(i,j) times (seed0, seed1,
 * seed2) at every volume point, in place.  Bit-exact for a given traceCov.
 * Asynchronous on the context's stream. */
int crava_hip_local_noise_realisation(crava_hip_ctx *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp,
                                      const double *traceCov, void *const *seed);

#ifdef __cplusplus
}
#endif

#endif /* CRAVA_HIP_H */
