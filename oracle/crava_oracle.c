/*
 * crava_oracle.c -- CPU restatement of CRAVA's Fourier-domain Bayesian AVO
 * inversion hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.  The
 * product (crava_amd/libcrava_hip.so) never links, loads or calls it.
 *
 * Each function names the reference file:line it restates.  Arithmetic is
 * float where the reference is float, in the reference's order, and the file
 * is built with -ffp-contract=off so no fused multiply-add changes a rounding
 * (the reference is built -O2 for plain x86-64, which has no FMA).  The one
 * deliberate difference: the 3-D transforms are computed in double and
 * rounded to float, where the reference runs FFTW 2.1.2 in float.
 *
 * Parity pins (tests/test_oracle.py):
 *   - lib_matr restatement + per-point loop: bit-exact against the reference's
 *     own lib_matr.c + FFTW 2.1.2 compiled from its sources (oracle/_ref);
 *   - FFT: closed forms (delta, single cosine mode), brute-force DFT on tiny
 *     odd/even sizes, Parseval, round trip, FFTW 2.1.2 (oracle/_ref);
 *   - posterior: dense brute-force Bayesian inversion on tiny grids (FP64),
 *     Woodbury identity, prior pass-through out of band and on failure;
 *   - RandomGen: LCG closed form, moments of rnorm01;
 *   - facies: an independent numpy trilinear lookup (clip-then-interpolate
 *     form), hand-computed node / cell-centre / clamp / negative-entry cases
 *     and the probability assembly (tests/test_oracle_facies.py);
 *   - NLL: central finite differences and an independent numpy evaluation;
 *   - non-constant-thickness per-k tables, Wavelet::resample and the
 *     energy-matching variances: closed-form difference-operator spectra,
 *     delta / zero wavelets, numpy's rfft with the 0.999 stretch, kriging
 *     exact at nodes and exact on a node-centred 20 Hz Ricker, numpy.linalg
 *     solves, the Toeplitz quadratic form and theo S/N = empirical S/N after
 *     matching (tests/test_oracle_wavelet.py).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

typedef struct { float re, im; } ocpx;   /* layout of fftw_complex (float) */
typedef struct { double re, im; } dcpx;

#define ORC_TOL 1e-20                    /* lib_matr.c:20 */
#define ORC_MAXT 8

/* ========================================================================
 * 1-D DFT in double (mixed radix, recursive on the smallest prime factor).
 * Sign -1 is FFTW's forward exponent (libs/fft/include/fftw.h:61).
 * ====================================================================== */
static int smallest_factor(int n)
{
  if (n % 4 == 0) return 4;
  for (int p = 2; p * p <= n; p++)
    if (n % p == 0) return p;
  return n;
}

static void dft_rec(dcpx *x, int n, int sign, dcpx *tmp)
{
  if (n == 1) return;
  int p = smallest_factor(n);
  int m = n / p;
  if (p == n) {                      /* prime (or 4): direct O(n^2) */
    for (int k = 0; k < n; k++) {
      double sr = 0.0, si = 0.0;
      for (int t = 0; t < n; t++) {
        long long q = ((long long)k * t) % n;
        double a = sign * 2.0 * M_PI * (double)q / (double)n;
        double c = cos(a), s = sin(a);
        sr += x[t].re * c - x[t].im * s;
        si += x[t].re * s + x[t].im * c;
      }
      tmp[k].re = sr; tmp[k].im = si;
    }
    memcpy(x, tmp, sizeof(dcpx) * n);
    return;
  }
  /* split into p decimated subsequences, transform each */
  dcpx *sub = (dcpx *)malloc(sizeof(dcpx) * n);
  for (int r = 0; r < p; r++)
    for (int q = 0; q < m; q++)
      sub[r * m + q] = x[r + p * q];
  for (int r = 0; r < p; r++)
    dft_rec(sub + r * m, m, sign, tmp);
  for (int k = 0; k < n; k++) {
    double sr = 0.0, si = 0.0;
    for (int r = 0; r < p; r++) {
      long long q = ((long long)r * k) % n;
      double a = sign * 2.0 * M_PI * (double)q / (double)n;
      double c = cos(a), s = sin(a);
      dcpx v = sub[r * m + (k % m)];
      sr += v.re * c - v.im * s;
      si += v.re * s + v.im * c;
    }
    x[k].re = sr; x[k].im = si;
  }
  free(sub);
}

void orc_dft1d(double *xy, int n, int sign)
{
  dcpx *tmp = (dcpx *)malloc(sizeof(dcpx) * (n > 0 ? n : 1));
  dft_rec((dcpx *)xy, n, sign, tmp);
  free(tmp);
}

/* transform every line of a [nz][ny][nx] complex array along one axis */
static void dft_axis(dcpx *a, int nx, int ny, int nz, int axis, int sign)
{
  int n = axis == 0 ? nx : (axis == 1 ? ny : nz);
  long long stride = axis == 0 ? 1 : (axis == 1 ? nx : (long long)nx * ny);
  dcpx *line = (dcpx *)malloc(sizeof(dcpx) * n);
  dcpx *tmp = (dcpx *)malloc(sizeof(dcpx) * n);
  int n1 = axis == 0 ? ny : nx;
  int n2 = axis == 2 ? ny : nz;
  for (int b2 = 0; b2 < n2; b2++)
    for (int b1 = 0; b1 < n1; b1++) {
      long long base;
      if (axis == 0) base = (long long)b1 * nx + (long long)b2 * nx * ny;
      else if (axis == 1) base = b1 + (long long)b2 * nx * ny;
      else base = b1 + (long long)b2 * nx;
      for (int t = 0; t < n; t++) line[t] = a[base + t * stride];
      dft_rec(line, n, sign, tmp);
      for (int t = 0; t < n; t++) a[base + t * stride] = line[t];
    }
  free(line);
  free(tmp);
}

/* ========================================================================
 * 3-D in-place real FFT on the FFTGrid layout (src/fftgrid.cpp:51-66):
 * real [nzp][nyp][rnxp] with rnxp = 2*(nxp/2+1), complex [nzp][nyp][cnxp].
 * ====================================================================== */

/* FFTGrid::fftInPlace (src/fftgrid.cpp:1804-1829): non-COVARIANCE grids are
 * first scaled by 1/sqrt(N) in float over the whole padded array. */
void orc_fft3d_r2c(float *g, int nxp, int nyp, int nzp, int is_cov)
{
  int cnxp = nxp / 2 + 1, rnxp = 2 * cnxp;
  long long rsize = (long long)rnxp * nyp * nzp;
  if (!is_cov) {
    float s = 1.0f / sqrtf((float)(nxp * nyp * nzp));
    for (long long t = 0; t < rsize; t++) g[t] *= s;
  }
  /* r2c along x: transform the full complex row and keep the half */
  long long hsize = (long long)cnxp * nyp * nzp;
  dcpx *h = (dcpx *)malloc(sizeof(dcpx) * hsize);
  dcpx *row = (dcpx *)malloc(sizeof(dcpx) * nxp);
  dcpx *tmp = (dcpx *)malloc(sizeof(dcpx) * nxp);
  for (long long r = 0; r < (long long)nyp * nzp; r++) {
    for (int i = 0; i < nxp; i++) { row[i].re = g[r * rnxp + i]; row[i].im = 0.0; }
    dft_rec(row, nxp, -1, tmp);
    for (int i = 0; i < cnxp; i++) h[r * cnxp + i] = row[i];
  }
  dft_axis(h, cnxp, nyp, nzp, 1, -1);
  dft_axis(h, cnxp, nyp, nzp, 2, -1);
  ocpx *c = (ocpx *)g;
  for (long long t = 0; t < hsize; t++) { c[t].re = (float)h[t].re; c[t].im = (float)h[t].im; }
  free(h); free(row); free(tmp);
}

/* FFTGrid::invFFTInPlace (src/fftgrid.cpp:1832-1862): complex inverse along
 * z and y on the half array, then a 1-D c2r along x that, like FFTW's
 * halfcomplex format, ignores Im of the k=0 and (even n) k=n/2 bins; finally
 * scale by 1/N (COVARIANCE) or 1/sqrt(N) in float. */
void orc_fft3d_c2r(float *g, int nxp, int nyp, int nzp, int is_cov)
{
  int cnxp = nxp / 2 + 1, rnxp = 2 * cnxp;
  long long hsize = (long long)cnxp * nyp * nzp;
  dcpx *h = (dcpx *)malloc(sizeof(dcpx) * hsize);
  ocpx *c = (ocpx *)g;
  for (long long t = 0; t < hsize; t++) { h[t].re = c[t].re; h[t].im = c[t].im; }
  dft_axis(h, cnxp, nyp, nzp, 2, +1);
  dft_axis(h, cnxp, nyp, nzp, 1, +1);
  dcpx *row = (dcpx *)malloc(sizeof(dcpx) * nxp);
  dcpx *tmp = (dcpx *)malloc(sizeof(dcpx) * nxp);
  for (long long r = 0; r < (long long)nyp * nzp; r++) {
    dcpx *hr = h + r * cnxp;
    for (int i = 0; i < nxp; i++) {
      if (i < cnxp) { row[i] = hr[i]; }
      else { row[i].re = hr[nxp - i].re; row[i].im = -hr[nxp - i].im; }
    }
    row[0].im = 0.0;
    if (nxp % 2 == 0) row[nxp / 2].im = 0.0;
    dft_rec(row, nxp, +1, tmp);
    for (int i = 0; i < nxp; i++) g[r * rnxp + i] = (float)row[i].re;
    for (int i = nxp; i < rnxp; i++) g[r * rnxp + i] = 0.0f;
  }
  float scale;
  if (is_cov) scale = (float)(1.0 / (double)(nxp * nyp * nzp));
  else scale = (float)(1.0 / (double)sqrtf((float)(nxp * nyp * nzp)));
  long long rsize = (long long)rnxp * nyp * nzp;
  for (long long t = 0; t < rsize; t++) g[t] *= scale;
  free(h); free(row); free(tmp);
}

/* Same transforms in double precision end to end (the FP64 oracle). */
void orc_fft3d_r2c_f64(double *g, int nxp, int nyp, int nzp, int is_cov)
{
  int cnxp = nxp / 2 + 1, rnxp = 2 * cnxp;
  long long rsize = (long long)rnxp * nyp * nzp;
  if (!is_cov) {
    double s = 1.0 / sqrt((double)nxp * nyp * nzp);
    for (long long t = 0; t < rsize; t++) g[t] *= s;
  }
  long long hsize = (long long)cnxp * nyp * nzp;
  dcpx *h = (dcpx *)malloc(sizeof(dcpx) * hsize);
  dcpx *row = (dcpx *)malloc(sizeof(dcpx) * nxp);
  dcpx *tmp = (dcpx *)malloc(sizeof(dcpx) * nxp);
  for (long long r = 0; r < (long long)nyp * nzp; r++) {
    for (int i = 0; i < nxp; i++) { row[i].re = g[r * rnxp + i]; row[i].im = 0.0; }
    dft_rec(row, nxp, -1, tmp);
    for (int i = 0; i < cnxp; i++) h[r * cnxp + i] = row[i];
  }
  dft_axis(h, cnxp, nyp, nzp, 1, -1);
  dft_axis(h, cnxp, nyp, nzp, 2, -1);
  memcpy(g, h, sizeof(dcpx) * hsize);
  free(h); free(row); free(tmp);
}

void orc_fft3d_c2r_f64(double *g, int nxp, int nyp, int nzp, int is_cov)
{
  int cnxp = nxp / 2 + 1, rnxp = 2 * cnxp;
  long long hsize = (long long)cnxp * nyp * nzp;
  dcpx *h = (dcpx *)malloc(sizeof(dcpx) * hsize);
  memcpy(h, g, sizeof(dcpx) * hsize);
  dft_axis(h, cnxp, nyp, nzp, 2, +1);
  dft_axis(h, cnxp, nyp, nzp, 1, +1);
  dcpx *row = (dcpx *)malloc(sizeof(dcpx) * nxp);
  dcpx *tmp = (dcpx *)malloc(sizeof(dcpx) * nxp);
  double scale = is_cov ? 1.0 / ((double)nxp * nyp * nzp) : 1.0 / sqrt((double)nxp * nyp * nzp);
  for (long long r = 0; r < (long long)nyp * nzp; r++) {
    dcpx *hr = h + r * cnxp;
    for (int i = 0; i < nxp; i++) {
      if (i < cnxp) { row[i] = hr[i]; }
      else { row[i].re = hr[nxp - i].re; row[i].im = -hr[nxp - i].im; }
    }
    row[0].im = 0.0;
    if (nxp % 2 == 0) row[nxp / 2].im = 0.0;
    dft_rec(row, nxp, +1, tmp);
    for (int i = 0; i < nxp; i++) g[r * rnxp + i] = row[i].re * scale;
    for (int i = nxp; i < rnxp; i++) g[r * rnxp + i] = 0.0;
  }
  free(h); free(row); free(tmp);
}

/* ========================================================================
 * lib_matr restatement (libs/lib/lib_matr.c), float, reference op order.
 * ====================================================================== */

/* lib_matrCholCpx, lib_matr.c:78-167 */
static int o_chol_cpx(int n, ocpx **x)
{
  int i, j, k;
  ocpx r;
  float help;
  float factor = x[0][0].re;
  if (factor <= 0) return 1;
  for (i = 0; i < n; i++)
    for (j = 0; j < n; j++) {
      x[i][j].re = x[i][j].re / factor;
      x[i][j].im = x[i][j].im / factor;
    }
  for (i = 0; i < n; i++) {
    if (x[i][i].re <= ORC_TOL) return 1;
    for (j = 0; j < i; j++) {
      r.re = 0.0f; r.im = 0.0f;
      for (k = 0; k < j; k++) {
        r.re += (x[i][k].re * x[j][k].re) + (x[i][k].im * x[j][k].im);
        r.im += -(x[i][k].re * x[j][k].im) + (x[i][k].im * x[j][k].re);
      }
      help = x[j][j].re * x[j][j].re + x[j][j].im * x[j][j].im;
      /* the second statement reads the already-updated .re, as written */
      x[i][j].re = ((x[i][j].re - r.re) * x[j][j].re + (x[i][j].im - r.im) * x[j][j].im) / help;
      x[i][j].im = ((x[i][j].im - r.im) * x[j][j].re - (x[i][j].re - r.re) * x[j][j].im) / help;
    }
    r.re = 0.0f; r.im = 0.0f;
    for (k = 0; k < i; k++)
      r.re += (x[i][k].re * x[i][k].re) + (x[i][k].im * x[i][k].im);
    r.re = x[i][i].re - r.re;
    if (r.re <= ORC_TOL) return 1;
    x[i][i].re = (float)(sqrt(r.re));
    x[i][i].im = 0.0f;
  }
  for (i = 0; i < n; i++)
    for (j = 0; j <= i; j++) {
      x[i][j].re *= (float)(sqrt(factor));
      x[i][j].im *= (float)(sqrt(factor));
    }
  return 0;
}

/* lib_matrAXeqBMatCpx, lib_matr.c:330-365 */
static void o_axeqb_mat_cpx(int n, ocpx **L, ocpx **X, int m)
{
  int li, lj, c;
  ocpx lx;
  float help;
  for (c = 0; c < m; c++) {
    for (li = 0; li < n; li++) {
      lx.re = X[li][c].re; lx.im = X[li][c].im;
      for (lj = 0; lj < li; lj++) {
        lx.re -= (X[lj][c].re * L[li][lj].re - X[lj][c].im * L[li][lj].im);
        lx.im -= (X[lj][c].im * L[li][lj].re + X[lj][c].re * L[li][lj].im);
      }
      help = (L[li][li].re * L[li][li].re + L[li][li].im * L[li][li].im);
      X[li][c].re = (lx.re * L[li][li].re + lx.im * L[li][li].im) / help;
      X[li][c].im = (lx.im * L[li][li].re - lx.re * L[li][li].im) / help;
    }
    for (li = n - 1; li >= 0; li--) {
      lx.re = X[li][c].re; lx.im = X[li][c].im;
      for (lj = n - 1; lj > li; lj--) {
        lx.re = lx.re - (X[lj][c].re * L[lj][li].re + X[lj][c].im * L[lj][li].im);
        lx.im = lx.im - (-X[lj][c].re * L[lj][li].im + X[lj][c].im * L[lj][li].re);
      }
      help = (L[li][li].re * L[li][li].re + L[li][li].im * L[li][li].im);
      X[li][c].re = (lx.re * L[li][li].re - lx.im * L[li][li].im) / help;
      X[li][c].im = (lx.im * L[li][li].re + lx.re * L[li][li].im) / help;
    }
  }
}

/* lib_matrProdCpx, lib_matr.c:544-566 */
static void o_prod_cpx(ocpx **a, ocpx **b, int n1, int n2, int n3, ocpx **out)
{
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n3; j++) {
      ocpx x = {0.0f, 0.0f};
      for (int k = 0; k < n2; k++) {
        x.re += a[i][k].re * b[k][j].re - a[i][k].im * b[k][j].im;
        x.im += a[i][k].im * b[k][j].re + a[i][k].re * b[k][j].im;
      }
      out[i][j] = x;
    }
}

/* lib_matrProdAdjointCpx, lib_matr.c:525-543: out = a * b^H */
static void o_prod_adjoint_cpx(ocpx **a, ocpx **b, int n1, int n2, int n3, ocpx **out)
{
  for (int i = 0; i < n1; i++)
    for (int j = 0; j < n3; j++) {
      ocpx x = {0.0f, 0.0f};
      for (int k = 0; k < n2; k++) {
        x.re += a[i][k].re * b[j][k].re + a[i][k].im * b[j][k].im;
        x.im += a[i][k].im * b[j][k].re - a[i][k].re * b[j][k].im;
      }
      out[i][j] = x;
    }
}

/* lib_matrProdMatVecCpx, lib_matr.c:720-735 */
static void o_prod_matvec_cpx(ocpx **m, const ocpx *v, int n1, int n2, ocpx *out)
{
  for (int i = 0; i < n1; i++) {
    ocpx x = {0.0f, 0.0f};
    for (int j = 0; j < n2; j++) {
      x.re += m[i][j].re * v[j].re - m[i][j].im * v[j].im;
      x.im += m[i][j].im * v[j].re + m[i][j].re * v[j].im;
    }
    out[i] = x;
  }
}

/* lib_matrProdAdjointMatVecCpx, lib_matr.c:761-778: out = m^H v */
static void o_prod_adjoint_matvec_cpx(ocpx **m, const ocpx *v, int n1, int n2, ocpx *out)
{
  for (int i = 0; i < n1; i++) {
    ocpx x = {0.0f, 0.0f};
    for (int j = 0; j < n2; j++) {
      x.re += m[j][i].re * v[j].re + m[j][i].im * v[j].im;
      x.im += -m[j][i].im * v[j].re + m[j][i].re * v[j].im;
    }
    out[i] = x;
  }
}

/* lib_matrProdCholVec, lib_matr.c:453-470 */
static void o_prod_chol_vec(int n, ocpx **m, ocpx *v)
{
  ocpx in[ORC_MAXT];
  for (int i = 0; i < n; i++) in[i] = v[i];
  for (int i = 0; i < n; i++) { v[i].re = 0.0f; v[i].im = 0.0f; }
  for (int i = 0; i < n; i++)
    for (int j = 0; j < i + 1; j++) {
      v[i].re += m[i][j].re * in[j].re - m[i][j].im * in[j].im;
      v[i].im += m[i][j].re * in[j].im + m[i][j].im * in[j].re;
    }
}

/* exported for the bit-exact check against the reference's own lib_matr.c */
int orc_chol_cpx(int n, float *mat /* n*n complex, row-major */)
{
  ocpx *rows[ORC_MAXT];
  for (int i = 0; i < n; i++) rows[i] = (ocpx *)mat + i * n;
  return o_chol_cpx(n, rows);
}

void orc_axeqb_mat_cpx(int n, float *L, float *X, int m)
{
  ocpx *lr[ORC_MAXT], *xr[ORC_MAXT];
  for (int i = 0; i < n; i++) { lr[i] = (ocpx *)L + i * n; xr[i] = (ocpx *)X + i * m; }
  o_axeqb_mat_cpx(n, lr, xr, m);
}

/* ========================================================================
 * Per-k tables (src/crava.cpp:1042-1057 constant-thickness branch,
 * :1337-1359 fillkW/fillkWNorm, src/wavelet.cpp:309-323 getCAmp,
 * src/wavelet1D.cpp:368-479 difference operators).
 * ====================================================================== */

/* 1-D r2c of an nzp real series into float spectrum[cnzp] (FFTW forward). */
static void spectrum_f(const float *w, int nzp, ocpx *spec)
{
  dcpx *x = (dcpx *)malloc(sizeof(dcpx) * nzp);
  dcpx *tmp = (dcpx *)malloc(sizeof(dcpx) * nzp);
  for (int t = 0; t < nzp; t++) { x[t].re = w[t]; x[t].im = 0.0; }
  dft_rec(x, nzp, -1, tmp);
  for (int t = 0; t < nzp / 2 + 1; t++) { spec[t].re = (float)x[t].re; spec[t].im = (float)x[t].im; }
  /* a real series has exactly real DC and Nyquist bins (FFTW's halfcomplex
   * format stores no imaginary part for them) */
  spec[0].im = 0.0f;
  if (nzp % 2 == 0) spec[nzp / 2].im = 0.0f;
  free(x); free(tmp);
}

/* Wavelet::getCAmp(k), src/wavelet.cpp:309-323 (Hermitian extension) */
static ocpx get_camp(const ocpx *spec, int nzp, int k)
{
  int cnzp = nzp / 2 + 1;
  ocpx v;
  if (k < cnzp) { v = spec[k]; }
  else { v.re = spec[nzp - k].re; v.im = -spec[nzp - k].im; }
  return v;
}

/* Wavelet::findNorm (time domain), src/wavelet.cpp:790-808 */
static float find_norm(const float *w, int nzp)
{
  double n2 = 0.0;
  for (int i = 0; i < nzp; i++) n2 += (double)(w[i] * w[i]);
  return (float)sqrt(n2);
}

/*
 * Build the per-k tables consumed by the per-point loop.
 *   wavelets   [ntheta][nzp]  scaled seismic wavelets in FFT order
 *   wnorm      [ntheta]       Wavelet::norm_ of each wavelet
 * Outputs, per k (all [nzp][ntheta] complex except invert):
 *   invert[k]  frequency gate, src/crava.cpp:1044,1098
 *   kWD        kD * conj(W(k))                        (:1050-1051)
 *   e1         conj(W(k)) / wnorm                      (:1055, fillkWNorm)
 *   e2         conj(E3(k)) / |e3|, e3 = central(backward(forward(w)))  (:1056)
 *   e3         1                                       (:1057)
 */
void orc_perk_tables(int ntheta, int nz, int nzp, float dz, double lz,
                     float low_cut, float high_cut, float min_rel_thick,
                     const float *wavelets, const float *wnorm,
                     unsigned char *invert, float *kWD, float *e1, float *e2, float *e3)
{
  int cnzp = nzp / 2 + 1;
  (void)dz;
  /* diff1 = forward difference operator: d[0] = -1, d[nzp-1] = +1 */
  float *d1 = (float *)calloc(nzp, sizeof(float));
  d1[0] = -1.0f; d1[nzp - 1] = 1.0f;
  ocpx *dspec = (ocpx *)malloc(sizeof(ocpx) * cnzp);
  spectrum_f(d1, nzp, dspec);
  ocpx *wspec = (ocpx *)malloc(sizeof(ocpx) * cnzp * ntheta);
  ocpx *e3spec = (ocpx *)malloc(sizeof(ocpx) * cnzp * ntheta);
  float *e3norm = (float *)malloc(sizeof(float) * ntheta);
  float *fw = (float *)malloc(sizeof(float) * nzp);
  float *bw = (float *)malloc(sizeof(float) * nzp);
  float *cw = (float *)malloc(sizeof(float) * nzp);
  for (int l = 0; l < ntheta; l++) {
    const float *w = wavelets + (long long)l * nzp;
    spectrum_f(w, nzp, wspec + l * cnzp);
    /* Wavelet1D(wavelet, FIRSTORDERFORWARDDIFF) etc., src/wavelet1D.cpp:431-460 */
    for (int i = 0; i < nzp; i++) fw[i] = (i == nzp - 1) ? (w[0] - w[i]) : (w[i + 1] - w[i]);
    for (int i = 0; i < nzp; i++) bw[i] = (i == 0) ? (fw[0] - fw[nzp - 1]) : (fw[i] - fw[i - 1]);
    for (int i = 0; i < nzp; i++) {
      if (i == 0) cw[i] = (float)(0.5 * (bw[i + 1] - bw[nzp - 1]));
      else if (i == nzp - 1) cw[i] = (float)(0.5 * (bw[0] - bw[i - 1]));
      else cw[i] = (float)(0.5 * (bw[i + 1] - bw[i - 1]));
    }
    e3norm[l] = find_norm(cw, nzp);
    spectrum_f(cw, nzp, e3spec + l * cnzp);
  }
  for (int k = 0; k < nzp; k++) {
    int kk = k < nzp - k ? k : nzp - k;
    float rf = (float)((nz * 1000.0f) / (lz * nzp) * kk);
    invert[k] = (rf > low_cut * min_rel_thick && rf < high_cut) ? 1 : 0;
    ocpx kD = get_camp(dspec, nzp, k);
    for (int l = 0; l < ntheta; l++) {
      ocpx w = get_camp(wspec + l * cnzp, nzp, k);
      ocpx kw; kw.re = w.re; kw.im = -w.im;                 /* fillkW */
      ocpx r;                                              /* lib_matrProdScalVecCpx */
      r.re = kD.re * kw.re - kD.im * kw.im;
      r.im = kD.re * kw.im + kD.im * kw.re;
      long long o = ((long long)k * ntheta + l) * 2;
      kWD[o] = r.re; kWD[o + 1] = r.im;
      e1[o] = w.re / wnorm[l]; e1[o + 1] = -(w.im / wnorm[l]);   /* fillkWNorm */
      ocpx s3 = get_camp(e3spec + l * cnzp, nzp, k);
      e2[o] = s3.re / e3norm[l]; e2[o + 1] = -(s3.im / e3norm[l]);
      e3[o] = 1.0f; e3[o + 1] = 0.0f;                      /* lib_matrFillOnesVecCpx */
    }
  }
  free(d1); free(dspec); free(wspec); free(e3spec); free(e3norm); free(fw); free(bw); free(cw);
}

/* ========================================================================
 * Per-k tables for a simbox that is not of constant thickness
 * (src/crava.cpp:1058-1083): the data were divided by the wavelet first, so
 * K = D A, the first error multiplier is 1/|w|, the second the spectrum of
 * the third-difference operator over |errorSmooth3|, and the third the
 * robust inverse wavelet amplitude of fillInverseAbskWRobust (:1361-1384).
 * ====================================================================== */

/* Wavelet::getCAmp(k, scale), src/wavelet.cpp:327-385: the stretched
 * wavelet's spectrum, FFT(w(s t)) = W(omega/s)/s, by linear interpolation of
 * the half spectrum at k/scale (zero at or beyond cnzp). */
static ocpx get_camp_scaled(const ocpx *spec, int nzp, int k, float scale)
{
  int cnzp = nzp / 2 + 1;
  int upper = k >= cnzp;
  int refk = upper ? nzp - k : k;
  float omega = (float)refk / scale;
  ocpx v = {0.0f, 0.0f};
  if (!(omega >= (float)cnzp)) {
    int omL = (int)floor(omega);
    int omU = omL + 1;
    if (omU >= cnzp) omU -= 1;
    float dOm = omega - (float)omL;
    v.re = spec[omL].re * (1.0f - dOm) + spec[omU].re * dOm;
    if (!upper) v.im = spec[omL].im * (1.0f - dOm) + spec[omU].im * dOm;
    else v.im = -spec[omL].im * (1.0f - dOm) - spec[omU].im * dOm;
  }
  v.re /= scale;
  v.im /= scale;
  return v;
}

/* forward, backward and central first differences on the periodic nzp
 * series, src/wavelet1D.cpp:431-460 */
static void diff_fwd(const float *x, int n, float *y)
{
  for (int i = 0; i < n; i++) y[i] = (i == n - 1) ? (x[0] - x[i]) : (x[i + 1] - x[i]);
}
static void diff_bwd(const float *x, int n, float *y)
{
  for (int i = 0; i < n; i++) y[i] = (i == 0) ? (x[0] - x[n - 1]) : (x[i] - x[i - 1]);
}
static void diff_cen(const float *x, int n, float *y)
{
  for (int i = 0; i < n; i++) {
    float up = (i == n - 1) ? x[0] : x[i + 1];
    float dn = (i == 0) ? x[n - 1] : x[i - 1];
    y[i] = (float)(0.5 * (up - dn));
  }
}

void orc_perk_tables_nonconst(int ntheta, int nz, int nzp, double lz,
                              float low_cut, float high_cut, float min_rel_thick,
                              const float *wavelets, const float *wnorm,
                              unsigned char *invert, float *kWD, float *e1, float *e2, float *e3)
{
  int cnzp = nzp / 2 + 1;
  float *a = (float *)calloc(nzp, sizeof(float));
  float *b = (float *)malloc(sizeof(float) * nzp);
  float *c = (float *)malloc(sizeof(float) * nzp);
  ocpx *d1spec = (ocpx *)malloc(sizeof(ocpx) * cnzp);
  ocpx *d3spec = (ocpx *)malloc(sizeof(ocpx) * cnzp);
  ocpx *wspec = (ocpx *)malloc(sizeof(ocpx) * cnzp * ntheta);
  float *s3norm = (float *)malloc(sizeof(float) * ntheta);
  /* diff1Operator, diff2Operator, diff3Operator (src/crava.cpp:938-944) */
  a[0] = -1.0f; a[nzp - 1] = 1.0f;
  spectrum_f(a, nzp, d1spec);
  diff_bwd(a, nzp, b);
  diff_cen(b, nzp, c);
  spectrum_f(c, nzp, d3spec);
  for (int l = 0; l < ntheta; l++) {
    const float *w = wavelets + (long long)l * nzp;
    spectrum_f(w, nzp, wspec + l * cnzp);             /* seisWaveletForNorm */
    diff_fwd(w, nzp, a);                              /* errorSmooth  (:958-961) */
    diff_bwd(a, nzp, b);                              /* errorSmooth2 */
    diff_cen(b, nzp, c);                              /* errorSmooth3 */
    s3norm[l] = find_norm(c, nzp);
  }
  for (int k = 0; k < nzp; k++) {
    int kk = k < nzp - k ? k : nzp - k;
    float rf = (float)((nz * 1000.0f) / (lz * nzp) * kk);
    invert[k] = (rf > low_cut * min_rel_thick && rf < high_cut) ? 1 : 0;
    ocpx kD = get_camp(d1spec, nzp, k);
    ocpx kD3 = get_camp(d3spec, nzp, k);
    for (int l = 0; l < ntheta; l++) {
      long long o = ((long long)k * ntheta + l) * 2;
      kWD[o] = kD.re; kWD[o + 1] = kD.im;             /* lib_matrFillValueVecCpx(kD) */
      e1[o] = 1.0f / wnorm[l]; e1[o + 1] = 0.0f;      /* ones / getNorm() */
      float mult = 1.0f / s3norm[l];
      e2[o] = kD3.re * mult; e2[o + 1] = kD3.im * mult;
      /* fillInverseAbskWRobust */
      ocpx v = get_camp(wspec + l * cnzp, nzp, k);
      ocpx vf = get_camp_scaled(wspec + l * cnzp, nzp, k, 0.999f);
      float mod = v.re * v.re + v.im * v.im;
      float modf = vf.re * vf.re + vf.im * vf.im;
      float mx = mod > modf ? mod : modf;
      /* sqrt of a float argument: the float overload under C++'s <math.h> */
      if (mx > 0.0f) e3[o] = (float)(1.0 / (double)sqrtf(mx));
      else e3[o] = wnorm[l] * nzp * nzp * 100.0f;
      e3[o + 1] = 0.0f;
    }
  }
  free(a); free(b); free(c); free(d1spec); free(d3spec); free(wspec); free(s3norm);
}

/* ========================================================================
 * Wavelet::resample (src/wavelet.cpp:416-463) with getWaveletValue
 * (:1078-1122): simple kriging of the six samples around z under a 20 Hz
 * Ricker correlation, solved by Cholesky (NRLib::CholeskySolve, LAPACK posv
 * on the upper triangle).  Returns 0, or -1 where the kriging matrix is not
 * positive definite (the reference throws).
 * ====================================================================== */
static double ricker20(double lag_ms)
{
  double dt = lag_ms * 0.001;
  double f = 20.0 * 20.0 * M_PI * M_PI * dt * dt;
  return (1.0 - 2.0 * f) * exp(-f);
}

static int krige6(float z, const float *w, int center, int nw, float dzw, float *value)
{
  int ind[6];
  double val[6], rhs[6], U[6][6];
  ind[2] = (int)floor(z / dzw);
  for (int k = 0; k < 6; k++) {
    ind[k] = ind[2] + k - 2;
    int i = ind[k] + center;
    val[k] = (i > -1 && i < nw) ? (double)w[i] : 0.0;   /* getArrayValueOrZero */
  }
  for (int k = 0; k < 6; k++) {
    rhs[k] = ricker20((double)(dzw * ind[k] - z));
    for (int l = 0; l <= k; l++) U[l][k] = ricker20((double)(dzw * ind[k] - dzw * ind[l]));
  }
  /* A = U^T U, U upper triangular (posv with uplo = U) */
  for (int j = 0; j < 6; j++) {
    double s = U[j][j];
    for (int i = 0; i < j; i++) s -= U[i][j] * U[i][j];
    if (!(s > 0.0)) return -1;
    U[j][j] = sqrt(s);
    for (int k = j + 1; k < 6; k++) {
      double t = U[j][k];
      for (int i = 0; i < j; i++) t -= U[i][j] * U[i][k];
      U[j][k] = t / U[j][j];
    }
  }
  double y[6], x[6];
  for (int i = 0; i < 6; i++) {               /* U^T y = rhs */
    double s = rhs[i];
    for (int k = 0; k < i; k++) s -= U[k][i] * y[k];
    y[i] = s / U[i][i];
  }
  for (int i = 5; i >= 0; i--) {              /* U x = y */
    double s = y[i];
    for (int k = i + 1; k < 6; k++) s -= U[i][k] * x[k];
    x[i] = s / U[i][i];
  }
  double v = 0.0;
  for (int k = 0; k < 6; k++) v += val[k] * x[k];
  *value = (float)v;
  return 0;
}

int orc_wavelet_resample(const float *w, int nw, int center, float dzw, float dz, int nzp,
                         float *out, float *norm)
{
  for (int k = 0; k < nzp; k++) {
    float z = (k < nzp / 2 + 1) ? (float)(dz * k) : (float)(dz * (k - nzp));
    if (krige6(z, w, center, nw, dzw, out + k) != 0) return -1;
  }
  double n2 = 0.0;
  for (int k = 0; k < nzp; k++) n2 += (double)(out[k] * out[k]);
  *norm = (float)sqrt(n2);
  return 0;
}

/* ========================================================================
 * Crava::computeVariances (src/crava.cpp:400-466) with computeWDCorrMVar
 * (:1763-1775): per angle, paramVar = A_l Sigma0 A_l^T, the variance of the
 * differenced wavelet under the temporal correlation, the model and signal
 * variances, the energy-matching gain and the theoretical S/N.  Float
 * arithmetic in the reference's order.
 * ====================================================================== */
void orc_wavelet_energy(int ntheta, int nzp, const float *wavelets, const float *A,
                        const double *prior_var0, const float *corr_t, const float *error_variance,
                        const float *emp_sn, const unsigned char *match,
                        float *gain, float *model_var, float *signal_var, float *theo_sn)
{
  float *wd = (float *)malloc(sizeof(float) * nzp);
  for (int l = 0; l < ntheta; l++) {
    diff_fwd(wavelets + (long long)l * nzp, nzp, wd);
    float pv = 0.0f;
    for (int i = 0; i < 3; i++)
      for (int m = 0; m < 3; m++)
        pv += (float)prior_var0[i * 3 + m] * A[l * 3 + i] * A[l * 3 + m];
    float wv = 0.0f;
    for (int i = 0; i < nzp; i++)
      for (int j = 0; j < nzp; j++) {
        int ci = i > j ? i - j : j - i;
        wv += wd[i] * corr_t[ci] * wd[j];
      }
    float mv = wv * pv;
    float sv = error_variance[l] + mv;
    float g = 1.0f;
    if (match && match[l]) {
      g = sqrtf((error_variance[l] / mv) * (emp_sn[l] - 1.0f));
      mv *= g * g;
      sv = error_variance[l] + mv;
    }
    gain[l] = g; model_var[l] = mv; signal_var[l] = sv;
    theo_sn[l] = sv / error_variance[l];
  }
  free(wd);
}

/* ========================================================================
 * Per-point posterior loop, src/crava.cpp:1042-1160 (constant thickness),
 * with getNextParameterCovariance (src/seismicparametersholder.cpp:356-427)
 * and getNextErrorVariance (src/crava.cpp:1298-1334).
 *
 * Grids are complex [nzp][nyp][cnxp]; cov order: aa, bb, rr, ab, ar, br.
 * A is [ntheta][3] float, err_theta_cov is [ntheta][ntheta] double,
 * sigma0 is the 3x3 prior covariance (double, row-major).
 * Returns the number of in-band points whose Cholesky failed.
 * ====================================================================== */

/* SeismicParametersHolder::getNextParameterCovariance */
static void o_param_cov(const double *s0, const ocpx *c6, ocpx pv[3][3])
{
  ocpx t[6];
  for (int q = 0; q < 6; q++) t[q] = c6[q];
  static const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  float v[6];
  for (int q = 0; q < 6; q++) {
    double s = s0[ia[q] * 3 + ib[q]];
    if (s != 0) t[q].re = t[q].re / (float)s;
    v[q] = (float)(sqrt(t[q].re * t[q].re));
  }
  for (int q = 0; q < 6; q++) {
    int a = ia[q], b = ib[q];
    float s = (float)s0[a * 3 + b];
    pv[a][b].re = v[q] * s; pv[a][b].im = 0.0f;
    if (a != b) { pv[b][a].re = v[q] * s; pv[b][a].im = -0.0f; }
  }
}

long long orc_posterior(int nxp, int nyp, int nzp, int ntheta,
                        const float *A, const double *err_theta_cov, float wnc,
                        const double *sigma0,
                        const unsigned char *invert, const float *kWD,
                        const float *e1, const float *e2, const float *e3,
                        float **mu, float **data, float **cov, const float *errcorr)
{
  int cnxp = nxp / 2 + 1;
  int n = ntheta;
  long long nfail = 0;
  ocpx Kb[ORC_MAXT][3], KSb[ORC_MAXT][3], KSccb[3][ORC_MAXT], pvb[3][3];
  ocpx mvb[ORC_MAXT][ORC_MAXT], evb[ORC_MAXT][ORC_MAXT], rvb[3][3];
  ocpx *K[ORC_MAXT], *KS[ORC_MAXT], *KScc[3], *pv[3], *mv[ORC_MAXT], *ev[ORC_MAXT], *rv[3];
  for (int i = 0; i < ORC_MAXT; i++) { K[i] = Kb[i]; KS[i] = KSb[i]; mv[i] = mvb[i]; ev[i] = evb[i]; }
  for (int i = 0; i < 3; i++) { KScc[i] = KSccb[i]; pv[i] = pvb[i]; rv[i] = rvb[i]; }
  ocpx m1[ORC_MAXT], m2[ORC_MAXT], m3[ORC_MAXT];
  ocpx ijkData[ORC_MAXT], ijkDataMean[ORC_MAXT], ijkRes[ORC_MAXT], ijkMean[3], ijkAns[3];
  ocpx **mu_c = (ocpx **)mu, **d_c = (ocpx **)data, **cv_c = (ocpx **)cov;
  const ocpx *ec = (const ocpx *)errcorr;
  for (int k = 0; k < nzp; k++) {
    for (int l = 0; l < n; l++) {
      long long o = ((long long)k * n + l) * 2;
      for (int m = 0; m < 3; m++) {                /* lib_matrProdDiagCpxR */
        K[l][m].re = kWD[o] * A[l * 3 + m];
        K[l][m].im = kWD[o + 1] * A[l * 3 + m];
      }
      m1[l].re = e1[o]; m1[l].im = e1[o + 1];
      m2[l].re = e2[o]; m2[l].im = e2[o + 1];
      m3[l].re = e3[o]; m3[l].im = e3[o + 1];
    }
    int inv = invert[k];
    for (int j = 0; j < nyp; j++)
      for (int i = 0; i < cnxp; i++) {
        long long p = ((long long)k * nyp + j) * cnxp + i;
        for (int a = 0; a < 3; a++) ijkMean[a] = mu_c[a][p];
        for (int l = 0; l < n; l++) { ijkData[l] = d_c[l][p]; ijkRes[l] = ijkData[l]; }
        ocpx c6[6];
        for (int q = 0; q < 6; q++) c6[q] = cv_c[q][p];
        o_param_cov(sigma0, c6, pvb);
        /* getNextErrorVariance */
        ocpx tmp = ec[p];
        float lam = (float)(sqrt(tmp.re * tmp.re));
        if (inv) {
          for (int l = 0; l < n; l++)
            for (int m = 0; m < n; m++) {
              ev[l][m].re = (float)(0.5 * (1.0 - wnc) * err_theta_cov[l * n + m] * lam *
                                    (m1[l].re * m1[m].re + m1[l].im * m1[m].im));
              ev[l][m].re += (float)(0.5 * (1.0 - wnc) * err_theta_cov[l * n + m] * lam *
                                     (m2[l].re * m2[m].re + m2[l].im * m2[m].im));
              if (l == m) {
                ev[l][m].re += (float)(wnc * err_theta_cov[l * n + m] * m3[l].re * m3[l].re);
                ev[l][m].im = 0.0f;
              } else {
                ev[l][m].im = (float)(0.5 * (1.0 - wnc) * err_theta_cov[l * n + m] * lam *
                                      (-m1[l].re * m1[m].im + m1[l].im * m1[m].re));
                ev[l][m].im += (float)(0.5 * (1.0 - wnc) * err_theta_cov[l * n + m] * lam *
                                       (-m2[l].re * m2[m].im + m2[l].im * m2[m].re));
              }
            }
          o_prod_cpx(K, pv, n, 3, 3, KS);
          o_prod_adjoint_cpx(KS, K, n, 3, n, mv);
          for (int l = 0; l < n; l++)                  /* lib_matrAddMatCpx */
            for (int m = 0; m < n; m++) { mv[l][m].re += ev[l][m].re; mv[l][m].im += ev[l][m].im; }
          int flag = o_chol_cpx(n, mv);
          if (flag == 0) {
            for (int a = 0; a < n; a++)               /* lib_matrAdjoint */
              for (int b = 0; b < 3; b++) { KScc[b][a].re = KS[a][b].re; KScc[b][a].im = -KS[a][b].im; }
            o_axeqb_mat_cpx(n, mv, KS, 3);
            o_prod_cpx(KScc, KS, 3, n, 3, rv);
            for (int a = 0; a < 3; a++)               /* lib_matrSubtMatCpx */
              for (int b = 0; b < 3; b++) { pv[a][b].re -= rv[a][b].re; pv[a][b].im -= rv[a][b].im; }
            o_prod_matvec_cpx(K, ijkMean, n, 3, ijkDataMean);
            for (int l = 0; l < n; l++) { ijkData[l].re -= ijkDataMean[l].re; ijkData[l].im -= ijkDataMean[l].im; }
            o_prod_adjoint_matvec_cpx(KS, ijkData, 3, n, ijkAns);
            for (int a = 0; a < 3; a++) { ijkMean[a].re += ijkAns[a].re; ijkMean[a].im += ijkAns[a].im; }
            o_prod_matvec_cpx(K, ijkMean, n, 3, ijkData);
            for (int l = 0; l < n; l++) { ijkRes[l].re -= ijkData[l].re; ijkRes[l].im -= ijkData[l].im; }
          } else {
            nfail++;
          }
        }
        for (int a = 0; a < 3; a++) mu_c[a][p] = ijkMean[a];
        cv_c[0][p] = pv[0][0]; cv_c[1][p] = pv[1][1]; cv_c[2][p] = pv[2][2];
        cv_c[3][p] = pv[0][1]; cv_c[4][p] = pv[0][2]; cv_c[5][p] = pv[1][2];
        for (int l = 0; l < n; l++) d_c[l][p] = ijkRes[l];
      }
  }
  return nfail;
}

/* ========================================================================
 * FP64 oracle of the same per-point mathematics (independent of the float
 * roundings): complex double, Cholesky without pre-normalisation, same
 * failure rule (pivot <= 1e-20 relative to M00) and pass-through semantics.
 * Tables are the float tables; grids are complex double.
 * ====================================================================== */
typedef struct { double re, im; } zc;

static int z_chol(int n, zc M[ORC_MAXT][ORC_MAXT])
{
  double f = M[0][0].re;
  if (f <= 0) return 1;
  for (int i = 0; i < n; i++) {
    for (int j = 0; j < i; j++) {
      double sr = M[i][j].re, si = M[i][j].im;
      for (int k = 0; k < j; k++) {   /* s -= L_ik conj(L_jk) */
        sr -= M[i][k].re * M[j][k].re + M[i][k].im * M[j][k].im;
        si -= M[i][k].im * M[j][k].re - M[i][k].re * M[j][k].im;
      }
      M[i][j].re = sr / M[j][j].re;
      M[i][j].im = si / M[j][j].re;
    }
    double d = M[i][i].re;
    for (int k = 0; k < i; k++) d -= M[i][k].re * M[i][k].re + M[i][k].im * M[i][k].im;
    if (d <= ORC_TOL * f) return 1;
    M[i][i].re = sqrt(d); M[i][i].im = 0.0;
  }
  return 0;
}

long long orc_posterior_f64(int nxp, int nyp, int nzp, int ntheta,
                            const float *A, const double *err_theta_cov, double wnc,
                            const double *sigma0,
                            const unsigned char *invert, const float *kWD,
                            const float *e1, const float *e2, const float *e3,
                            double **mu, double **data, double **cov, const double *errcorr)
{
  int cnxp = nxp / 2 + 1, n = ntheta;
  long long nfail = 0;
  static const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  for (int k = 0; k < nzp; k++) {
    zc K[ORC_MAXT][3], m1[ORC_MAXT], m2[ORC_MAXT], m3[ORC_MAXT];
    for (int l = 0; l < n; l++) {
      long long o = ((long long)k * n + l) * 2;
      for (int m = 0; m < 3; m++) { K[l][m].re = (double)kWD[o] * A[l * 3 + m]; K[l][m].im = (double)kWD[o + 1] * A[l * 3 + m]; }
      m1[l].re = e1[o]; m1[l].im = e1[o + 1];
      m2[l].re = e2[o]; m2[l].im = e2[o + 1];
      m3[l].re = e3[o]; m3[l].im = e3[o + 1];
    }
    for (int j = 0; j < nyp; j++)
      for (int i = 0; i < cnxp; i++) {
        long long p = ((long long)k * nyp + j) * cnxp + i;
        double S[3][3];
        for (int q = 0; q < 6; q++) {
          double s = sigma0[ia[q] * 3 + ib[q]];
          double v = cov[q][2 * p];
          if (s != 0) v = fabs(v / s) * s; else v = fabs(v) * s;
          S[ia[q]][ib[q]] = v; S[ib[q]][ia[q]] = v;
        }
        if (!invert[k]) {
          for (int q = 0; q < 6; q++) { cov[q][2 * p] = S[ia[q]][ib[q]]; cov[q][2 * p + 1] = 0.0; }
          continue;
        }
        double lam = fabs(errcorr[2 * p]);
        zc M[ORC_MAXT][ORC_MAXT], KS[ORC_MAXT][3];
        for (int l = 0; l < n; l++)
          for (int b = 0; b < 3; b++) {
            double sr = 0, si = 0;
            for (int a = 0; a < 3; a++) { sr += K[l][a].re * S[a][b]; si += K[l][a].im * S[a][b]; }
            KS[l][b].re = sr; KS[l][b].im = si;
          }
        for (int l = 0; l < n; l++)
          for (int m = 0; m < n; m++) {
            double sr = 0, si = 0;
            for (int b = 0; b < 3; b++) {   /* KS_lb conj(K_mb) */
              sr += KS[l][b].re * K[m][b].re + KS[l][b].im * K[m][b].im;
              si += KS[l][b].im * K[m][b].re - KS[l][b].re * K[m][b].im;
            }
            double c = 0.5 * (1.0 - wnc) * err_theta_cov[l * n + m] * lam;
            sr += c * (m1[l].re * m1[m].re + m1[l].im * m1[m].im + m2[l].re * m2[m].re + m2[l].im * m2[m].im);
            if (l == m) sr += wnc * err_theta_cov[l * n + m] * m3[l].re * m3[l].re;
            else si += c * (-m1[l].re * m1[m].im + m1[l].im * m1[m].re - m2[l].re * m2[m].im + m2[l].im * m2[m].re);
            M[l][m].re = sr; M[l][m].im = (l == m) ? 0.0 : si;
          }
        if (z_chol(n, M)) {
          nfail++;
          for (int q = 0; q < 6; q++) { cov[q][2 * p] = S[ia[q]][ib[q]]; cov[q][2 * p + 1] = 0.0; }
          continue;
        }
        zc X[ORC_MAXT][3];
        for (int b = 0; b < 3; b++) {
          for (int l = 0; l < n; l++) {          /* L y = ks */
            double sr = KS[l][b].re, si = KS[l][b].im;
            for (int m = 0; m < l; m++) {
              sr -= M[l][m].re * X[m][b].re - M[l][m].im * X[m][b].im;
              si -= M[l][m].re * X[m][b].im + M[l][m].im * X[m][b].re;
            }
            X[l][b].re = sr / M[l][l].re; X[l][b].im = si / M[l][l].re;
          }
          for (int l = n - 1; l >= 0; l--) {     /* L^H x = y */
            double sr = X[l][b].re, si = X[l][b].im;
            for (int m = l + 1; m < n; m++) {  /* conj(L_ml) x_m */
              sr -= M[m][l].re * X[m][b].re + M[m][l].im * X[m][b].im;
              si -= M[m][l].re * X[m][b].im - M[m][l].im * X[m][b].re;
            }
            X[l][b].re = sr / M[l][l].re; X[l][b].im = si / M[l][l].re;
          }
        }
        /* Sigma_post = S - KS^H X */
        zc P[3][3];
        for (int a = 0; a < 3; a++)
          for (int b = 0; b < 3; b++) {
            double sr = 0, si = 0;
            for (int l = 0; l < n; l++) {     /* conj(KS_la) X_lb */
              sr += KS[l][a].re * X[l][b].re + KS[l][a].im * X[l][b].im;
              si += KS[l][a].re * X[l][b].im - KS[l][a].im * X[l][b].re;
            }
            P[a][b].re = S[a][b] - sr; P[a][b].im = -si;
          }
        for (int q = 0; q < 6; q++) { cov[q][2 * p] = P[ia[q]][ib[q]].re; cov[q][2 * p + 1] = P[ia[q]][ib[q]].im; }
        /* mean: mu += X^H (d - K mu); residual d - K mu_post */
        zc m0[3], d[ORC_MAXT], r0[ORC_MAXT];
        for (int a = 0; a < 3; a++) { m0[a].re = mu[a][2 * p]; m0[a].im = mu[a][2 * p + 1]; }
        for (int l = 0; l < n; l++) {
          d[l].re = data[l][2 * p]; d[l].im = data[l][2 * p + 1];
          double sr = 0, si = 0;
          for (int a = 0; a < 3; a++) { sr += K[l][a].re * m0[a].re - K[l][a].im * m0[a].im; si += K[l][a].re * m0[a].im + K[l][a].im * m0[a].re; }
          r0[l].re = d[l].re - sr; r0[l].im = d[l].im - si;
        }
        for (int a = 0; a < 3; a++) {
          double sr = 0, si = 0;
          for (int l = 0; l < n; l++) { sr += X[l][a].re * r0[l].re + X[l][a].im * r0[l].im; si += X[l][a].re * r0[l].im - X[l][a].im * r0[l].re; }
          m0[a].re += sr; m0[a].im += si;
          mu[a][2 * p] = m0[a].re; mu[a][2 * p + 1] = m0[a].im;
        }
        for (int l = 0; l < n; l++) {
          double sr = 0, si = 0;
          for (int a = 0; a < 3; a++) { sr += K[l][a].re * m0[a].re - K[l][a].im * m0[a].im; si += K[l][a].re * m0[a].im + K[l][a].im * m0[a].re; }
          data[l][2 * p] = d[l].re - sr; data[l][2 * p + 1] = d[l].im - si;
        }
      }
  }
  return nfail;
}

/* ========================================================================
 * RandomGen (libs/lib/random.cpp:38-119, random.h:11-14) and
 * FFTGrid::fillInComplexNoise (src/fftgrid.cpp:1214-1263).
 * ====================================================================== */
static double o_unif01(uint32_t *seed)
{
  *seed = 69069u * (*seed) + 1u;
  return (double)(*seed) * ((1.0 / (double)(256 * 256 * 256 * 64)) / 4.0);
}

static double o_g(double x)
{
  double a = 17.49731196, b = 2.36785163, c = 2.15787544;
  double ax = fabs(x);
  double r = a * exp(-x * x / 2.0);
  if (ax < 1.0) r -= 2.0 * b * (3.0 - x * x) + c * (1.5 - ax);
  else if (ax < 1.5) r -= b * (3.0 - ax) * (3.0 - ax) + c * (1.5 - ax);
  else r -= b * (3.0 - ax) * (3.0 - ax);
  return r;
}

double orc_rnorm01(uint32_t *seed)
{
  double u, u1, u2, u3, c, x, v1, v2, w, s, t;
  u = o_unif01(seed);
  if (u < 0.8638) {
    u1 = o_unif01(seed); u2 = o_unif01(seed); u3 = o_unif01(seed);
    x = 2.0 * (u1 + u2 + u3) - 3.0;
  } else if (u < 0.9745) {
    u1 = o_unif01(seed); u2 = o_unif01(seed);
    x = 1.5 * (u1 + u2 - 1.0);
  } else if (u < 0.9973002039) {
    do {
      u1 = o_unif01(seed); u2 = o_unif01(seed);
      x = 6.0 * u1 - 3.0;
    } while (0.358 * u2 > o_g(x));
  } else {
    do {
      do {
        u1 = o_unif01(seed); u2 = o_unif01(seed);
        v1 = 2.0 * u1 - 1.0; v2 = 2.0 * u2 - 1.0;
        w = v1 * v1 + v2 * v2;
      } while (w >= 1.0);
      c = sqrt((9.0 - 2.0 * log(w)) / w);
      s = c * v1; t = c * v2;
    } while ((fabs(s) <= 3.0) && (fabs(t) <= 3.0));
    x = (fabs(s) > 3.0) ? s : t;
  }
  return x;
}

double orc_unif01(uint32_t *seed) { return o_unif01(seed); }

void orc_fill_complex_noise(float *grid, int nxp, int nyp, int nzp, uint32_t *seed)
{
  int cnxp = nxp / 2 + 1;
  long long csize = (long long)cnxp * nyp * nzp;
  ocpx *cv = (ocpx *)grid;
  float std = (float)(1 / sqrt(2.0));
  for (long long i = 0; i < csize; i++) {
    if (((i % cnxp) == 0) || (((i % cnxp) == cnxp - 1) && ((i % 2) == 1))) {
      long long xshift = i % cnxp;
      long long jk = (i - xshift) / cnxp;
      long long jind = jk % nyp, kind = jk / nyp;
      long long jcc = jind == 0 ? 0 : nyp - jind;
      long long kcc = kind == 0 ? 0 : nzp - kind;
      long long jkcc = jcc + kcc * nyp;
      if (jkcc == jk) {
        cv[i].re = (float)(orc_rnorm01(seed));
        cv[i].im = 0;
      } else if (jkcc > jk) {
        cv[i].re = (float)(std * orc_rnorm01(seed));
        cv[i].im = (float)(std * orc_rnorm01(seed));
      } else {
        long long cci = jkcc * cnxp + xshift;
        cv[i].re = cv[cci].re;
        cv[i].im = -cv[cci].im;
      }
    } else {
      cv[i].re = (float)(std * orc_rnorm01(seed));
      cv[i].im = (float)(std * orc_rnorm01(seed));
    }
  }
}

/* Crava::simulate per-point loop, src/crava.cpp:1482-1521: seed <- chol(Sigma_post) seed,
 * zero on failure.  Returns the number of failed factorisations. */
long long orc_simulate_points(long long npts, float **cov, float **seed)
{
  long long nfail = 0;
  ocpx pcb[3][3];
  ocpx *pc[3] = {pcb[0], pcb[1], pcb[2]};
  ocpx **cv = (ocpx **)cov, **sd = (ocpx **)seed;
  for (long long p = 0; p < npts; p++) {
    pc[0][0] = cv[0][p]; pc[1][1] = cv[1][p]; pc[2][2] = cv[2][p];
    pc[0][1] = cv[3][p]; pc[0][2] = cv[4][p]; pc[1][2] = cv[5][p];
    pc[1][0].re = pc[0][1].re; pc[1][0].im = -pc[0][1].im;
    pc[2][0].re = pc[0][2].re; pc[2][0].im = -pc[0][2].im;
    pc[2][1].re = pc[1][2].re; pc[2][1].im = -pc[1][2].im;
    ocpx s[3] = {sd[0][p], sd[1][p], sd[2][p]};
    if (o_chol_cpx(3, pc) == 0) {
      o_prod_chol_vec(3, pc, s);
    } else {
      nfail++;
      for (int l = 0; l < 3; l++) { s[l].re = 0.0f; s[l].im = 0.0f; }
    }
    sd[0][p] = s[0]; sd[1][p] = s[1]; sd[2][p] = s[2];
  }
  return nfail;
}

/* ========================================================================
 * Facies probabilities, src/faciesprob.cpp:2198-2265 with
 * FindDensityFromPosteriorPDF (:1277-1317, one density dimension, no angle
 * weights) and PosteriorElasticPDF3D::FindDensity
 * (src/posteriorelasticpdf3d.cpp:511-604) on an unrotated box volume
 * (Simbox::getInterpolationIndexes, src/simbox.cpp:186-201).
 *
 * tables: nfac tables, each an FFTGrid(n1,n2,n3,n1,n2,n3) real array
 * [n3][n2][2*(n1/2+1)].  vol = {x0, y0, dx, dy, ztop, zbot} (double),
 * volume cell counts vn = {nx, ny, nz}.
 * Inputs alpha/beta/rho are padded real grids [nzp][nyp][rnxp]; outputs are
 * nfac+1 unpadded-layout real grids [nz][ny][2*(nx/2+1)] (last = undefined).
 * ====================================================================== */
static double o_find_density(double vp, double vs, double rho, const float *tab,
                             int tn1, int tn2, const double *vol, const int *vn)
{
  double value = 0.0;
  double rx = (vp - vol[0]) * 1.0 + (vs - vol[1]) * 0.0;
  double ry = -(vp - vol[0]) * 0.0 + (vs - vol[1]) * 1.0;
  double jF = rx / vol[2] - 0.5;
  double kF = ry / vol[3] - 0.5;
  double lF = (double)(vn[2]) * (rho - vol[4]) / (vol[5] - vol[4]) - 0.5;
  int j1, k1, l1, j2, k2, l2;
  float wj, wk, wl;
  j1 = (int)floor(jF);
  if (j1 < 0) { j1 = 0; j2 = 0; wj = 0; }
  else if (j1 >= vn[0] - 1) { j1 = vn[0] - 1; j2 = j1; wj = 0; }
  else { j2 = j1 + 1; wj = (float)(jF - j1); }
  k1 = (int)floor(kF);
  if (k1 < 0) { k1 = 0; k2 = 0; wk = 0; }
  else if (k1 >= vn[1] - 1) { k1 = vn[1] - 1; k2 = k1; wk = 0; }
  else { k2 = k1 + 1; wk = (float)(kF - k1); }
  l1 = (int)floor(lF);
  if (l1 < 0) { l1 = 0; l2 = 0; wl = 0; }
  else if (l1 >= vn[2] - 1) { l1 = vn[2] - 1; l2 = l1; wl = 0; }
  else { l2 = l1 + 1; wl = (float)(lF - l1); }
  int rn = 2 * (tn1 / 2 + 1);
#define TV(a, b, c) (tab[(a) + (long long)rn * (b) + (long long)rn * tn2 * (c)])
  float v1 = TV(j1, k1, l1) > 0 ? TV(j1, k1, l1) : 0.0f;
  float v2 = TV(j1, k1, l2) > 0 ? TV(j1, k1, l2) : 0.0f;
  float v3 = TV(j1, k2, l1) > 0 ? TV(j1, k2, l1) : 0.0f;
  float v4 = TV(j1, k2, l2) > 0 ? TV(j1, k2, l2) : 0.0f;
  float v5 = TV(j2, k1, l1) > 0 ? TV(j2, k1, l1) : 0.0f;
  float v6 = TV(j2, k1, l2) > 0 ? TV(j2, k1, l2) : 0.0f;
  float v7 = TV(j2, k2, l1) > 0 ? TV(j2, k2, l1) : 0.0f;
  float v8 = TV(j2, k2, l2) > 0 ? TV(j2, k2, l2) : 0.0f;
#undef TV
  value += (1.0f - wj) * (1.0f - wk) * (1.0f - wl) * v1;
  value += (1.0f - wj) * (1.0f - wk) * (wl) * v2;
  value += (1.0f - wj) * (wk) * (1.0f - wl) * v3;
  value += (1.0f - wj) * (wk) * (wl) * v4;
  value += (wj) * (1.0f - wk) * (1.0f - wl) * v5;
  value += (wj) * (1.0f - wk) * (wl) * v6;
  value += (wj) * (wk) * (1.0f - wl) * v7;
  value += (wj) * (wk) * (wl) * v8;
  return value;
}

/* One density lookup (PosteriorElasticPDF3D::FindDensity), exported so the
 * tests can check it against hand-computed cases and a brute-force lookup. */
double orc_facies_find_density(double vp, double vs, double rho, const float *tab,
                               int tn1, int tn2, const double *vol, const int *vn)
{
  return o_find_density(vp, vs, rho, tab, tn1, tn2, vol, vn);
}

void orc_facies(int nx, int ny, int nz, int nxp, int nyp, int nzp,
                const float *alpha, const float *beta, const float *rho,
                int nfac, const float *tables, int tn1, int tn2, int tn3,
                const double *vol, const int *vn, const float *prior_facies,
                float p_undefined, float **prob)
{
  int rnxp = 2 * (nxp / 2 + 1);
  int smallrnxp = 2 * (nx / 2 + 1);
  long long tsize = (long long)2 * (tn1 / 2 + 1) * tn2 * tn3;
  float undefSum = p_undefined / (vn[0] * vn[1] * vn[2]);
  float value[64];
  long long in = 0, out = 0;
  for (int i = 0; i < nzp; i++)
    for (int j = 0; j < nyp; j++)
      for (int k = 0; k < rnxp; k++, in++) {
        float a = alpha[in], b = beta[in], r = rho[in];
        if (k < smallrnxp && j < ny && i < nz) {
          float sum = undefSum;
          for (int l = 0; l < nfac; l++) {
            float dens;
            if (k < nx) {
              float v = (float)o_find_density(a, b, r, tables + l * tsize, tn1, tn2, vol, vn);
              dens = v > 0.0f ? v : 0.0f;
            } else {
              dens = 1.0;
            }
            value[l] = prior_facies[l] * dens;
            sum = sum + value[l];
          }
          for (int l = 0; l < nfac; l++) {
            float help = value[l] / sum;
            prob[l][out] = (k < nx) ? help : -99999.0f;   /* RMISSING */
          }
          prob[nfac][out] = (k < nx) ? undefSum / sum : -99999.0f;
          out++;
        }
      }
}

/* CalculateFaciesProbFromPosteriorElasticPDF with the local-noise density
 * dimension (src/faciesprob.cpp:2153-2265, FindDensityFromPosteriorPDF
 * :1277-1317): ndens = 2^nang density sets (tables [ndens][nfac], lattices
 * vols [ndens][6] / vns [ndens][3]); the angle weights from the noise-scale
 * tables noise [nang][ny*nx] (min / max skipping RMISSING, 0 when constant),
 * with one weight per angle (the reference's own vectors are sized before the
 * angles are counted). */
void orc_facies_densdim(int nx, int ny, int nz, int nxp, int nyp, int nzp,
                        const float *alpha, const float *beta, const float *rho,
                        int nfac, int nang, const float *tables, int tn1, int tn2, int tn3,
                        const double *vols, const int *vns, const double *noise,
                        const float *prior_facies, float p_undefined, float **prob)
{
  int ndens = 1 << nang;
  int rnxp = 2 * (nxp / 2 + 1);
  int smallrnxp = 2 * (nx / 2 + 1);
  long long tsize = (long long)2 * (tn1 / 2 + 1) * tn2 * tn3;
  long long ntr = (long long)nx * ny;
  float undefSum = p_undefined / (vns[0] * vns[1] * vns[2]);
  double *tg = (double *)malloc(sizeof(double) * (nang > 0 ? nang : 1) * ntr);
  for (int g = 0; g < nang; g++) {
    const double *tab = noise + g * ntr;
    double mn = tab[0], mx = tab[0];
    for (long long t = 0; t < ntr; t++) {            /* Grid2D::FindMin / FindMax(RMISSING) */
      if ((mn == -99999.0 || tab[t] < mn) && tab[t] != -99999.0) mn = tab[t];
      if ((mx == -99999.0 || tab[t] > mx) && tab[t] != -99999.0) mx = tab[t];
    }
    for (long long t = 0; t < ntr; t++) tg[g * ntr + t] = (mn == mx) ? 0.0 : (tab[t] - mn) / (mx - mn);
  }
  float value[64], t[8], dv[256];
  long long in = 0, out = 0;
  for (int i = 0; i < nzp; i++)
    for (int j = 0; j < nyp; j++)
      for (int k = 0; k < rnxp; k++, in++) {
        float a = alpha[in], b = beta[in], r = rho[in];
        if (k < smallrnxp && j < ny && i < nz) {
          float sum = undefSum;
          for (int l = 0; l < nfac; l++) {
            float dens;
            if (k < nx) {
              for (int g = 0; g < nang; g++) t[g] = (float)tg[g * ntr + (long long)j * nx + k];
              for (int d = 0; d < ndens; d++)
                dv[d] = (float)o_find_density(a, b, r, tables + ((long long)d * nfac + l) * tsize, tn1, tn2,
                                              vols + 6 * d, vns + 3 * d);
              float valuesum = 0;
              for (int d = 0; d < ndens; d++) {
                int factor = 1;
                for (int g = 0; g < nang; g++) {
                  if (g > 0) factor *= 2;
                  if ((d & factor) > 0) dv[d] *= t[g];
                  else dv[d] *= (1 - t[g]);
                }
                valuesum += dv[d];
              }
              dens = valuesum > 0.0f ? valuesum : 0.0f;
            } else {
              dens = 1.0;
            }
            value[l] = prior_facies[l] * dens;
            sum = sum + value[l];
          }
          for (int l = 0; l < nfac; l++) {
            float help = value[l] / sum;
            prob[l][out] = (k < nx) ? help : -99999.0f;
          }
          prob[nfac][out] = (k < nx) ? undefSum / sum : -99999.0f;
          out++;
        }
      }
  free(tg);
}

/* ========================================================================
 * Grid fills and operators (the callers either side of the transforms).
 * ====================================================================== */

/* FFTGrid::fillInParamCorr (src/fftgrid.cpp:1119-1164), then
 * multiplyByScalar(scale) over all rsize values (src/fftgrid.cpp:1992-1999,
 * as SeismicParametersHolder::initializeCorrelations applies it). */
void orc_fill_param_corr(float *grid, int nxp, int nyp, int nzp, const float *corr_xy,
                         const float *circ_t, float grad_i, float grad_j, float scale)
{
  int rnxp = 2 * (nxp / 2 + 1);
  long long p = 0;
  for (int k = 0; k < nzp; k++)
    for (int j = 0; j < nyp; j++)
      for (int i = 0; i < rnxp; i++, p++) {
        float value;
        if (i < nxp) {
          int ci = (i < nxp - i) ? i : i - nxp;   /* signed lateral lag */
          int cj = (j < nyp - j) ? j : j - nyp;
          float t = (float)k + (float)ci * grad_i;
          t = t + (float)cj * grad_j;
          int b = (int)floorf(t);
          float f = t - (float)b;
          b = ((b % nzp) + nzp) % nzp;
          value = (1.0f - f) * circ_t[b];
          value = value + f * circ_t[b == nzp - 1 ? 0 : b + 1];
          value = value * corr_xy[i + nxp * j];
        } else {
          value = -99999.0f;                       /* RMISSING */
        }
        grid[p] = value * scale;
      }
}

/* FFTGrid::fillInErrCorr (src/fftgrid.cpp:1065-1116) with range 1: the
 * lateral correlation lands on the time samples within one sample of the
 * shifted zero lag (cyclically), weighted by 1 - distance. */
void orc_fill_err_corr(float *grid, int nxp, int nyp, int nzp, const float *corr_xy,
                       float grad_i, float grad_j)
{
  int rnxp = 2 * (nxp / 2 + 1);
  long long p = 0;
  for (int k = 0; k < nzp; k++)
    for (int j = 0; j < nyp; j++)
      for (int i = 0; i < rnxp; i++, p++) {
        int ci = (i < nxp - i) ? i : i - nxp;
        int cj = (j < nyp - j) ? j : j - nyp;
        float t = (float)k + (float)ci * grad_i;
        t = t + (float)cj * grad_j;
        float value = 0.0f;
        if (i >= nxp) {
          value = -99999.0f;
        } else if (fabsf(t) < 1.0f || fabsf(t - (float)nzp) < 1.0f) {
          int b = (int)t;                 /* truncation toward zero */
          float f = t - (float)b;
          /* b is 0 or nzp-1 (or -1 / nzp for tiny negative t); reduce to
           * {-1, 0} and weight by the distance to the zero lag */
          while (b < -1) b += nzp;
          while (b >= 1) b -= nzp;
          if (b < 0) f = 1.0f - f;
          value = (1.0f - f) * corr_xy[i + nxp * j];
        }
        grid[p] = value;
      }
}

/* Element-wise operators of FFTGrid (src/fftgrid.cpp:1877-1999) and the
 * copy (op 8); op codes as CRAVA_HIP_OP_*; n floats (real) or n complex
 * values (transformed). */
void orc_grid_op(int op, float *a, const float *b, long long n, int transformed, float s)
{
  for (long long p = 0; p < n; p++) {
    if (transformed) {
      float ar = a[2 * p], ai = a[2 * p + 1];
      float br = b ? b[2 * p] : 0.0f, bi = b ? b[2 * p + 1] : 0.0f;
      switch (op) {
        case 0: ar = ar + br; ai = ai + bi; break;
        case 1: ar = ar - br; ai = ai - bi; break;
        case 2: { float r = br * ar - bi * ai; float im = bi * ar + br * ai; ar = r; ai = im; break; }
        case 3: ar = -ar; ai = -ai; break;
        case 4: ai = -ai; break;
        case 8: ar = br; ai = bi; break;
        default: break;
      }
      a[2 * p] = ar;
      a[2 * p + 1] = ai;
    } else {
      float v = a[p];
      switch (op) {
        case 0: v = v + b[p]; break;
        case 1: v = v - b[p]; break;
        case 2: v = v * b[p]; break;
        case 3: v = -v; break;
        case 5: v = v * s; break;
        case 6: v = v + s; break;
        case 7: v = s; break;
        case 8: v = b[p]; break;
        default: break;
      }
      a[p] = v;
    }
  }
}

/* Crava::computeDataVariance (src/crava.cpp:322-363) for one stack. */
float orc_data_variance(const float *grid, int nx, int ny, int nz, int nxp, int nyp, int nzp,
                        long long *ndata)
{
  int rnxp = 2 * (nxp / 2 + 1);
  double tot = 0.0;
  long long cnt = 0, p = 0;
  for (int k = 0; k < nzp; k++) {
    double plane = 0.0;
    for (int j = 0; j < nyp; j++) {
      double row = 0.0;
      for (int i = 0; i < rnxp; i++, p++) {
        float t = grid[p];
        if (k < nz && j < ny && i < nx && t != 0.0f) {
          row += (double)(t * t);
          cnt++;
        }
      }
      plane += row;
    }
    tot += plane;
  }
  if (ndata) *ndata = cnt;
  return cnt == 0 ? 0.0f : (float)(tot / (double)cnt);
}

/* FaciesProb::calculateFaciesProbGeomodel (src/faciesprob.cpp:2286-2330),
 * constant priors; grids nz x ny x 2*(nx/2+1), prob[nfac] = undefined. */
void orc_facies_geomodel(int nx, int ny, int nz, int nfac, const float *prior, float **prob)
{
  int rnx = 2 * (nx / 2 + 1);
  long long p = 0;
  for (int k = 0; k < nz; k++)
    for (int j = 0; j < ny; j++)
      for (int i = 0; i < rnx; i++, p++) {
        float u = prob[nfac][p];
        for (int l = 0; l < nfac; l++) {
          float v = prior[l] * u + prob[l][p];
          prob[l][p] = (i < nx) ? v : -99999.0f;
        }
      }
}

/* ========================================================================
 * Forward modelling: Crava::computeSyntSeismic (src/crava.cpp:1667-1759)
 * with computeSeismicImpedance (:1632-1663), the Wavelet1D difference
 * operator (src/wavelet1D.cpp:420-436), the local wavelet
 * (Wavelet1D::createLocalWavelet1D / shiftAndScale, src/wavelet1D.cpp:
 * 501-561) and Wavelet::getCAmp(k, scale) (src/wavelet.cpp:328-384).
 * The 1-D transforms are double DFTs rounded to float (the reference runs
 * FFTW 2.1.2 in float); everything else is the reference's float order.
 *   A [nt][3], wavelets [nt][nzp]; shift/gain/sf: NULL or nt pointers, each
 *   NULL or [ny][nx]; alpha/beta/rho/out[nt]: real padded grids.
 * ====================================================================== */

/* c2r of a half spectrum (FFTW's backward sign, unnormalised; the imaginary
 * parts of the DC and even-length Nyquist bins are not read), rounded to
 * float. */
static void c2r_f(const ocpx *spec, int nzp, float *out)
{
  dcpx *x = (dcpx *)malloc(sizeof(dcpx) * nzp);
  dcpx *tmp = (dcpx *)malloc(sizeof(dcpx) * nzp);
  int cn = nzp / 2 + 1;
  for (int k = 0; k < nzp; k++) {
    if (k < cn) { x[k].re = spec[k].re; x[k].im = spec[k].im; }
    else { x[k].re = spec[nzp - k].re; x[k].im = -spec[nzp - k].im; }
  }
  x[0].im = 0.0;
  if (nzp % 2 == 0) x[nzp / 2].im = 0.0;
  dft_rec(x, nzp, +1, tmp);
  for (int t = 0; t < nzp; t++) out[t] = (float)x[t].re;
  free(x); free(tmp);
}

void orc_synthetic_seismic(int nx, int ny, int nz, int nxp, int nyp, int nzp, int nt, const float *A,
                           const float *wavelets, float dz, const float *const *shift,
                           const float *const *gain, const float *const *sf, const float *alpha,
                           const float *beta, const float *rho, float **out)
{
  const int rnxp = 2 * (nxp / 2 + 1), cn = nzp / 2 + 1;
  const long long plane = (long long)rnxp * nyp, rsize = plane * nzp;
  float *imp = (float *)malloc(sizeof(float) * nzp);
  float *r = (float *)malloc(sizeof(float) * nzp);
  ocpx *R = (ocpx *)malloc(sizeof(ocpx) * cn);
  ocpx *W = (ocpx *)malloc(sizeof(ocpx) * cn);
  ocpx *L = (ocpx *)malloc(sizeof(ocpx) * cn);
  for (int l = 0; l < nt; l++) {
    const float *Al = A + 3 * l;
    float *g = out[l];
    /* computeSeismicImpedance: every cell of the padded grid */
    for (long long t = 0; t < rsize; t++) {
      float v = 0;
      v += alpha[t] * Al[0];
      v += beta[t] * Al[1];
      v += rho[t] * Al[2];
      g[t] = v;
    }
    spectrum_f(wavelets + (long long)l * nzp, nzp, W);
    for (int i = 0; i < nx; i++)
      for (int j = 0; j < ny; j++) {
        const long long tr = (long long)j * rnxp + i, ti = (long long)j * nx + i;
        int k;
        for (k = 0; k < nz; k++) imp[k] = g[k * plane + tr];
        float fac = 1.0f / (float)(nzp - nz - 1);                 /* taper, :1692-1696 */
        for (; k < nzp; k++) imp[k] = fac * ((k - nz) * imp[0] + (nzp - k - 1) * imp[nz - 1]);
        for (k = 0; k < nzp; k++)                                  /* FIRSTORDERFORWARDDIFF */
          r[k] = (k == nzp - 1) ? imp[0] - imp[k] : imp[k + 1] - imp[k];
        spectrum_f(r, nzp, R);
        /* local wavelet: shiftAndScale of the global spectrum */
        float ts = (shift && shift[l]) ? shift[l][ti] : 0.0f;
        float gn = (gain && gain[l]) ? gain[l][ti] : 1.0f;
        float ishift = ts / dz;
        if (!(shift && shift[l])) ishift = 0.0f;
        for (int q = 0; q < cn; q++) {
          ocpx am, tmp;
          am.re = (float)(gn * cos(2.0 * (M_PI * (ishift) * q) / (float)nzp));
          am.im = (float)(gn * sin(-2.0 * (M_PI * (ishift) * q) / (float)nzp));
          tmp.re = am.re * W[q].re - am.im * W[q].im;
          tmp.im = am.im * W[q].re + am.re * W[q].im;
          L[q] = tmp;
        }
        float scale = (sf && sf[l]) ? sf[l][ti] : 1.0f;
        for (k = 0; k < cn; k++) {
          ocpx w;                                                   /* getCAmp(k, scale), k < cnzp */
          float omega = (float)k / scale;
          if (omega >= cn) { w.re = 0.0f; w.im = 0.0f; }
          else {
            int omL = (int)floorf(omega), omU = (int)floorf(omega) + 1;
            if (omU >= cn) omU -= 1;
            float dO = omega - (float)omL;
            w.re = L[omL].re * (1.0f - dO) + L[omU].re * dO;
            w.im = (L[omL].im * (1.0f - dO) + L[omU].im * dO);
          }
          w.re /= scale;
          w.im /= scale;
          ocpx s;                                                   /* :1711-1714 */
          s.re = R[k].re * w.re + R[k].im * w.im;
          s.im = -R[k].re * w.im + R[k].im * w.re;
          R[k] = s;
        }
        c2r_f(R, nzp, r);
        const double inv = 1.0 / (double)nzp;                       /* invFFT1DInPlace scale */
        for (k = 0; k < nzp; k++) g[k * plane + tr] = (float)(r[k] * inv);
      }
  }
  free(imp); free(r); free(R); free(W); free(L);
}

/* ========================================================================
 * Time-lapse merge, State4D::merge mean loop (src/state4d.cpp:117-136) and
 * State4D::mergeCov (src/state4d.cpp:161-236), n complex points per grid.
 * ====================================================================== */
void orc_merge_mean_4d(long long n, float **ms, float **md, float **mu)
{
  for (long long p = 0; p < n; p++) {
    ocpx full[6], cur[3];
    for (int l = 0; l < 3; l++) {
      full[l].re = ms[l][2 * p]; full[l].im = ms[l][2 * p + 1];
      full[l + 3].re = md[l][2 * p]; full[l + 3].im = md[l][2 * p + 1];
    }
    for (int l = 0; l < 3; l++) {
      cur[l].re = full[l].re + full[l + 3].re;
      cur[l].im = full[l].im + full[l + 3].im;
    }
    for (int l = 0; l < 3; l++) { mu[l][2 * p] = cur[l].re; mu[l][2 * p + 1] = cur[l].im; }
  }
}

/* ss[6], dd[6]: upper triangles aa ab ar bb br rr; sd[9]: sd[3a+b]; sigma[6] */
void orc_merge_cov_4d(long long n, float **ss, float **dd, float **sd, float **sigma)
{
  for (long long p = 0; p < n; p++) {
    ocpx F[6][6], C[3][3];
    int q = 0;
    for (int a = 0; a < 3; a++)
      for (int b = a; b < 3; b++, q++) {
        F[a][b].re = ss[q][2 * p]; F[a][b].im = ss[q][2 * p + 1];
        F[a + 3][b + 3].re = dd[q][2 * p]; F[a + 3][b + 3].im = dd[q][2 * p + 1];
      }
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) {
        F[a][b + 3].re = sd[3 * a + b][2 * p];
        F[a][b + 3].im = sd[3 * a + b][2 * p + 1];
      }
    for (int l = 0; l < 6; l++)
      for (int m = l + 1; m < 6; m++) {
        F[m][l].re = F[l][m].re;
        F[m][l].im = -F[l][m].im;
      }
    for (int l = 0; l < 3; l++)
      for (int m = 0; m < 3; m++) {
        C[l][m].re = F[l][m].re;
        C[l][m].re += F[l + 3][m + 3].re;
        C[l][m].re += F[l + 3][m].re;
        C[l][m].re += F[l][m + 3].re;
        C[l][m].im = F[l][m].im;
        C[l][m].im += F[l + 3][m + 3].im;
        C[l][m].im += F[l + 3][m].im;
        C[l][m].im += F[l][m + 3].im;
      }
    q = 0;
    for (int a = 0; a < 3; a++)
      for (int b = a; b < 3; b++, q++) { sigma[q][2 * p] = C[a][b].re; sigma[q][2 * p + 1] = C[a][b].im; }
  }
}

/* ========================================================================
 * Wavelet division of the non-constant-thickness path:
 * Crava::divideDataByScaleWavelet (src/crava.cpp:592-714) with
 * Crava::computeAdjustmentFactor (:726-817),
 * computeElasticImpedanceTimeCovariance (:476-487),
 * computeReflectionCoefficientTimeCovariance (:490-504),
 * FFTGrid::getDistToBoundary (src/fftgrid.cpp:1374-1412), the Wavelet1D
 * copy + shiftAndScale (src/wavelet.cpp:53-88, src/wavelet1D.cpp:497-561),
 * the difference operators (src/wavelet1D.cpp:420-470),
 * Wavelet::findNormWithinFrequencyBand (src/wavelet.cpp:811-833),
 * Wavelet::multiplyRAmpByConstant (:466-474), Wavelet::getCAmp (:309-384).
 * One-dimensional wavelets: the local wavelet is the global one shifted and
 * scaled with the shift / gain of the mirrored trace (iInd, jInd), read
 * through the stretch sf = relThick(i, j).
 * The reference re-derives the global wavelet's spectrum once per trace
 * through a float c2r / r2c round trip (Wavelet's copy constructor and
 * fft1DInPlace); this restatement keeps the spectrum of the given samples.
 * ====================================================================== */
#define ORC_RMISSING (-99999.0f)                 /* src/definitions.h:35 */

/* FFTGrid::getDistToBoundary, src/fftgrid.cpp:1374-1412 */
float orc_dist_to_boundary(int i, int n, int np)
{
  float dist = 0.0f;
  if (i < np) {
    if (i >= n) {
      const int mn = n < np - n ? n : np - n;
      const float taper = (float)(mn / 2.1);
      const int below = np - i, above = i - (n - 1);
      if (above < below) dist = (float)above / taper;
      else dist = (float)below / taper;
    }
  } else {
    dist = ORC_RMISSING;
  }
  return dist;
}

/* FFTGrid::getFillNumber, src/fftgrid.cpp:1319-1357 (i < np) */
int orc_fill_number(int i, int n, int np)
{
  if (i >= np) return -99999;
  if (i < n) return i;
  const int below = np - i - 1, above = i - n + 1;
  if (above < below) return (n - above) > n / 2 ? (n - above) : n / 2;
  return below < n / 2 ? below : n / 2;
}

/* trace index of the wavelet's local properties, src/crava.cpp:626-642 */
static int swd_mirror(int i, int n, int np)
{
  int r = i;
  if (r > 3 * n - 1) r = 0;
  if (r > (np + n) / 2) r = np - r;
  if (r >= n) r = 2 * n - r - 1;
  return r;
}

/* Wavelet::getCAmp(k, scale) for k < nzp/2 + 1, src/wavelet.cpp:328-357,381-383 */
static ocpx camp_scaled(const ocpx *c, int cn, int k, float scale)
{
  ocpx v;
  const float omega = (float)k / scale;
  if (omega >= cn) { v.re = 0.0f; v.im = 0.0f; }
  else {
    const int omL = (int)floorf(omega);
    int omU = (int)floorf(omega) + 1;
    if (omU >= cn) omU -= 1;
    const float d = omega - (float)omL;
    v.re = c[omL].re * (1.0f - d) + c[omU].re * d;
    v.im = (c[omL].im * (1.0f - d) + c[omU].im * d);
  }
  v.re /= scale;
  v.im /= scale;
  return v;
}

/*
 *   wavelets [nt][nzp], wnorm [nt] (Wavelet::getNorm), A [nt][3],
 *   var0 [3][3] priorVar0_, corr_t [nzp] getPriorCorrTFiltered(nz, nzp),
 *   err_var [nt] errThetaCov_[l][l], rel_thick [nyp][nxp] Simbox::getRelThick,
 *   shift / gain: NULL or nt pointers, each NULL or [ny][nx];
 *   grids [nt] real padded [nzp][nyp][rnxp], divided in place;
 *   traces [i0, i1) x [0, nyp) only (i1 <= 0: all), the others untouched.
 */
void orc_divide_by_scale_wavelet(int nx, int ny, int nz, int nxp, int nyp, int nzp, int nt,
                                 const float *wavelets, const float *wnorm, const float *A,
                                 const double *var0, const float *corr_t, const float *err_var,
                                 const double *rel_thick, const float *const *shift,
                                 const float *const *gain, float dz, double lz, float low_cut,
                                 float high_cut, float wnc, float **grids, int i0, int i1)
{
  const int rnxp = 2 * (nxp / 2 + 1), cn = nzp / 2 + 1;
  const long long plane = (long long)rnxp * nyp;
  const float tol_fac = 0.05f;                                   /* :734 */
  if (i1 <= 0 || i1 > nxp) i1 = nxp;
  float *r = (float *)malloc(sizeof(float) * nzp);
  float *rc = (float *)malloc(sizeof(float) * nzp);
  float *wt = (float *)malloc(sizeof(float) * nzp);
  float *fw = (float *)malloc(sizeof(float) * nzp);
  float *bw = (float *)malloc(sizeof(float) * nzp);
  float *cw = (float *)malloc(sizeof(float) * nzp);
  ocpx *W = (ocpx *)malloc(sizeof(ocpx) * cn);
  ocpx *RC = (ocpx *)malloc(sizeof(ocpx) * cn);
  ocpx *R = (ocpx *)malloc(sizeof(ocpx) * cn);
  ocpx *WL = (ocpx *)malloc(sizeof(ocpx) * cn);
  ocpx *E3 = (ocpx *)malloc(sizeof(ocpx) * cn);
  const float sq = sqrtf((float)nzp);
  for (int l = 0; l < nt; l++) {
    const float *Al = A + 3 * l;
    float *g = grids[l];
    spectrum_f(wavelets + (long long)l * nzp, nzp, W);
    /* rcSpecIntens: computeReflectionCoefficientTimeCovariance, then r2c */
    double ei_var = 0.0;
    for (int a = 0; a < 3; a++)
      for (int b = 0; b < 3; b++) ei_var += (double)(Al[a] * (float)var0[3 * a + b] * Al[b]);
    for (int k = 0; k < nzp; k++) rc[k] = (float)(ei_var * corr_t[k]);
    {
      const float first = rc[0];
      float prev = rc[nzp - 1];
      for (int t = 0; t < nzp - 1; t++) {
        const float curr = rc[t], next = rc[t + 1];
        rc[t] = 2 * curr - next - prev;
        prev = curr;
      }
      rc[nzp - 1] = 2 * rc[nzp - 1] - first - prev;
    }
    spectrum_f(rc, nzp, RC);
    float mod_w = wnorm[l];                                        /* :765-768 */
    mod_w *= mod_w;
    const float maxfreq = (float)((nzp / 2) * 1000.0 * nz) / (float)(lz * nzp);
    mod_w *= maxfreq / high_cut;
    const float ev = err_var[l], wn = wnorm[l];
    const float d_om = (float)(1.0 / ((double)((float)nzp * dz) * 0.001));
    int lo_i = (int)floorf(low_cut / d_om), hi_i = (int)ceilf(high_cut / d_om);
    if (lo_i < 0) lo_i = 0;
    if (hi_i > cn) hi_i = cn;
    for (int i = i0; i < i1; i++)
      for (int j = 0; j < nyp; j++) {
        const int ii = swd_mirror(i, nx, nxp), jj = swd_mirror(j, ny, nyp);
        const long long tr = (long long)j * rnxp + i;
        int k;
        for (k = 0; k < nzp; k++) {                                /* :644-657 */
          r[k] = g[k * plane + tr] / sq;
          if (k > nz) {
            const float dist = orc_dist_to_boundary(k, nz, nzp);
            const float f = 1 - dist * dist;
            r[k] *= (f < 0.0f) ? 0.0f : f;
          }
        }
        spectrum_f(r, nzp, R);
        /* local wavelet (Wavelet::getLocalTimeshift / getLocalGainFactor) */
        float ts = 0.0f, gn = 1.0f;
        if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) {
          if (shift && shift[l]) ts = shift[l][(long long)jj * nx + ii];
          if (gain && gain[l]) gn = gain[l][(long long)jj * nx + ii];
        }
        const float ishift = ts / dz;
        for (int q = 0; q < cn; q++) {
          ocpx am, tmp;
          am.re = (float)(gn * cos(2.0 * (M_PI * (ishift) * q) / (float)nzp));
          am.im = (float)(gn * sin(-2.0 * (M_PI * (ishift) * q) / (float)nzp));
          tmp.re = am.re * W[q].re - am.im * W[q].im;
          tmp.im = am.im * W[q].re + am.re * W[q].im;
          WL[q] = tmp;
        }
        /* errorSmooth3 = central(backward(forward(wLocal))) in time; the
         * copy takes wLocal back to its samples (invFFT1DInPlace) */
        c2r_f(WL, nzp, wt);
        for (int t = 0; t < nzp; t++) wt[t] = (float)(wt[t] * (1.0 / (double)nzp));
        for (int t = 0; t < nzp; t++) fw[t] = (t == nzp - 1) ? (wt[0] - wt[t]) : (wt[t + 1] - wt[t]);
        for (int t = 0; t < nzp; t++) bw[t] = (t == 0) ? (fw[0] - fw[nzp - 1]) : (fw[t] - fw[t - 1]);
        for (int t = 0; t < nzp; t++) {
          if (t == 0) cw[t] = (float)(0.5 * (bw[t + 1] - bw[nzp - 1]));
          else if (t == nzp - 1) cw[t] = (float)(0.5 * (bw[0] - bw[t - 1]));
          else cw[t] = (float)(0.5 * (bw[t + 1] - bw[t - 1]));
        }
        spectrum_f(cw, nzp, E3);
        spectrum_f(wt, nzp, WL);                                   /* wLocal->fft1DInPlace() */
        double n2 = 0.0;
        const double fac = 1.0 / (double)nzp;
        for (int q = lo_i; q < hi_i; q++) n2 += 2.0 * fac * (double)(E3[q].re * E3[q].re + E3[q].im * E3[q].im);
        if (lo_i == 0) n2 -= fac * (double)(E3[0].re * E3[0].re + E3[0].im * E3[0].im);
        const float c = 1.0f / (float)sqrt(n2);
        /* multiplyRAmpByConstant: floats of index < nzp scaled, the rest of
         * the interleaved half spectrum set to RMISSING */
        float *e3f = (float *)E3;
        for (int t = 0; t < 2 * cn; t++) {
          if (t < nzp) e3f[t] *= c;
          else e3f[t] = ORC_RMISSING;
        }
        const double rel_t = rel_thick[(long long)j * nxp + i];
        const float sf = (float)(rel_t * 1.0f);                   /* getLocalStretch = 1 */
        const double delta_f = (double)nz * 1000.0 / (rel_t * lz * (double)nzp);
        for (k = 0; k < cn; k++) {                                 /* :772-814 */
          const ocpx w_loc = camp_scaled(WL, cn, k, sf);
          const float w_loc2 = (w_loc.re * w_loc.re + w_loc.im * w_loc.im);
          const ocpx w_glob = W[k];
          const float w_glob2 = (w_glob.re * w_glob.re + w_glob.im * w_glob.im);
          ocpx e = camp_scaled(WL, cn, k, sf);
          float s_loc = (e.re * e.re + e.im * e.im) / (wn * wn);
          e = camp_scaled(E3, cn, k, sf);
          s_loc += e.re * e.re + e.im * e.im;
          e = WL[k];
          float s_glob = (e.re * e.re + e.im * e.im) / (wn * wn);
          e = E3[k];
          s_glob += e.re * e.re + e.im * e.im;
          float sig_loc = ev * (wnc + (1.0f - wnc) * 0.5f * s_loc);
          const float sig_glob = ev * (wnc + (1.0f - wnc) * 0.5f * s_glob);
          const float tau2 = fabsf(RC[k].re);
          float wg2a = w_glob2, wl2a = w_loc2;
          if (w_glob2 < tol_fac * mod_w) wg2a = (float)(0.5f * (w_glob2 + tol_fac * mod_w));
          if (w_loc2 < tol_fac * mod_w) wl2a = (float)(0.5f * (w_loc2 + tol_fac * mod_w));
          if (wl2a < tol_fac * wg2a) {
            wl2a = (float)(0.5f * (w_loc2 + tol_fac * wg2a));
            sig_loc *= wl2a / w_loc2;
          }
          ocpx adj;
          adj.re = (w_loc.re / (tau2 * wl2a + sig_loc)) / (wg2a / (tau2 * w_glob2 + sig_glob));
          adj.im = (w_loc.im / (tau2 * wl2a + sig_loc)) / (wg2a / (tau2 * w_glob2 + sig_glob));
          if ((delta_f * k < high_cut) && (delta_f * k > low_cut)) {
            const float tmp = R[k].re * adj.re - R[k].im * adj.im;
            R[k].im = R[k].im * adj.re + R[k].re * adj.im;
            R[k].re = tmp;
          } else {
            R[k].im = 0.0f;
            R[k].re = 0.0f;
          }
        }
        c2r_f(R, nzp, r);
        for (k = 0; k < nzp; k++) g[k * plane + tr] = r[k] / sq;
      }
  }
  free(r); free(rc); free(wt); free(fw); free(bw); free(cw);
  free(W); free(RC); free(R); free(WL); free(E3);
}

/* ========================================================================
 * FFTGrid::interpolateSeismic (src/fftgrid.cpp:2564-2655) with
 * interpolateTrace (:2658-2696) and extrapolateSeismic (:2699-2733) on a
 * real padded grid [nzp][nyp][rnxp].  Returns the number of traces flagged
 * for interpolation; *n_zero receives those with zero energy.
 * ====================================================================== */
static int isq_interp_trace(float *g, long long plane, int rnxp, int nx, int ny, int nzp, int index,
                            short *flags, int i, int j)
{
  int nt = 0;
  const int left = (i > 0 && (flags[index - 1] % 2) == 0);
  const int right = (i < nx - 1 && (flags[index + 1] % 2) == 0);
  const int up = (j > 0 && (flags[index - nx] % 2) == 0);
  const int down = (j < ny - 1 && (flags[index + nx] % 2) == 0);
  nt = left + right + up + down;
  if (nt > 1) {
    const long long t = (long long)j * rnxp + i;
    for (int k = 0; k < nzp; k++) {
      const long long o = k * plane + t;
      float m = left ? g[o - 1] : 0.0f;
      if (right) m += g[o + 1];
      if (up) m += g[o - rnxp];
      if (down) m += g[o + rnxp];
      g[o] = m / (float)nt;
    }
    flags[index]--;
  }
  return flags[index];
}

int orc_interpolate_seismic(int nx, int ny, int nz, int nxp, int nyp, int nzp, float threshold, float *g,
                            int *n_zero)
{
  const int rnxp = 2 * (nxp / 2 + 1);
  const long long plane = (long long)rnxp * nyp;
  short *flags = (short *)malloc(sizeof(short) * nx * ny);
  float *emap = (float *)malloc(sizeof(float) * nx * ny);
  int flag = 0, imin = nx, imax = 0, jmin = ny, jmax = 0, index = 0;
  float total = 0.0f;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      float e = 0.0f;
      for (int k = 0; k < nz; k++) {
        const float v = g[k * plane + (long long)j * rnxp + i];
        e += v * v;
      }
      emap[index] = e;
      total += e;
      index++;
    }
  const float limit = threshold * total / (float)(nx * ny);
  int n_inter = 0, n_inter0 = 0;
  index = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      int cur = 0;
      if (emap[index] <= limit) {
        cur = 1;
        n_inter++;
        if (emap[index] == 0.0f) n_inter0++;
      }
      flags[index] = (short)(flag + cur);
      if (cur == 1) flag = 2;
      else {
        if (i < imin) imin = i;
        if (i > imax) imax = i;
        if (j < jmin) jmin = j;
        if (j > jmax) jmax = j;
      }
      index++;
    }
  int ci = 0;
  for (int j = 0; j < ny; j++)
    for (int i = 0; i < nx; i++) {
      if ((flags[ci] % 2) == 1) {
        if ((isq_interp_trace(g, plane, rnxp, nx, ny, nzp, ci, flags, i, j) % 2) == 0) {
          index = ci - 1;
          while (index >= 0 && flags[index] > 1) {
            if ((flags[index] % 2) == 1)
              isq_interp_trace(g, plane, rnxp, nx, ny, nzp, index, flags, index % nx, index / nx);
            index--;
          }
          if (index < 0) index = 0;
          index++;
          while (flags[index - 1] == 0 && index <= ci) {
            if (flags[index] > 1) flags[index] -= 2;
            index++;
          }
        }
      }
      ci++;
    }
  /* extrapolateSeismic */
  for (int j = 0; j < nyp; j++) {
    int refj = orc_fill_number(j, ny, nyp);
    if (refj < jmin) refj = jmin;
    else if (refj > jmax) refj = jmax;
    for (int i = 0; i < nxp; i++) {
      if (i < imin || i > imax || j < jmin || j > jmax) {
        int refi = orc_fill_number(i, nx, nxp);
        if (refi < imin) refi = imin;
        else if (refi > imax) refi = imax;
        for (int k = 0; k < nzp; k++) {
          const float value = g[k * plane + (long long)refj * rnxp + refi];
          const float dx = orc_dist_to_boundary(i, nx, nxp);
          const float dy = orc_dist_to_boundary(j, ny, nyp);
          const float dzz = orc_dist_to_boundary(k, nz, nzp);
          double b = 1.0 - dx * dx - dy * dy - dzz * dzz;
          if (b < 0.0) b = 0.0;
          const float mult = (float)pow(b, 3);
          g[k * plane + (long long)j * rnxp + i] = mult * value;
        }
      }
    }
  }
  free(flags);
  free(emap);
  if (n_zero) *n_zero = n_inter0;
  return n_inter;
}
