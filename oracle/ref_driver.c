/*
 * ref_driver.c -- drives the reference's OWN C sources (FFTW 2.1.2 float and
 * libs/lib/lib_matr.c, compiled from /root/reference by oracle/Makefile into
 * oracle/_ref/libcrava_ref.so) in the order src/crava.cpp uses them.
 *
 * TEST INFRASTRUCTURE ONLY (the checker and the CPU baseline): the product
 * never links or loads it.  The C++ glue around the C calls
 * (getNextParameterCovariance, getNextErrorVariance, fftInPlace scaling) is
 * restated here because the reference's C++ units need cblas/LAPACK headers
 * that this image does not have.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "rfftw.h"
#include "lib/lib_matr.h"

#define REF_MAXT 8

/* FFTGrid::fftInPlace, src/fftgrid.cpp:1804-1829 */
void ref_fft3d_r2c(float *g, int nxp, int nyp, int nzp, int is_cov)
{
  int cnxp = nxp / 2 + 1, rnxp = 2 * cnxp;
  long long rsize = (long long)rnxp * nyp * nzp;
  if (!is_cov) {
    float s = 1.0f / sqrtf((float)(nxp * nyp * nzp));
    for (long long t = 0; t < rsize; t++) g[t] *= s;
  }
  rfftwnd_plan plan = rfftw3d_create_plan(nzp, nyp, nxp, FFTW_REAL_TO_COMPLEX,
                                          FFTW_ESTIMATE | FFTW_IN_PLACE);
  rfftwnd_one_real_to_complex(plan, (fftw_real *)g, (fftw_complex *)g);
  fftwnd_destroy_plan(plan);
}

/* FFTGrid::invFFTInPlace, src/fftgrid.cpp:1832-1862 */
void ref_fft3d_c2r(float *g, int nxp, int nyp, int nzp, int is_cov)
{
  int cnxp = nxp / 2 + 1, rnxp = 2 * cnxp;
  long long rsize = (long long)rnxp * nyp * nzp;
  float scale;
  if (is_cov) scale = (float)(1.0 / (nxp * nyp * nzp));
  else scale = (float)(1.0 / sqrtf((float)(nxp * nyp * nzp)));
  rfftwnd_plan plan = rfftw3d_create_plan(nzp, nyp, nxp, FFTW_COMPLEX_TO_REAL,
                                          FFTW_ESTIMATE | FFTW_IN_PLACE);
  rfftwnd_one_complex_to_real(plan, (fftw_complex *)g, (fftw_real *)g);
  fftwnd_destroy_plan(plan);
  for (long long t = 0; t < rsize; t++) g[t] *= scale;
}

/* SeismicParametersHolder::computeCircCorrT + makeCircCorrTPosDef
 * (src/seismicparametersholder.cpp:431-499) on the reference's FFTW 1-D
 * transforms (FFTGrid::fft1DzInPlace / invFFT1DzInPlace, src/fftgrid.cpp:
 * 2002-2037): symmetric circular extension, |spectrum| with k <= minIntFq
 * zeroed, unnormalised inverse, scaled to 1 at lag 0.  out: nzp floats.
 * Returns 1 where the reference would exit (c(0) <= 1e-5). */
int ref_circ_corr_t(const float *prior, int n, int min_int_fq, int nzp, float *out)
{
  int cn = nzp / 2 + 1, k;
  fftw_real *c = (fftw_real *)malloc(sizeof(fftw_real) * 2 * cn);
  for (k = 0; k < 2 * cn; k++) {
    if (k < nzp) {
      int refk = k < nzp / 2 + 1 ? k : nzp - k;
      c[k] = refk < n ? prior[refk] : 0.0f;
    } else {
      c[k] = -99999.0f;   /* RMISSING */
    }
  }
  rfftwnd_plan plan = rfftwnd_create_plan(1, &nzp, FFTW_REAL_TO_COMPLEX, FFTW_ESTIMATE | FFTW_IN_PLACE);
  rfftwnd_one_real_to_complex(plan, c, (fftw_complex *)c);
  fftwnd_destroy_plan(plan);
  fftw_complex *f = (fftw_complex *)c;
  for (k = 0; k < cn; k++) {
    if (k <= min_int_fq) f[k].re = 0.0f;
    else f[k].re = (float)sqrtf(f[k].re * f[k].re + f[k].im * f[k].im);
    f[k].im = 0.0f;
  }
  plan = rfftwnd_create_plan(1, &nzp, FFTW_COMPLEX_TO_REAL, FFTW_ESTIMATE | FFTW_IN_PLACE);
  rfftwnd_one_complex_to_real(plan, f, c);
  fftwnd_destroy_plan(plan);
  int bad = !(c[0] > 1.e-5f);
  float scale = bad ? 0.0f : (float)(1.0f / c[0]);
  for (k = 0; k < nzp; k++) out[k] = c[k] * scale;
  free(c);
  return bad;
}

/* 1-D r2c (Wavelet::fft1DInPlace, src/wavelet.cpp:246-261) */
void ref_fft1d_r2c(float *w, int nzp)
{
  rfftwnd_plan plan = rfftwnd_create_plan(1, &nzp, FFTW_REAL_TO_COMPLEX,
                                          FFTW_ESTIMATE | FFTW_IN_PLACE);
  rfftwnd_one_real_to_complex(plan, (fftw_real *)w, (fftw_complex *)w);
  fftwnd_destroy_plan(plan);
}

int ref_chol_cpx(int n, float *mat)
{
  fftw_complex *rows[REF_MAXT];
  for (int i = 0; i < n; i++) rows[i] = (fftw_complex *)mat + i * n;
  return lib_matrCholCpx(n, rows);
}

void ref_axeqb_mat_cpx(int n, float *L, float *X, int m)
{
  fftw_complex *lr[REF_MAXT], *xr[REF_MAXT];
  for (int i = 0; i < n; i++) { lr[i] = (fftw_complex *)L + i * n; xr[i] = (fftw_complex *)X + i * m; }
  lib_matrAXeqBMatCpx(n, lr, xr, m);
}

/* getNextParameterCovariance, src/seismicparametersholder.cpp:356-427 */
static void ref_param_cov(const double *s0, const fftw_complex *c6, fftw_complex **pv)
{
  static const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  fftw_complex t[6];
  float v[6];
  for (int q = 0; q < 6; q++) t[q] = c6[q];
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

/* Crava::computePostMeanResidAndFFTCov loop, src/crava.cpp:1042-1160,
 * calling the reference's lib_matr routines. */
long long ref_posterior(int nxp, int nyp, int nzp, int ntheta,
                        const float *A, const double *err_theta_cov, float wnc,
                        const double *sigma0,
                        const unsigned char *invert, const float *kWD,
                        const float *e1, const float *e2, const float *e3,
                        float **mu, float **data, float **cov, const float *errcorr)
{
  int cnxp = nxp / 2 + 1, n = ntheta;
  long long nfail = 0;
  fftw_complex Kb[REF_MAXT][3], KSb[REF_MAXT][3], KSccb[3][REF_MAXT], pvb[3][3];
  fftw_complex mvb[REF_MAXT][REF_MAXT], evb[REF_MAXT][REF_MAXT], rvb[3][3];
  fftw_complex *K[REF_MAXT], *KS[REF_MAXT], *KScc[3], *pv[3], *mv[REF_MAXT], *ev[REF_MAXT], *rv[3];
  for (int i = 0; i < REF_MAXT; i++) { K[i] = Kb[i]; KS[i] = KSb[i]; mv[i] = mvb[i]; ev[i] = evb[i]; }
  for (int i = 0; i < 3; i++) { KScc[i] = KSccb[i]; pv[i] = pvb[i]; rv[i] = rvb[i]; }
  fftw_complex m1[REF_MAXT], m2[REF_MAXT], m3[REF_MAXT], kW[REF_MAXT];
  fftw_complex ijkData[REF_MAXT], ijkDataMean[REF_MAXT], ijkRes[REF_MAXT], ijkMean[3], ijkAns[3];
  fftw_complex **mu_c = (fftw_complex **)mu, **d_c = (fftw_complex **)data, **cv_c = (fftw_complex **)cov;
  float **Ar = (float **)malloc(sizeof(float *) * n);
  for (int l = 0; l < n; l++) Ar[l] = (float *)A + 3 * l;
  const fftw_complex *ec = (const fftw_complex *)errcorr;
  for (int k = 0; k < nzp; k++) {
    for (int l = 0; l < n; l++) {
      long long o = ((long long)k * n + l) * 2;
      kW[l].re = kWD[o]; kW[l].im = kWD[o + 1];
      m1[l].re = e1[o]; m1[l].im = e1[o + 1];
      m2[l].re = e2[o]; m2[l].im = e2[o + 1];
      m3[l].re = e3[o]; m3[l].im = e3[o + 1];
    }
    lib_matrProdDiagCpxR(kW, Ar, n, 3, K);
    int inv = invert[k];
    for (int j = 0; j < nyp; j++)
      for (int i = 0; i < cnxp; i++) {
        long long p = ((long long)k * nyp + j) * cnxp + i;
        for (int a = 0; a < 3; a++) ijkMean[a] = mu_c[a][p];
        for (int l = 0; l < n; l++) { ijkData[l] = d_c[l][p]; ijkRes[l] = ijkData[l]; }
        fftw_complex c6[6];
        for (int q = 0; q < 6; q++) c6[q] = cv_c[q][p];
        ref_param_cov(sigma0, c6, pv);
        fftw_complex tmp = ec[p];
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
          lib_matrProdCpx(K, pv, n, 3, 3, KS);
          lib_matrProdAdjointCpx(KS, K, n, 3, n, mv);
          lib_matrAddMatCpx(ev, n, n, mv);
          if (lib_matrCholCpx(n, mv) == 0) {
            lib_matrAdjoint(KS, n, 3, KScc);
            lib_matrAXeqBMatCpx(n, mv, KS, 3);
            lib_matrProdCpx(KScc, KS, 3, n, 3, rv);
            lib_matrSubtMatCpx(rv, 3, 3, pv);
            lib_matrProdMatVecCpx(K, ijkMean, n, 3, ijkDataMean);
            lib_matrSubtVecCpx(ijkDataMean, n, ijkData);
            lib_matrProdAdjointMatVecCpx(KS, ijkData, 3, n, ijkAns);
            lib_matrAddVecCpx(ijkAns, 3, ijkMean);
            lib_matrProdMatVecCpx(K, ijkMean, n, 3, ijkData);
            lib_matrSubtVecCpx(ijkData, n, ijkRes);
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
  free(Ar);
  return nfail;
}

/* Crava::simulate per-point loop, src/crava.cpp:1482-1521 */
long long ref_simulate_points(long long npts, float **cov, float **seed)
{
  long long nfail = 0;
  fftw_complex pcb[3][3];
  fftw_complex *pc[3] = {pcb[0], pcb[1], pcb[2]};
  fftw_complex **cv = (fftw_complex **)cov, **sd = (fftw_complex **)seed;
  for (long long p = 0; p < npts; p++) {
    pc[0][0] = cv[0][p]; pc[1][1] = cv[1][p]; pc[2][2] = cv[2][p];
    pc[0][1] = cv[3][p]; pc[0][2] = cv[4][p]; pc[1][2] = cv[5][p];
    pc[1][0].re = pc[0][1].re; pc[1][0].im = -pc[0][1].im;
    pc[2][0].re = pc[0][2].re; pc[2][0].im = -pc[0][2].im;
    pc[2][1].re = pc[1][2].re; pc[2][1].im = -pc[1][2].im;
    fftw_complex s[3] = {sd[0][p], sd[1][p], sd[2][p]};
    if (lib_matrCholCpx(3, pc) == 0) {
      lib_matrProdCholVec(3, pc, s);
    } else {
      nfail++;
      for (int l = 0; l < 3; l++) { s[l].re = 0.0f; s[l].im = 0.0f; }
    }
    sd[0][p] = s[0]; sd[1][p] = s[1]; sd[2][p] = s[2];
  }
  return nfail;
}

/* Wall time of the reference CPU pipeline on one core (the CPU baseline):
 * nfwd forward transforms, the per-point loop, ninv inverse transforms. */
double ref_wall_seconds(void)
{
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

/* Crava::computeSyntSeismic trace loop (src/crava.cpp:1667-1759) with
 * computeSeismicImpedance (:1632-1663), run on the reference's FFTW 2.1.2
 * 1-D transforms (Wavelet::fft1DInPlace / invFFT1DInPlace,
 * src/wavelet.cpp:246-280) and the local wavelet of Wavelet1D::shiftAndScale
 * (src/wavelet1D.cpp:539-561) and Wavelet::getCAmp(k, scale)
 * (src/wavelet.cpp:328-384).  Arguments as orc_synthetic_seismic.  Traces
 * [i0, i1) x [0, ny) only (i1 <= 0: all), so the CPU baseline can time a
 * bounded sample. */
void ref_synthetic_seismic(int nx, int ny, int nz, int nxp, int nyp, int nzp, int nt, const float *A,
                           const float *wavelets, float dz, const float *const *shift,
                           const float *const *gain, const float *const *sf, const float *alpha,
                           const float *beta, const float *rho, float **out, int i0, int i1)
{
  const int rnxp = 2 * (nxp / 2 + 1), cn = nzp / 2 + 1, rn = 2 * cn;
  const long long plane = (long long)rnxp * nyp, rsize = plane * nzp;
  float *imp = (float *)malloc(sizeof(float) * rn);
  float *r = (float *)malloc(sizeof(float) * rn);
  float *w = (float *)malloc(sizeof(float) * rn);
  float *loc = (float *)malloc(sizeof(float) * rn);
  fftw_complex *R = (fftw_complex *)r, *Wc = (fftw_complex *)w, *Lc = (fftw_complex *)loc;
  rfftwnd_plan p1 = rfftwnd_create_plan(1, &nzp, FFTW_REAL_TO_COMPLEX, FFTW_ESTIMATE | FFTW_IN_PLACE);
  rfftwnd_plan p2 = rfftwnd_create_plan(1, &nzp, FFTW_COMPLEX_TO_REAL, FFTW_ESTIMATE | FFTW_IN_PLACE);
  if (i1 <= 0 || i1 > nx) i1 = nx;
  for (int l = 0; l < nt; l++) {
    const float *Al = A + 3 * l;
    float *g = out[l];
    for (long long t = 0; t < rsize; t++) {
      float v = 0;
      v += alpha[t] * Al[0];
      v += beta[t] * Al[1];
      v += rho[t] * Al[2];
      g[t] = v;
    }
    memcpy(w, wavelets + (long long)l * nzp, sizeof(float) * nzp);
    rfftwnd_one_real_to_complex(p1, (fftw_real *)w, Wc);
    for (int i = i0; i < i1; i++)
      for (int j = 0; j < ny; j++) {
        const long long tr = (long long)j * rnxp + i, ti = (long long)j * nx + i;
        int k;
        for (k = 0; k < nz; k++) imp[k] = g[k * plane + tr];
        float fac = 1.0f / (float)(nzp - nz - 1);
        for (; k < nzp; k++) imp[k] = fac * ((k - nz) * imp[0] + (nzp - k - 1) * imp[nz - 1]);
        for (k = 0; k < nzp; k++) r[k] = (k == nzp - 1) ? imp[0] - imp[k] : imp[k + 1] - imp[k];
        rfftwnd_one_real_to_complex(p1, (fftw_real *)r, R);
        float ts = (shift && shift[l]) ? shift[l][ti] : 0.0f;
        float gn = (gain && gain[l]) ? gain[l][ti] : 1.0f;
        float ishift = (shift && shift[l]) ? ts / dz : 0.0f;
        for (int q = 0; q < cn; q++) {
          fftw_complex am, tmp;
          am.re = (float)(gn * cos(2.0 * (M_PI * (ishift) * q) / (float)nzp));
          am.im = (float)(gn * sin(-2.0 * (M_PI * (ishift) * q) / (float)nzp));
          tmp.re = am.re * Wc[q].re - am.im * Wc[q].im;
          tmp.im = am.im * Wc[q].re + am.re * Wc[q].im;
          Lc[q] = tmp;
        }
        float scale = (sf && sf[l]) ? sf[l][ti] : 1.0f;
        for (k = 0; k < cn; k++) {
          fftw_complex v, s;
          float omega = (float)k / scale;
          if (omega >= cn) { v.re = 0.0f; v.im = 0.0f; }
          else {
            int omL = (int)floor(omega), omU = (int)floor(omega) + 1;
            if (omU >= cn) omU -= 1;
            float dO = omega - (float)omL;
            v.re = Lc[omL].re * (1.0f - dO) + Lc[omU].re * dO;
            v.im = (Lc[omL].im * (1.0f - dO) + Lc[omU].im * dO);
          }
          v.re /= scale;
          v.im /= scale;
          s.re = R[k].re * v.re + R[k].im * v.im;
          s.im = -R[k].re * v.im + R[k].im * v.re;
          R[k] = s;
        }
        rfftwnd_one_complex_to_real(p2, R, (fftw_real *)r);
        const double inv = 1.0 / (double)nzp;
        for (k = 0; k < nzp; k++) g[k * plane + tr] = (float)(r[k] * inv);
      }
  }
  fftwnd_destroy_plan(p1);
  fftwnd_destroy_plan(p2);
  free(imp); free(r); free(w); free(loc);
}

/* Crava::divideDataByScaleWavelet (src/crava.cpp:592-714) with
 * computeAdjustmentFactor (:726-817) on the reference's FFTW 2.1.2 1-D
 * transforms, including the global wavelet's state changes: Wavelet's copy
 * constructor (src/wavelet.cpp:53-88) takes the global wavelet back to its
 * samples (invFFT1DInPlace) and computeAdjustmentFactor transforms it again
 * (fft1DInPlace), so from the second trace on its spectrum is re-derived
 * through a float c2r / r2c round trip per trace.  Arguments as
 * orc_divide_by_scale_wavelet; traces [i0, i1) x [0, nyp) only (i1 <= 0:
 * all), so the CPU baseline can time a bounded sample. */
static int ref_mirror(int i, int n, int np)
{
  int r = i;
  if (r > 3 * n - 1) r = 0;
  if (r > (np + n) / 2) r = np - r;
  if (r >= n) r = 2 * n - r - 1;
  return r;
}

static float ref_dist(int i, int n, int np)
{
  float dist = 0.0f;
  if (i < np) {
    if (i >= n) {
      const int mn = n < np - n ? n : np - n;
      const float taper = (float)(mn / 2.1);
      const int below = np - i, above = i - (n - 1);
      dist = (above < below) ? (float)above / taper : (float)below / taper;
    }
  } else {
    dist = -99999.0f;
  }
  return dist;
}

static fftw_complex ref_camp(const fftw_complex *c, int cn, int k, float scale)
{
  fftw_complex v;
  const float omega = (float)k / scale;
  if (omega >= cn) { v.re = 0.0f; v.im = 0.0f; }
  else {
    const int omL = (int)floor(omega);
    int omU = (int)floor(omega) + 1;
    if (omU >= cn) omU -= 1;
    const float d = omega - (float)omL;
    v.re = c[omL].re * (1.0f - d) + c[omU].re * d;
    v.im = (c[omL].im * (1.0f - d) + c[omU].im * d);
  }
  v.re /= scale;
  v.im /= scale;
  return v;
}

void ref_divide_by_scale_wavelet(int nx, int ny, int nz, int nxp, int nyp, int nzp, int nt,
                                 const float *wavelets, const float *wnorm, const float *A,
                                 const double *var0, const float *corr_t, const float *err_var,
                                 const double *rel_thick, const float *const *shift,
                                 const float *const *gain, float dz, double lz, float low_cut,
                                 float high_cut, float wnc, float **grids, int i0, int i1)
{
  const int rnxp = 2 * (nxp / 2 + 1), cn = nzp / 2 + 1, rn = 2 * cn;
  const long long plane = (long long)rnxp * nyp;
  const float tol_fac = 0.05f;
  float *r = (float *)malloc(sizeof(float) * rn), *rc = (float *)malloc(sizeof(float) * rn);
  float *gw = (float *)malloc(sizeof(float) * rn), *lw = (float *)malloc(sizeof(float) * rn);
  float *fw = (float *)malloc(sizeof(float) * rn), *bw = (float *)malloc(sizeof(float) * rn);
  float *cw = (float *)malloc(sizeof(float) * rn);
  fftw_complex *R = (fftw_complex *)r, *RC = (fftw_complex *)rc, *GW = (fftw_complex *)gw;
  fftw_complex *LW = (fftw_complex *)lw, *E3 = (fftw_complex *)cw;
  rfftwnd_plan p1 = rfftwnd_create_plan(1, &nzp, FFTW_REAL_TO_COMPLEX, FFTW_ESTIMATE | FFTW_IN_PLACE);
  rfftwnd_plan p2 = rfftwnd_create_plan(1, &nzp, FFTW_COMPLEX_TO_REAL, FFTW_ESTIMATE | FFTW_IN_PLACE);
  const float sq = sqrtf((float)nzp);
  const double inv_n = 1.0 / (double)nzp;
  if (i1 <= 0 || i1 > nxp) i1 = nxp;
  for (int l = 0; l < nt; l++) {
    const float *Al = A + 3 * l;
    float *g = grids[l];
    memcpy(gw, wavelets + (long long)l * nzp, sizeof(float) * nzp);
    int gw_real = 1;
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
    rfftwnd_one_real_to_complex(p1, (fftw_real *)rc, RC);
    float mod_w = wnorm[l];
    mod_w *= mod_w;
    const float maxfreq = (float)((nzp / 2) * 1000.0 * nz) / (float)(lz * nzp);
    mod_w *= maxfreq / high_cut;
    const float ev = err_var[l], wn = wnorm[l];
    const float d_om = (float)(1.0 / ((double)((float)nzp * dz) * 0.001));
    int lo_i = (int)floor(low_cut / d_om), hi_i = (int)ceil(high_cut / d_om);
    if (lo_i < 0) lo_i = 0;
    if (hi_i > cn) hi_i = cn;
    for (int i = i0; i < i1; i++)
      for (int j = 0; j < nyp; j++) {
        const int ii = ref_mirror(i, nx, nxp), jj = ref_mirror(j, ny, nyp);
        const long long tr = (long long)j * rnxp + i;
        int k;
        for (k = 0; k < nzp; k++) {
          r[k] = g[k * plane + tr] / sq;
          if (k > nz) {
            const float dist = ref_dist(k, nz, nzp);
            const float f = 1 - dist * dist;
            r[k] *= (f < 0.0f) ? 0.0f : f;
          }
        }
        rfftwnd_one_real_to_complex(p1, (fftw_real *)r, R);
        /* createLocalWavelet1D: copy of the global wavelet's samples */
        if (!gw_real) {
          rfftwnd_one_complex_to_real(p2, GW, (fftw_real *)gw);
          for (int t = 0; t < nzp; t++) gw[t] = (float)(gw[t] * inv_n);
          gw_real = 1;
        }
        for (int t = 0; t < rn; t++) lw[t] = t < nzp ? gw[t] : 0.0f;
        float ts = 0.0f, gn = 1.0f;
        if (ii >= 0 && ii < nx && jj >= 0 && jj < ny) {
          if (shift && shift[l]) ts = shift[l][(long long)jj * nx + ii];
          if (gain && gain[l]) gn = gain[l][(long long)jj * nx + ii];
        }
        const float ishift = ts / dz;
        rfftwnd_one_real_to_complex(p1, (fftw_real *)lw, LW);
        for (int q = 0; q < cn; q++) {
          fftw_complex am, tmp;
          am.re = (float)(gn * cos(2.0 * (M_PI * (ishift) * q) / (float)nzp));
          am.im = (float)(gn * sin(-2.0 * (M_PI * (ishift) * q) / (float)nzp));
          tmp.re = am.re * LW[q].re - am.im * LW[q].im;
          tmp.im = am.im * LW[q].re + am.re * LW[q].im;
          LW[q] = tmp;
        }
        /* errorSmooth chain from wLocal's samples */
        rfftwnd_one_complex_to_real(p2, LW, (fftw_real *)lw);
        for (int t = 0; t < nzp; t++) lw[t] = (float)(lw[t] * inv_n);
        for (int t = 0; t < nzp; t++) fw[t] = (t == nzp - 1) ? (lw[0] - lw[t]) : (lw[t + 1] - lw[t]);
        for (int t = 0; t < nzp; t++) bw[t] = (t == 0) ? (fw[0] - fw[nzp - 1]) : (fw[t] - fw[t - 1]);
        for (int t = 0; t < nzp; t++) {
          if (t == 0) cw[t] = (float)(0.5 * (bw[t + 1] - bw[nzp - 1]));
          else if (t == nzp - 1) cw[t] = (float)(0.5 * (bw[0] - bw[t - 1]));
          else cw[t] = (float)(0.5 * (bw[t + 1] - bw[t - 1]));
        }
        rfftwnd_one_real_to_complex(p1, (fftw_real *)cw, E3);
        double n2 = 0.0;
        const double fac = 1.0 / (double)nzp;
        for (int q = lo_i; q < hi_i; q++) n2 += 2.0 * fac * (double)(E3[q].re * E3[q].re + E3[q].im * E3[q].im);
        if (lo_i == 0) n2 -= fac * (double)(E3[0].re * E3[0].re + E3[0].im * E3[0].im);
        const float c = 1.0f / (float)sqrt(n2);
        for (int t = 0; t < rn; t++) {
          if (t < nzp) cw[t] *= c;
          else cw[t] = -99999.0f;
        }
        rfftwnd_one_real_to_complex(p1, (fftw_real *)lw, LW);   /* wLocal->fft1DInPlace() */
        if (gw_real) {                                           /* wGlobal->fft1DInPlace() */
          rfftwnd_one_real_to_complex(p1, (fftw_real *)gw, GW);
          gw_real = 0;
        }
        const double rel_t = rel_thick[(long long)j * nxp + i];
        const float sf = (float)(rel_t * 1.0f);
        const double delta_f = (double)nz * 1000.0 / (rel_t * lz * (double)nzp);
        for (k = 0; k < cn; k++) {
          const fftw_complex w_loc = ref_camp(LW, cn, k, sf);
          const float w_loc2 = (w_loc.re * w_loc.re + w_loc.im * w_loc.im);
          const fftw_complex w_glob = GW[k];
          const float w_glob2 = (w_glob.re * w_glob.re + w_glob.im * w_glob.im);
          fftw_complex e = ref_camp(LW, cn, k, sf);
          float s_loc = (e.re * e.re + e.im * e.im) / (wn * wn);
          e = ref_camp(E3, cn, k, sf);
          s_loc += e.re * e.re + e.im * e.im;
          e = LW[k];
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
          fftw_complex adj;
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
        rfftwnd_one_complex_to_real(p2, R, (fftw_real *)r);
        for (k = 0; k < nzp; k++) g[k * plane + tr] = r[k] / sq;
      }
  }
  fftwnd_destroy_plan(p1);
  fftwnd_destroy_plan(p2);
  free(r); free(rc); free(gw); free(lw); free(fw); free(bw); free(cw);
}
