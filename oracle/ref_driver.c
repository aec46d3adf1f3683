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
