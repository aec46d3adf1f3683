// Host-side per-k preparation (see tables.hpp).
#include "tables.hpp"

#include "eig3.hpp"

#include <algorithm>
#include <cmath>
#include <cstdint>
#include <vector>

namespace crava {

namespace {

struct c64 {
  float re, im;
};

// Half spectrum (nzp/2+1 bins) of a real series, forward sign, double
// accumulation with exactly reduced twiddle angles.
// cos and sin of 2 pi m / n, m < n, evaluated in long double and rounded to
// double; kept per length (and per thread), since a workload's calls share
// a few lengths and the long-double evaluation costs more than the sums.
struct TrigTable {
  std::vector<double> cs, sn;
};

const TrigTable &trig_table(int n) {
  thread_local std::vector<std::pair<int, TrigTable>> cache;
  for (const auto &e : cache)
    if (e.first == n) return e.second;
  TrigTable t;
  t.cs.resize(n);
  t.sn.resize(n);
  for (int m = 0; m < n; ++m) {
    const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
    t.cs[m] = (double)cosl(a);
    t.sn[m] = (double)sinl(a);
  }
  if (cache.size() >= 16) cache.erase(cache.begin());
  cache.emplace_back(n, std::move(t));
  return cache.back().second;
}

std::vector<c64> half_spectrum(const float *w, int nzp) {
  const int cn = nzp / 2 + 1;
  const TrigTable &tt = trig_table(nzp);
  const double *cs = tt.cs.data(), *sn = tt.sn.data();
  std::vector<c64> out(cn);
  for (int k = 0; k < cn; ++k) {
    double re = 0.0, im = 0.0;
    // m = k t mod nzp, stepped by k
    int m = 0;
    for (int t = 0; t < nzp; ++t) {
      re += (double)w[t] * cs[m];
      im -= (double)w[t] * sn[m];
      m += k;
      if (m >= nzp) m -= nzp;
    }
    out[k].re = (float)re;
    out[k].im = (float)im;
  }
  // DC and (even length) Nyquist bins of a real series are exactly real
  out[0].im = 0.0f;
  if (nzp % 2 == 0) out[nzp / 2].im = 0.0f;
  return out;
}

c64 camp(const std::vector<c64> &s, int nzp, int k) {
  const int cn = nzp / 2 + 1;
  if (k < cn) return s[k];
  c64 v = s[nzp - k];
  v.im = -v.im;
  return v;
}

}  // namespace

void wavelet_half_spectrum(const float *w, int nzp, float *re_im) {
  const std::vector<c64> s = half_spectrum(w, nzp);
  for (int k = 0; k < nzp / 2 + 1; ++k) {
    re_im[2 * k] = s[k].re;
    re_im[2 * k + 1] = s[k].im;
  }
}

void perk_vectors(int ntheta, int nz, int nzp, double lz, float low_cut, float high_cut, float min_rel_thick,
                  const float *wavelets, const float *wnorm, unsigned char *invert, float *kWD, float *e1,
                  float *e2, float *e3) {
  // forward-difference operator d[0] = -1, d[nzp-1] = +1 (src/wavelet1D.cpp:386-390)
  std::vector<float> d1(nzp, 0.0f);
  d1[0] = -1.0f;
  d1[nzp - 1] = 1.0f;
  const std::vector<c64> dsp = half_spectrum(d1.data(), nzp);

  std::vector<std::vector<c64>> wsp(ntheta), e3sp(ntheta);
  std::vector<float> e3norm(ntheta);
  std::vector<float> fw(nzp), bw(nzp), cw(nzp);
  for (int l = 0; l < ntheta; ++l) {
    const float *w = wavelets + (long long)l * nzp;
    wsp[l] = half_spectrum(w, nzp);
    // difference-operator wavelets, src/wavelet1D.cpp:431-460
    for (int i = 0; i < nzp; ++i) fw[i] = (i == nzp - 1) ? w[0] - w[i] : w[i + 1] - w[i];
    for (int i = 0; i < nzp; ++i) bw[i] = (i == 0) ? fw[0] - fw[nzp - 1] : fw[i] - fw[i - 1];
    for (int i = 0; i < nzp; ++i) {
      const float up = bw[(i + 1) % nzp], dn = bw[(i + nzp - 1) % nzp];
      cw[i] = (float)(0.5 * (up - dn));
    }
    double n2 = 0.0;  // Wavelet::findNorm, time domain
    for (int i = 0; i < nzp; ++i) n2 += (double)(cw[i] * cw[i]);
    e3norm[l] = (float)std::sqrt(n2);
    e3sp[l] = half_spectrum(cw.data(), nzp);
  }
  for (int k = 0; k < nzp; ++k) {
    const int kk = k < nzp - k ? k : nzp - k;
    const float rf = (float)((nz * 1000.0f) / (lz * nzp) * kk);
    invert[k] = (rf > low_cut * min_rel_thick && rf < high_cut) ? 1 : 0;
    const c64 kD = camp(dsp, nzp, k);
    for (int l = 0; l < ntheta; ++l) {
      const long long o = ((long long)k * ntheta + l) * 2;
      const c64 w = camp(wsp[l], nzp, k);
      const float kwr = w.re, kwi = -w.im;  // conj
      kWD[o] = kD.re * kwr - kD.im * kwi;
      kWD[o + 1] = kD.re * kwi + kD.im * kwr;
      e1[o] = w.re / wnorm[l];
      e1[o + 1] = -(w.im / wnorm[l]);
      const c64 s3 = camp(e3sp[l], nzp, k);
      e2[o] = s3.re / e3norm[l];
      e2[o + 1] = -(s3.im / e3norm[l]);
      e3[o] = 1.0f;
      e3[o + 1] = 0.0f;
    }
  }
}

// Spectrum value of a real series at k with Wavelet::getCAmp(k, scale)'s
// stretch (src/wavelet.cpp:327-385): linear interpolation of the half
// spectrum at k / scale, zero beyond the last bin, divided by scale; float
// arithmetic as the reference's.
static c64 camp_scaled(const std::vector<c64> &s, int nzp, int k, float scale) {
  const int cn = nzp / 2 + 1;
  const bool mirror = k >= cn;
  const int kk = mirror ? nzp - k : k;
  c64 v{0.0f, 0.0f};
  const float omega = float(kk) / scale;
  if (!(omega >= (float)cn)) {
    const int omL = (int)std::floor(omega);
    int omU = omL + 1;
    if (omU >= cn) omU -= 1;
    const float dOmega = omega - float(omL);
    v.re = s[omL].re * (1.0f - dOmega) + s[omU].re * dOmega;
    v.im = s[omL].im * (1.0f - dOmega) + s[omU].im * dOmega;
    if (mirror) v.im = -s[omL].im * (1.0f - dOmega) - s[omU].im * dOmega;
  }
  v.re /= scale;
  v.im /= scale;
  return v;
}

// Central o backward o forward difference of a series (the Wavelet1D
// difference constructors, src/wavelet1D.cpp:420-460), in float.
static void diff3(const float *w, int nzp, std::vector<float> &fw, std::vector<float> &bw, std::vector<float> &cw) {
  fw.resize(nzp);
  bw.resize(nzp);
  cw.resize(nzp);
  for (int i = 0; i < nzp; ++i) fw[i] = (i == nzp - 1) ? w[0] - w[i] : w[i + 1] - w[i];
  for (int i = 0; i < nzp; ++i) bw[i] = (i == 0) ? fw[0] - fw[nzp - 1] : fw[i] - fw[i - 1];
  for (int i = 0; i < nzp; ++i) {
    const float up = bw[(i + 1) % nzp], dn = bw[(i + nzp - 1) % nzp];
    cw[i] = (float)(0.5 * (up - dn));
  }
}

void perk_vectors_nonconst(int ntheta, int nz, int nzp, double lz, float low_cut, float high_cut,
                           float min_rel_thick, const float *wavelets, const float *wnorm, unsigned char *invert,
                           float *kWD, float *e1, float *e2, float *e3) {
  // forward-difference operator and its central o backward difference
  // (diff1Operator / diff3Operator, src/crava.cpp:938-944)
  std::vector<float> d1(nzp, 0.0f), fw, bw, cw;
  d1[0] = -1.0f;
  d1[nzp - 1] = 1.0f;
  const std::vector<c64> dsp = half_spectrum(d1.data(), nzp);
  std::vector<float> d2(nzp), d3(nzp);
  for (int i = 0; i < nzp; ++i) d2[i] = (i == 0) ? d1[0] - d1[nzp - 1] : d1[i] - d1[i - 1];
  for (int i = 0; i < nzp; ++i) d3[i] = (float)(0.5 * (d2[(i + 1) % nzp] - d2[(i + nzp - 1) % nzp]));
  const std::vector<c64> d3sp = half_spectrum(d3.data(), nzp);
  std::vector<std::vector<c64>> wsp(ntheta);
  std::vector<float> e3norm(ntheta);
  for (int l = 0; l < ntheta; ++l) {
    const float *w = wavelets + (long long)l * nzp;
    wsp[l] = half_spectrum(w, nzp);
    diff3(w, nzp, fw, bw, cw);   // errorSmooth3, whose norm scales errMult2
    double n2 = 0.0;
    for (int i = 0; i < nzp; ++i) n2 += (double)(cw[i] * cw[i]);
    e3norm[l] = (float)std::sqrt(n2);
  }
  for (int k = 0; k < nzp; ++k) {
    const int kk = k < nzp - k ? k : nzp - k;
    const float rf = (float)((nz * 1000.0f) / (lz * nzp) * kk);
    invert[k] = (rf > low_cut * min_rel_thick && rf < high_cut) ? 1 : 0;
    const c64 kD = camp(dsp, nzp, k), kD3 = camp(d3sp, nzp, k);
    for (int l = 0; l < ntheta; ++l) {
      const long long o = ((long long)k * ntheta + l) * 2;
      // K = D A: the data were divided by the wavelet (src/crava.cpp:1060-1064)
      kWD[o] = kD.re;
      kWD[o + 1] = kD.im;
      e1[o] = 1.0f / wnorm[l];                       // :1067-1071
      e1[o + 1] = 0.0f;
      const float m = 1.0f / e3norm[l];              // :1073-1080
      e2[o] = kD3.re * m;
      e2[o + 1] = kD3.im * m;
      // fillInverseAbskWRobust, :1361-1384
      const c64 v = camp(wsp[l], nzp, k), vf = camp_scaled(wsp[l], nzp, k, 0.999f);
      const float mod = v.re * v.re + v.im * v.im, modf = vf.re * vf.re + vf.im * vf.im;
      const float mx = std::max(mod, modf);
      e3[o] = mx > 0.0f ? float(1.0 / std::sqrt(mx)) : wnorm[l] * nzp * nzp * 100.0f;
      e3[o + 1] = 0.0f;
    }
  }
}

// Wavelet::getWaveletValue (src/wavelet.cpp:1078-1122): simple kriging of the
// six samples around z with a 20 Hz Ricker correlation, in double.  False
// where the kriging matrix is not positive definite.
static bool wavelet_value(float z, const float *w, int center, int nw, float dz, float *value) {
  int ind[6];
  double val[6], cov[6], C[6][6];
  ind[2] = (int)std::floor(z / dz);
  for (int k = 0; k < 6; ++k) ind[k] = ind[2] + k - 2;
  for (int k = 0; k < 6; ++k) {
    const int i = ind[k] + center;
    val[k] = (i > -1 && i < nw) ? (double)w[i] : 0.0;
  }
  const double nu = 20.0, pi = 3.14159265358979323846;
  for (int k = 0; k < 6; ++k) {
    // lags in float (dz * index, then the difference), as the reference forms them
    double dt = (double)(dz * (float)ind[k] - z) * 0.001;
    double f = nu * nu * pi * pi * dt * dt;
    cov[k] = (1 - 2 * f) * std::exp(-f);
    for (int l = 0; l <= k; ++l) {
      dt = (double)(dz * (float)ind[k] - dz * (float)ind[l]) * 0.001;
      f = nu * nu * pi * pi * dt * dt;
      C[l][k] = C[k][l] = (1 - 2 * f) * std::exp(-f);
    }
  }
  // Cholesky solve C x = cov
  double Lm[6][6] = {{0}}, y[6], x[6];
  for (int i = 0; i < 6; ++i) {
    for (int j = 0; j <= i; ++j) {
      double s = C[i][j];
      for (int k = 0; k < j; ++k) s -= Lm[i][k] * Lm[j][k];
      if (i == j && !(s > 0.0)) return false;
      Lm[i][j] = (i == j) ? std::sqrt(s) : s / Lm[j][j];
    }
  }
  for (int i = 0; i < 6; ++i) {
    double s = cov[i];
    for (int k = 0; k < i; ++k) s -= Lm[i][k] * y[k];
    y[i] = s / Lm[i][i];
  }
  for (int i = 5; i >= 0; --i) {
    double s = y[i];
    for (int k = i + 1; k < 6; ++k) s -= Lm[k][i] * x[k];
    x[i] = s / Lm[i][i];
  }
  double v = 0.0;
  for (int k = 0; k < 6; ++k) v += val[k] * x[k];
  *value = (float)v;
  return true;
}

bool wavelet_resample(const float *w, int nw, int center, float dzw, float dz, int nzp, float *out, float *norm) {
  for (int k = 0; k < nzp; ++k) {
    const float z = k < nzp / 2 + 1 ? (float)(dz * k) : (float)(dz * (k - nzp));
    if (!wavelet_value(z, w, center, nw, dzw, out + k)) return false;
  }
  double n2 = 0.0;
  for (int k = 0; k < nzp; ++k) n2 += (double)(out[k] * out[k]);
  *norm = (float)std::sqrt(n2);
  return true;
}

void wavelet_energy(int ntheta, int nzp, const float *wavelets, const float *A, const double *prior_var0,
                    const float *corr_t, const float *error_variance, const float *emp_sn,
                    const unsigned char *match, float *gain, float *model_var, float *signal_var,
                    float *theo_sn) {
  std::vector<float> wd(nzp);
  for (int l = 0; l < ntheta; ++l) {
    const float *w = wavelets + (long long)l * nzp;
    // errorSmooth = forward difference of the wavelet (src/crava.cpp:411-419)
    for (int i = 0; i < nzp; ++i) wd[i] = (i == nzp - 1) ? w[0] - w[i] : w[i + 1] - w[i];
    // paramVar = A_l Sigma0 A_l^T, float accumulation (:423-429)
    float pv = 0.0f;
    for (int a = 0; a < 3; ++a)
      for (int b = 0; b < 3; ++b) pv += (float)prior_var0[a * 3 + b] * A[l * 3 + a] * A[l * 3 + b];
    // computeWDCorrMVar (:1763-1775): w_d^T circulant(corrT) w_d, float accumulation
    float var = 0.0f;
    for (int i = 0; i < nzp; ++i)
      for (int j = 0; j < nzp; ++j) var += wd[i] * corr_t[i > j ? i - j : j - i] * wd[j];
    float mv = var * pv;
    float sv = error_variance[l] + mv;
    float g = 1.0f;
    if (match && match[l]) {                          // energy matching, :443-459
      g = std::sqrt((error_variance[l] / mv) * (emp_sn[l] - 1.0f));
      mv *= g * g;
      sv = error_variance[l] + mv;
    }
    gain[l] = g;
    model_var[l] = mv;
    signal_var[l] = sv;
    theo_sn[l] = sv / error_variance[l];
  }
}

// ---------------------------------------------------------------------------
// Padding sizes (FFTGrid::findClosestFactorableNumber, src/fftgrid.cpp:
// 2837-2866; ModelGeneral::setPaddingSize and estimate*PaddingSize(s),
// src/modelgeneral.cpp:581-587,1337-1425), with the reference's float
// logarithms for the loop bounds.

int closest_factorable(int leastint) {
  const float lf = std::log((float)leastint);
  const int m2 = (int)std::ceil((double)lf / std::log(2.0f));
  const int m3 = (int)std::ceil((double)lf / std::log(3.0f));
  const int m5 = (int)std::ceil((double)lf / std::log(5.0f));
  const int m7 = (int)std::ceil((double)lf / std::log(7.0f));
  int closest = (int)std::pow(2.0f, m2);
  for (int i = 0; i <= m2; ++i)
    for (int j = 0; j <= m3; ++j)
      for (int k = 0; k <= m5; ++k)
        for (int l = 0; l <= m7; ++l) {
          const int f = (int)(std::pow(2.0, i) * std::pow(3.0, j) * std::pow(5.0, k) * std::pow(7.0, l));
          if (f >= leastint && f < closest) closest = f;
        }
  return closest;
}

int padding_size(int n, double pad_fac) {
  return closest_factorable((int)std::ceil(n * (1.0f + pad_fac)));
}

void estimate_xy_padding(int nx, int ny, double dx, double dy, double lx, double ly, double x_pad_fac,
                         double y_pad_fac, bool estimate, float range1, float range2, float angle, int *nxpad,
                         int *nypad, double *true_x_pad_fac, double *true_y_pad_fac) {
  if (estimate) {
    const double factor = 0.5;   // half a lateral range
    double xpad = factor * std::max(std::fabs(range1 * std::cos(angle)), std::fabs(range2 * std::sin(angle)));
    double ypad = factor * std::max(std::fabs(range1 * std::sin(angle)), std::fabs(range2 * std::cos(angle)));
    xpad = std::max(xpad, dx);   // at least one cell
    ypad = std::max(ypad, dy);
    x_pad_fac = std::min(1.0, xpad / lx);
    y_pad_fac = std::min(1.0, ypad / ly);
  }
  *nxpad = padding_size(nx, x_pad_fac);
  *nypad = padding_size(ny, y_pad_fac);
  *true_x_pad_fac = (double)(*nxpad - nx) / (double)nx;
  *true_y_pad_fac = (double)(*nypad - ny) / (double)ny;
}

void estimate_z_padding(int nz, double lz, double min_rel_thick, double z_pad_fac, bool estimate,
                        float default_wavelet_length, int *nzpad, double *true_z_pad_fac) {
  const double min_lz = lz * min_rel_thick;
  if (estimate) {
    const double zpad = (double)default_wavelet_length / 1.0;
    z_pad_fac = std::min(1.0, zpad / min_lz);
  }
  *nzpad = padding_size(nz, z_pad_fac);
  *true_z_pad_fac = (double)(*nzpad - nz) / (double)nz;
}

// ---------------------------------------------------------------------------
// SeismicParametersHolder::computeCircCorrT + makeCircCorrTPosDef
// (src/seismicparametersholder.cpp:431-499).  The reference's steps in float
// -- symmetric circular extension, r2c, |.| in float with the lowest bins
// zeroed, unnormalised c2r, scaling by 1/c(0) -- with each transform
// evaluated in double and rounded once to float.

bool circ_corr_t(const float *prior, int n, int min_int_fq, int nzp, float *out) {
#pragma clang fp contract(off)
  const int cn = nzp / 2 + 1;
  std::vector<double> c(nzp);
  for (int k = 0; k < nzp; ++k) {
    const int refk = k < nzp / 2 + 1 ? k : nzp - k;
    c[k] = refk < n ? (double)prior[refk] : 0.0;
  }
  const double w = 2.0 * 3.14159265358979323846 / (double)nzp;
  std::vector<float> mag(cn);
  for (int k = 0; k < cn; ++k) {
    double sr = 0.0, si = 0.0;
    for (int t = 0; t < nzp; ++t) {
      const double a = w * (double)(((long long)k * t) % nzp);
      sr += c[t] * std::cos(a);
      si -= c[t] * std::sin(a);
    }
    const float re = (float)sr, im = (float)si;
    const float p2 = re * re + im * im;
    mag[k] = k <= min_int_fq ? 0.0f : std::sqrt(p2);
  }
  // c2r of a real, even half spectrum: bins 1..cn-1 twice, except an even
  // length's Nyquist bin
  std::vector<float> r(nzp);
  for (int t = 0; t < nzp; ++t) {
    double v = (double)mag[0];
    for (int k = 1; k < cn; ++k) {
      const double m = (double)mag[k] * std::cos(w * (double)(((long long)k * t) % nzp));
      v += (nzp % 2 == 0 && k == nzp / 2) ? m : 2.0 * m;
    }
    r[t] = (float)v;
  }
  if (!(r[0] > 1.e-5f)) return false;
  const float scale = 1.0f / r[0];
  for (int k = 0; k < nzp; ++k) out[k] = r[k] * scale;
  return true;
}

// ---------------------------------------------------------------------------
// SeismicParametersHolder::setCorrelationParameters' robustification of
// Sigma0 (src/seismicparametersholder.cpp:114-135).

namespace {

// eigenvalues of a symmetric 3x3 matrix (cyclic Jacobi), for the condition
// test only
void sym3_eigenvalues(const double v[3][3], double ev[3]) {
  double m[3][3];
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) m[a][b] = 0.5 * (v[a][b] + v[b][a]);
  for (int sweep = 0; sweep < 50; ++sweep) {
    const double off = m[0][1] * m[0][1] + m[0][2] * m[0][2] + m[1][2] * m[1][2];
    if (off <= 1e-300) break;
    for (int p = 0; p < 2; ++p)
      for (int q = p + 1; q < 3; ++q) {
        if (m[p][q] == 0.0) continue;
        const double th = 0.5 * (m[q][q] - m[p][p]) / m[p][q];
        const double t = (th >= 0 ? 1.0 : -1.0) / (std::fabs(th) + std::sqrt(th * th + 1.0));
        const double c = 1.0 / std::sqrt(t * t + 1.0), sn = t * c;
        for (int k = 0; k < 3; ++k) {
          const double kp = m[k][p], kq = m[k][q];
          m[k][p] = c * kp - sn * kq;
          m[k][q] = sn * kp + c * kq;
        }
        for (int k = 0; k < 3; ++k) {
          const double pk = m[p][k], qk = m[q][k];
          m[p][k] = c * pk - sn * qk;
          m[q][k] = sn * pk + c * qk;
        }
      }
  }
  for (int a = 0; a < 3; ++a) ev[a] = m[a][a];
}

}  // namespace

int prior_var0_robustify(const double in[9], double out[9]) {
  double m[3][3], wr[3], vr[3][3];
  bool diag = true;
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      m[r][c] = in[r * 3 + c];
      out[r * 3 + c] = in[r * 3 + c];
      if (r != c && m[r][c] != 0.0) diag = false;
    }
  bool have_vectors = dgeev3(m, wr, vr);
  if (!have_vectors && diag) {
    // DGEEV on a diagonal matrix: DGEBAL isolates every eigenvalue without
    // exchanges, so the eigenvalues are the diagonal in order and the
    // eigenvectors the unit vectors
    for (int r = 0; r < 3; ++r) {
      wr[r] = m[r][r];
      for (int c = 0; c < 3; ++c) vr[r][c] = r == c ? 1.0 : 0.0;
    }
    have_vectors = true;
  }
  if (!have_vectors) sym3_eigenvalues(m, wr);
  double mx = wr[0];
  for (int i = 1; i < 3; ++i)
    if (mx < wr[i]) mx = wr[i];
  bool lift = false;
  for (int k = 0; k < 3; ++k) lift = lift || wr[k] < mx * 0.001;
  if (!lift) return 0;
  if (!have_vectors) return -1;
  // the reference forms the update from eVec(k, i) eVec(k, j): row k of the
  // eigenvector matrix (component k of every eigenvector), kept as written
  for (int k = 0; k < 3; ++k) {
    if (!(wr[k] < mx * 0.001)) continue;
    for (int i = 0; i < 3; ++i)
      for (int j = 0; j < 3; ++j) out[i * 3 + j] += vr[k][i] * vr[k][j] * (mx * 0.0011 - wr[k]);
  }
  return 1;
}

// ---------------------------------------------------------------------------
// The reference's sequential noise: RandomGen (libs/lib/random.cpp:38-119,
// constants libs/lib/random.h:11-14) driving FFTGrid::fillInComplexNoise
// (src/fftgrid.cpp:1214-1263).

namespace {

struct RefRandom {
  uint32_t seed;
  double unif01() {                        // 32-bit LCG, multiplier 69069, increment 1
    seed = 69069u * seed + 1u;
    return (double)seed * (1.0 / 4294967296.0);
  }
  static double g(double x) {              // majorising difference of the Marsaglia-Bray method
    const double a = 17.49731196, b = 2.36785163, c = 2.15787544;
    const double ax = std::fabs(x);
    double r = a * std::exp(-x * x / 2.0);
    if (ax < 1.0) r -= 2.0 * b * (3.0 - x * x) + c * (1.5 - ax);
    else if (ax < 1.5) r -= b * (3.0 - ax) * (3.0 - ax) + c * (1.5 - ax);
    else r -= b * (3.0 - ax) * (3.0 - ax);
    return r;
  }
  double rnorm01() {                       // Marsaglia-Bray (Ripley, Stochastic Simulation, p. 84)
    const double u = unif01();
    if (u < 0.8638) {
      const double u1 = unif01(), u2 = unif01(), u3 = unif01();
      return 2.0 * (u1 + u2 + u3) - 3.0;
    }
    if (u < 0.9745) {
      const double u1 = unif01(), u2 = unif01();
      return 1.5 * (u1 + u2 - 1.0);
    }
    if (u < 0.9973002039) {
      double x;
      for (;;) {
        const double u1 = unif01(), u2 = unif01();
        x = 6.0 * u1 - 3.0;
        if (!(0.358 * u2 > g(x))) return x;
      }
    }
    for (;;) {
      double v1, v2, w;
      do {
        const double u1 = unif01(), u2 = unif01();
        v1 = 2.0 * u1 - 1.0;
        v2 = 2.0 * u2 - 1.0;
        w = v1 * v1 + v2 * v2;
      } while (w >= 1.0);
      const double c = std::sqrt((9.0 - 2.0 * std::log(w)) / w);
      const double s = c * v1, t = c * v2;
      if (!((std::fabs(s) <= 3.0) && (std::fabs(t) <= 3.0))) return std::fabs(s) > 3.0 ? s : t;
    }
  }
};

}  // namespace

void reference_complex_noise(int nxp, int nyp, int nzp, uint32_t *seed, float *re_im) {
  RefRandom rg{*seed};
  const int cnxp = nxp / 2 + 1;
  const long long csize = (long long)cnxp * nyp * nzp;
  const float sd = float(1 / std::sqrt(2.0));
  for (long long i = 0; i < csize; ++i) {
    const int xs = (int)(i % cnxp);
    // x planes 0 and (when the index is odd, as the reference tests it) cnxp - 1
    if (xs == 0 || (xs == cnxp - 1 && (i % 2) == 1)) {
      const long long jk = (i - xs) / cnxp;
      const int j = (int)(jk % nyp), k = (int)(jk / nyp);
      const int jc = j == 0 ? 0 : nyp - j, kc = k == 0 ? 0 : nzp - k;
      const long long jkc = jc + (long long)kc * nyp;
      if (jkc == jk) {                       // its own conjugate: real
        re_im[2 * i] = float(rg.rnorm01());
        re_im[2 * i + 1] = 0;
      } else if (jkc > jk) {                 // first of a conjugate pair
        re_im[2 * i] = float(sd * rg.rnorm01());
        re_im[2 * i + 1] = float(sd * rg.rnorm01());
      } else {                               // conjugate of an earlier point
        const long long cc = jkc * cnxp + xs;
        re_im[2 * i] = re_im[2 * cc];
        re_im[2 * i + 1] = -re_im[2 * cc + 1];
      }
    } else {
      re_im[2 * i] = float(sd * rg.rnorm01());
      re_im[2 * i + 1] = float(sd * rg.rnorm01());
    }
  }
  *seed = rg.seed;
}

}  // namespace crava
