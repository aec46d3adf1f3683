// Host-side per-k preparation (see tables.hpp).
#include "tables.hpp"

#include <cmath>
#include <vector>

namespace crava {

namespace {

struct c64 {
  float re, im;
};

// Half spectrum (nzp/2+1 bins) of a real series, forward sign, double
// accumulation with exactly reduced twiddle angles.
std::vector<c64> half_spectrum(const float *w, int nzp) {
  const int cn = nzp / 2 + 1;
  std::vector<double> cs(nzp), sn(nzp);
  for (int m = 0; m < nzp; ++m) {
    const long double a = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)nzp;
    cs[m] = (double)cosl(a);
    sn[m] = (double)sinl(a);
  }
  std::vector<c64> out(cn);
  for (int k = 0; k < cn; ++k) {
    double re = 0.0, im = 0.0;
    for (int t = 0; t < nzp; ++t) {
      const int m = (int)(((long long)k * t) % nzp);
      re += (double)w[t] * cs[m];
      im -= (double)w[t] * sn[m];
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

}  // namespace crava
