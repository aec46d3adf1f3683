// SeismicParametersHolder of the C++ host mirror (see crava_host.hpp).
#include <cmath>

#include "crava_host.hpp"

#include <algorithm>

namespace crava_hip {

SeismicParametersHolder::~SeismicParametersHolder() {
  for (FFTGrid *g : mu_) delete g;
  for (FFTGrid *g : cov_) delete g;
}

void SeismicParametersHolder::allocateGrids(int nx, int ny, int nz, int nxp, int nyp, int nzp) {
  for (int c = 0; c < 3; ++c) {
    delete mu_[c];
    mu_[c] = new FFTGrid(ctx_, nx, ny, nz, nxp, nyp, nzp);
    mu_[c]->setType(FFTGrid::PARAMETER);
    mu_[c]->createRealGrid();
    mu_[c]->fillInConstant(0.0f);
  }
  for (int q = 0; q < 6; ++q) {
    delete cov_[q];
    cov_[q] = new FFTGrid(ctx_, nx, ny, nz, nxp, nyp, nzp);
    cov_[q]->setType(FFTGrid::COVARIANCE);
    cov_[q]->createRealGrid();
    cov_[q]->fillInConstant(0.0f);
  }
}


// setCorrelationParameters (src/seismicparametersholder.cpp:92-149): Sigma0
// with the reference's lift of eigenvalues below 0.1 % of the largest
// (crava_hip_prior_var0_robustify); a lift that library cannot reproduce
// throws.
void SeismicParametersHolder::setPriorVar0(const double v[3][3]) {
  double out[9];
  int changed = 0;
  check(crava_hip_prior_var0_robustify(&v[0][0], out, &changed), "setPriorVar0");
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) priorVar0_[a][b] = out[a * 3 + b];
}

void SeismicParametersHolder::setCovTransformPrecision(int precision) {
  for (FFTGrid *g : cov_)
    if (g) g->setTransformPrecision(precision);
}

// Real spectrum of a length-n real sequence by a direct DFT in double
// (n is a padded trace length, a few hundred; this runs once per inversion).
static void dft_real(const std::vector<double> &x, std::vector<double> &re, std::vector<double> &im) {
  const int n = (int)x.size();
  const int cn = n / 2 + 1;
  re.assign(cn, 0.0);
  im.assign(cn, 0.0);
  const double w = -2.0 * 3.14159265358979323846 / n;
  for (int k = 0; k < cn; ++k) {
    double sr = 0.0, si = 0.0;
    for (int t = 0; t < n; ++t) {
      const long long m = ((long long)k * t) % n;
      sr += x[t] * std::cos(w * (double)m);
      si += x[t] * std::sin(w * (double)m);
    }
    re[k] = sr;
    im[k] = si;
  }
}

float SeismicParametersHolder::getOrigin(FFTGrid *grid) const {
  if (!grid || grid->getIsTransformed()) throw Error(CRAVA_HIP_E_STATE, "getOrigin needs a grid in the time domain");
  return grid->readTrace(0, 0, 0, 1)[0];
}

void SeismicParametersHolder::updatePriorVar() {
  priorVar0_[0][0] = getOrigin(cov_[0]);
  priorVar0_[1][1] = getOrigin(cov_[1]);
  priorVar0_[2][2] = getOrigin(cov_[2]);
  priorVar0_[0][1] = getOrigin(cov_[3]);
  priorVar0_[1][0] = priorVar0_[0][1];
  priorVar0_[2][0] = getOrigin(cov_[4]);
  priorVar0_[0][2] = priorVar0_[2][0];
  priorVar0_[2][1] = getOrigin(cov_[5]);
  priorVar0_[1][2] = priorVar0_[2][1];
}

std::vector<float> SeismicParametersHolder::createPostCov00(FFTGrid *postCov) const {
  if (!postCov || postCov->getIsTransformed())
    throw Error(CRAVA_HIP_E_STATE, "createPostCov00 needs a grid in the time domain");
  return postCov->readTrace(0, 0, 0, postCov->getNz());
}

std::vector<float> SeismicParametersHolder::extractParamCorrFromCovAlpha(int nzp) const {
  FFTGrid *c = cov_[0];
  if (!c || c->getIsTransformed())
    throw Error(CRAVA_HIP_E_STATE, "extractParamCorrFromCovAlpha needs covAlpha in the time domain");
  std::vector<float> circ = c->readTrace(0, 0, 0, nzp);
  const float constant = circ[0];
  for (int k = 0; k < nzp; ++k) circ[k] = circ[k] / constant;
  return circ;
}

std::vector<float> SeismicParametersHolder::getPriorCorrTFiltered(int nz, int nzp) const {
  const std::vector<float> circ = extractParamCorrFromCovAlpha(nzp);
  std::vector<float> filtered(nzp);
  for (int i = 0; i < nzp; ++i) {
    const int refk = i < nzp / 2 + 1 ? i : nzp - i;
    filtered[i] = refk < nz ? circ[refk] : 0.0f;
  }
  return filtered;
}

std::vector<float> SeismicParametersHolder::computeCircCorrT(const std::vector<float> &priorCorrT, int minIntFq,
                                                             int nzp) const {
  if (priorCorrT.empty() || priorCorrT[0] == 0.0f) throw Error(CRAVA_HIP_E_INVALID, "priorCorrT[0] must be nonzero");
  const int n = (int)priorCorrT.size();
  std::vector<float> c(nzp);
  for (int k = 0; k < nzp; ++k) {
    const int refk = k < nzp / 2 + 1 ? k : nzp - k;   // symmetric circular extension
    c[k] = refk < n ? priorCorrT[refk] : 0.0f;
  }
  makeCircCorrTPosDef(c, minIntFq, nzp);
  return c;
}

// :463-499: |spectrum| with the lowest frequencies (k <= minIntFq) removed,
// back to a correlation normalised to 1 at lag 0.
void SeismicParametersHolder::makeCircCorrTPosDef(std::vector<float> &c, int minIntFq, int nzp) const {
  std::vector<double> x(c.begin(), c.begin() + nzp), re, im;
  dft_real(x, re, im);
  const int cn = nzp / 2 + 1;
  std::vector<double> mag(cn);
  for (int k = 0; k < cn; ++k) mag[k] = k <= minIntFq ? 0.0 : std::sqrt(re[k] * re[k] + im[k] * im[k]);
  // unnormalised inverse of the real, even spectrum (ifft(fft(f)) = n f, as
  // the reference's 1-D transforms): c(t) = sum_k m_k e^{+2 pi i k t/n}
  std::vector<double> out(nzp, 0.0);
  for (int t = 0; t < nzp; ++t) {
    double s = mag[0];
    for (int k = 1; k < cn; ++k) {
      const bool self = (nzp % 2 == 0) && (k == nzp / 2);
      const double v = mag[k] * std::cos(2.0 * 3.14159265358979323846 * (double)(((long long)k * t) % nzp) / nzp);
      s += self ? v : 2.0 * v;
    }
    out[t] = s;
  }
  const float c0 = (float)out[0];
  if (!(c0 > 1e-5f)) throw Error(CRAVA_HIP_E_INVALID, "circular temporal correlation undefined (increase nz)");
  const float scale = 1.0f / c0;
  for (int k = 0; k < nzp; ++k) c[k] = (float)out[k] * scale;
}

void SeismicParametersHolder::initializeCorrelations(const std::vector<float> &priorCorrXY,
                                                     const std::vector<float> &priorCorrT, float corrGradI,
                                                     float corrGradJ, int lowIntCut, int nzp) {
  if (!cov_[0]) throw Error(CRAVA_HIP_E_STATE, "allocateGrids first");
  const std::vector<float> circ = computeCircCorrT(priorCorrT, lowIntCut, nzp);
  FFTGrid *g0 = cov_[0];
  // one launch fills the six grids: fillInParamCorr then multiplyByScalar(Sigma0_ab)
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  float scale[6];
  void *grids[6];
  for (int q = 0; q < 6; ++q) {
    scale[q] = static_cast<float>(priorVar0_[ia[q]][ib[q]]);
    grids[q] = cov_[q]->devicePtr();
  }
  const int nxp = g0->getNxp(), nyp = g0->getNyp();
  if ((long long)priorCorrXY.size() < (long long)nxp * nyp) throw Error(CRAVA_HIP_E_INVALID, "corrXY too small");
  void *cxy = nullptr, *ct = nullptr;
  check(crava_hip_grid_alloc(ctx_->handle(), 2 * nxp * nyp, 1, 1, CRAVA_HIP_F32, &cxy), "alloc");
  int rc = crava_hip_grid_alloc(ctx_->handle(), 2 * nzp, 1, 1, CRAVA_HIP_F32, &ct);
  if (rc == CRAVA_HIP_OK) rc = crava_hip_grid_upload(ctx_->handle(), cxy, priorCorrXY.data(), sizeof(float) * nxp * nyp);
  if (rc == CRAVA_HIP_OK) rc = crava_hip_grid_upload(ctx_->handle(), ct, circ.data(), sizeof(float) * nzp);
  if (rc == CRAVA_HIP_OK)
    rc = crava_hip_fill_param_corr(ctx_->handle(), grids, 6, scale, nxp, nyp, g0->getNzp(),
                                   static_cast<const float *>(cxy), static_cast<const float *>(ct), corrGradI,
                                   corrGradJ);
  if (rc == CRAVA_HIP_OK) rc = crava_hip_synchronize(ctx_->handle());
  crava_hip_grid_free(cxy);
  crava_hip_grid_free(ct);
  check(rc, "initializeCorrelations");
  for (FFTGrid *g : cov_) g->setIsTransformed(false);
}

void SeismicParametersHolder::transformSet(FFTGrid *const *grids, int n, bool inverse) {
  std::vector<void *> ptr;
  for (int i = 0; i < n; ++i)
    if (grids[i]->getIsTransformed() == inverse) ptr.push_back(grids[i]->devicePtr());
  if (ptr.empty()) return;
  FFTGrid *g = grids[0];
  const int type = g->getType() == FFTGrid::COVARIANCE ? CRAVA_HIP_COVARIANCE : CRAVA_HIP_PARAMETER;
  const int prec = inverse ? CRAVA_HIP_F32 : g->getTransformPrecision();
  if (inverse)
    check(crava_hip_fft_c2r_many(ctx_->handle(), ptr.data(), (int)ptr.size(), g->getNxp(), g->getNyp(), g->getNzp(),
                                 type, prec),
          "invFFT grid set");
  else
    check(crava_hip_fft_r2c_many(ctx_->handle(), ptr.data(), (int)ptr.size(), g->getNxp(), g->getNyp(), g->getNzp(),
                                 type, prec),
          "FFT grid set");
  for (int i = 0; i < n; ++i) grids[i]->setIsTransformed(!inverse);
}

void SeismicParametersHolder::FFTCovGrids() { transformSet(cov_, 6, false); }
void SeismicParametersHolder::invFFTCovGrids() { transformSet(cov_, 6, true); }

void SeismicParametersHolder::FFTAllGrids() {
  transformSet(mu_, 3, false);
  FFTCovGrids();
}

void SeismicParametersHolder::invFFTAllGrids() {
  transformSet(mu_, 3, true);
  invFFTCovGrids();
}

// ---------------------------------------------------------------------------
// State4D (src/state4d.cpp:90-236)
State4D::State4D(Context *ctx, std::vector<FFTGrid *> muStatic, std::vector<FFTGrid *> muDynamic,
                 std::vector<FFTGrid *> sigmaStaticStatic, std::vector<FFTGrid *> sigmaDynamicDynamic,
                 std::vector<FFTGrid *> sigmaStaticDynamic)
    : ctx_(ctx), ms_(std::move(muStatic)), md_(std::move(muDynamic)), ss_(std::move(sigmaStaticStatic)),
      dd_(std::move(sigmaDynamicDynamic)), sd_(std::move(sigmaStaticDynamic)) {
  if (ms_.size() != 3 || md_.size() != 3 || ss_.size() != 6 || dd_.size() != 6 || sd_.size() != 9)
    throw Error(CRAVA_HIP_E_INVALID, "State4D needs 3 + 3 means and 6 + 6 + 9 covariance grids");
}

static void need_transformed(const std::vector<FFTGrid *> &g) {
  for (const FFTGrid *x : g)
    if (!x || !x->getIsTransformed()) throw Error(CRAVA_HIP_E_STATE, "State4D merges transformed grids");
}

void State4D::merge(SeismicParametersHolder &current) {
  need_transformed(ms_);
  need_transformed(md_);
  FFTGrid *mu[3] = {current.GetMuAlpha(), current.GetMuBeta(), current.GetMuRho()};
  void *s[3], *d[3], *o[3];
  for (int l = 0; l < 3; ++l) {
    s[l] = ms_[l]->devicePtr();
    d[l] = md_[l]->devicePtr();
    o[l] = mu[l]->devicePtr();
  }
  const FFTGrid *g = ms_[0];
  check(crava_hip_merge_mean_4d(ctx_->handle(), g->getNxp(), g->getNyp(), g->getNzp(), s, d, o), "merge");
  for (FFTGrid *m : mu) m->setIsTransformed(true);
  mergeCov({current.GetCovAlpha(), current.GetCrCovAlphaBeta(), current.GetCrCovAlphaRho(), current.GetCovBeta(),
            current.GetCrCovBetaRho(), current.GetCovRho()});
}

void State4D::mergeCov(const std::vector<FFTGrid *> &sigma) {
  if (sigma.size() != 6) throw Error(CRAVA_HIP_E_INVALID, "mergeCov needs 6 grids");
  need_transformed(ss_);
  need_transformed(dd_);
  need_transformed(sd_);
  void *a[6], *b[6], *c[9], *o[6];
  for (int q = 0; q < 6; ++q) {
    a[q] = ss_[q]->devicePtr();
    b[q] = dd_[q]->devicePtr();
    o[q] = sigma[q]->devicePtr();
  }
  for (int q = 0; q < 9; ++q) c[q] = sd_[q]->devicePtr();
  const FFTGrid *g = ss_[0];
  check(crava_hip_merge_cov_4d(ctx_->handle(), g->getNxp(), g->getNyp(), g->getNzp(), a, b, c, o), "mergeCov");
  for (FFTGrid *x : sigma) x->setIsTransformed(true);
}

}  // namespace crava_hip
