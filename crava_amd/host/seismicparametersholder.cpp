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
  if (priorCorrT.empty()) throw Error(CRAVA_HIP_E_INVALID, "priorCorrT must not be empty");
  std::vector<float> c(nzp);
  check(crava_hip_circ_corr_t(priorCorrT.data(), (int)priorCorrT.size(), minIntFq, nzp, c.data()),
        "computeCircCorrT");
  return c;
}

// :463-499 on a circularly extended correlation (c[k] = c[nzp - k], as
// computeCircCorrT builds it): the C ABI re-extends its first half.
void SeismicParametersHolder::makeCircCorrTPosDef(std::vector<float> &c, int minIntFq, int nzp) const {
  if ((int)c.size() < nzp) throw Error(CRAVA_HIP_E_INVALID, "makeCircCorrTPosDef: fewer than nzp values");
  std::vector<float> out(nzp);
  check(crava_hip_circ_corr_t(c.data(), nzp / 2 + 1, minIntFq, nzp, out.data()), "makeCircCorrTPosDef");
  for (int k = 0; k < nzp; ++k) c[k] = out[k];
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
