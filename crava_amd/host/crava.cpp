// Crava (posterior core) of the C++ host mirror (see crava_host.hpp).
#include <algorithm>
#include <cmath>
#include <cstring>

#include "crava_host.hpp"

namespace crava_hip {

Crava::Crava(Context *ctx, const Settings &s, std::vector<FFTGrid *> seisData)
    : ctx_(ctx), s_(s), seisData_(std::move(seisData)) {
  const int nt = s_.ntheta;
  if (nt < 1 || nt > 8 || (int)seisData_.size() != nt || (int)s_.A.size() != 3 * nt ||
      (int)s_.wavelets.size() != nt * s_.nzp || (int)s_.waveletNorm.size() != nt || (int)s_.snr.size() != nt ||
      (int)s_.thetaDeg.size() != nt)
    throw Error(CRAVA_HIP_E_INVALID, "inconsistent Crava settings");
  dataVariance_.assign(nt, 0.0f);
  errorVariance_.assign(nt, 0.0f);
  errThetaCov_.assign((size_t)nt * nt, 0.0);
}

Crava::~Crava() {
  if (traceCov_) (void)crava_hip_grid_free(traceCov_);
}

void Crava::computeDataVariance() {
  for (int l = 0; l < s_.ntheta; ++l) {
    FFTGrid *g = seisData_[l];
    long long n = 0;
    check(crava_hip_data_variance(ctx_->handle(), g->devicePtr(), s_.nx, s_.ny, s_.nz, s_.nxp, s_.nyp, s_.nzp,
                                  &dataVariance_[l], &n),
          "computeDataVariance");
  }
}

// src/crava.cpp:366-401 without local noise scales.  S/N < 1.1 is an error
// (the reference exits).
void Crava::setupErrorCorrelation(const std::function<double(float)> &angularCorr) {
  const int nt = s_.ntheta;
  for (int l = 0; l < nt; ++l) {
    if (s_.snr[l] < 1.1f) throw Error(CRAVA_HIP_E_INVALID, "signal-to-noise ratio below 1.1");
    errorVariance_[l] = dataVariance_[l] / s_.snr[l];
  }
  for (int i = 0; i < nt; ++i)
    for (int j = 0; j < nt; ++j) {
      const float dTheta = s_.thetaDeg[i] - s_.thetaDeg[j];
      errThetaCov_[(size_t)i * nt + j] = static_cast<float>(std::sqrt(errorVariance_[i]) *
                                                            std::sqrt(errorVariance_[j]) * angularCorr(dTheta));
    }
}

void Crava::computeVariances(const std::vector<float> &corrT, const SeismicParametersHolder &sph,
                             const std::function<double(float)> &angularCorr) {
  const int nt = s_.ntheta, nzp = s_.nzp;
  if ((int)corrT.size() < nzp) throw Error(CRAVA_HIP_E_INVALID, "computeVariances needs nzp values of corrT");
  if (!s_.matchEnergies.empty() && (int)s_.matchEnergies.size() != nt)
    throw Error(CRAVA_HIP_E_INVALID, "matchEnergies needs one flag per angle");
  computeDataVariance();
  setupErrorCorrelation(angularCorr);
  std::vector<float> gain(nt);
  modelVariance_.assign(nt, 0.0f);
  signalVariance_.assign(nt, 0.0f);
  theoSNRatio_.assign(nt, 0.0f);
  check(crava_hip_wavelet_energy(nt, nzp, s_.wavelets.data(), s_.A.data(), &sph.getPriorVar0()[0][0], corrT.data(),
                                 errorVariance_.data(), s_.snr.data(),
                                 s_.matchEnergies.empty() ? nullptr : s_.matchEnergies.data(), gain.data(),
                                 modelVariance_.data(), signalVariance_.data(), theoSNRatio_.data()),
        "computeVariances");
  // Wavelet::scale (src/wavelet.cpp:477-486): the samples times the gain;
  // norm_ keeps its value, so the error multipliers that divide by the norm
  // (fillkWNorm, :1067-1071) see the scaled wavelet over the unscaled norm
  for (int l = 0; l < nt; ++l) {
    if (gain[l] == 1.0f) continue;
    for (int k = 0; k < nzp; ++k) s_.wavelets[(size_t)l * nzp + k] *= gain[l];
  }
}

long long Crava::computePostMeanResidAndFFTCov(SeismicParametersHolder &sph) {
  const int nt = s_.ntheta, nzp = s_.nzp;
  crava_hip_ctx *h = ctx_->handle();
  // per-k vectors (src/crava.cpp:938-1083)
  std::vector<unsigned char> invert(nzp);
  std::vector<float> kWD((size_t)nzp * nt * 2), e1(kWD.size()), e2(kWD.size()), e3(kWD.size());
  const double lz = (double)s_.nz * s_.dz;
  if (s_.relThick.empty()) {
    check(crava_hip_perk_tables(nt, s_.nz, nzp, lz, s_.lowCut, s_.highCut, 1.0f, s_.wavelets.data(),
                                s_.waveletNorm.data(), invert.data(), kWD.data(), e1.data(), e2.data(), e3.data()),
          "perk tables");
  } else {
    double minRel = s_.minRelThick;
    if (!(minRel > 0.0)) minRel = *std::min_element(s_.relThick.begin(), s_.relThick.end());
    check(crava_hip_perk_tables_nonconst(nt, s_.nz, nzp, lz, s_.lowCut, s_.highCut, (float)minRel,
                                         s_.wavelets.data(), s_.waveletNorm.data(), invert.data(), kWD.data(),
                                         e1.data(), e2.data(), e3.data()),
          "perk tables (non-constant thickness)");
  }
  crava_hip_post_params p;
  std::memset(&p, 0, sizeof(p));
  p.nxp = s_.nxp;
  p.nyp = s_.nyp;
  p.nzp = nzp;
  p.ntheta = nt;
  p.wnc = s_.wnc;
  const double(&v0)[3][3] = sph.getPriorVar0();
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) p.sigma0[a * 3 + b] = v0[a][b];
  for (int i = 0; i < 3 * nt; ++i) p.A[i] = s_.A[i];
  for (int i = 0; i < nt * nt; ++i) p.err_theta_cov[i] = errThetaCov_[i];
  p.mixed_threshold = s_.mixedThreshold;
  crava_hip_post *post = nullptr;
  check(crava_hip_post_create(h, &p, invert.data(), kWD.data(), e1.data(), e2.data(), e3.data(), &post),
        "crava_hip_post_create");

  // error correlation grid (src/crava.cpp:154-157,1012)
  FFTGrid errCorr(ctx_, s_.nx, s_.ny, s_.nz, s_.nxp, s_.nyp, nzp);
  errCorr.setType(FFTGrid::COVARIANCE);
  errCorr.setTransformPrecision(sph.GetCovAlpha()->getTransformPrecision());
  errCorr.createRealGrid();
  long long nfail = 0;
  try {
    errCorr.fillInErrCorr(s_.corrXY, s_.corrGradI, s_.corrGradJ);
    // forward transforms: data and means share the 1/sqrt(N) scaling, one set
    std::vector<void *> dm;
    for (FFTGrid *g : seisData_)
      if (!g->getIsTransformed()) dm.push_back(g->devicePtr());
    for (int c = 0; c < 3; ++c)
      if (!sph.mu(c)->getIsTransformed()) dm.push_back(sph.mu(c)->devicePtr());
    if (!dm.empty())
      check(crava_hip_fft_r2c_many(h, dm.data(), (int)dm.size(), s_.nxp, s_.nyp, nzp, CRAVA_HIP_DATA, CRAVA_HIP_F32),
            "forward transforms");
    for (FFTGrid *g : seisData_) g->setIsTransformed(true);
    for (int c = 0; c < 3; ++c) sph.mu(c)->setIsTransformed(true);
    sph.FFTCovGrids();
    errCorr.fftInPlace();

    void *mu[3], *data[8], *cov[6];
    for (int c = 0; c < 3; ++c) mu[c] = sph.mu(c)->devicePtr();
    for (int l = 0; l < nt; ++l) data[l] = seisData_[l]->devicePtr();
    for (int q = 0; q < 6; ++q) cov[q] = sph.cov(q)->devicePtr();
    long long n64 = 0;
    check(crava_hip_post_run(post, s_.precision, mu, data, cov, errCorr.devicePtr(), &nfail, &n64), "posterior");
    nFP64_ = n64;

    // posterior mean and residuals back to the time domain (src/crava.cpp:1181-1201)
    std::vector<void *> back;
    for (int c = 0; c < 3; ++c) back.push_back(mu[c]);
    for (int l = 0; l < nt; ++l) back.push_back(data[l]);
    check(crava_hip_fft_c2r_many(h, back.data(), (int)back.size(), s_.nxp, s_.nyp, nzp, CRAVA_HIP_DATA,
                                 CRAVA_HIP_F32),
          "inverse transforms");
    for (int c = 0; c < 3; ++c) sph.mu(c)->setIsTransformed(false);
    for (FFTGrid *g : seisData_) g->setIsTransformed(false);
  } catch (...) {
    crava_hip_post_destroy(post);
    throw;
  }
  crava_hip_post_destroy(post);
  ctx_->synchronize();
  return nfail;
}

long long Crava::simulate(SeismicParametersHolder &sph, unsigned long long seed, unsigned long long stream,
                          FFTGrid *real[3]) {
  void *seeds[3];
  for (int c = 0; c < 3; ++c) {
    seeds[c] = real[c]->devicePtr();
    check(crava_hip_fill_complex_noise(ctx_->handle(), seeds[c], s_.nxp, s_.nyp, s_.nzp, seed, stream * 3 + c),
          "noise");
  }
  return realise(sph, real);
}

long long Crava::simulate(SeismicParametersHolder &sph, uint32_t &randomGenSeed, FFTGrid *real[3]) {
  for (int c = 0; c < 3; ++c) real[c]->fillInComplexNoise(randomGenSeed);
  return realise(sph, real);
}

// :1482-1573: the per-point step, the inverse transforms, the local-noise
// factor when correctAlphaBetaRho has run (:1547-1566), then the means.
long long Crava::realise(SeismicParametersHolder &sph, FFTGrid *real[3]) {
  crava_hip_ctx *h = ctx_->handle();
  void *seeds[3], *cov[6], *mean[3];
  for (int c = 0; c < 3; ++c) {
    seeds[c] = real[c]->devicePtr();
    mean[c] = (traceCov_ || sph.mu(c)->getIsTransformed()) ? nullptr : sph.mu(c)->devicePtr();
  }
  for (int q = 0; q < 6; ++q) {
    if (!sph.cov(q)->getIsTransformed()) throw Error(CRAVA_HIP_E_STATE, "simulate needs transformed covariances");
    cov[q] = sph.cov(q)->devicePtr();
  }
  long long nfail = 0;
  check(crava_hip_simulate(h, s_.nxp, s_.nyp, s_.nzp, cov, seeds, mean, &nfail), "simulate");
  if (traceCov_) {
    check(crava_hip_local_noise_realisation(h, s_.nx, s_.ny, s_.nz, s_.nxp, s_.nyp, s_.nzp,
                                            static_cast<const double *>(traceCov_), seeds),
          "local noise realisation");
    for (int c = 0; c < 3; ++c)
      if (!sph.mu(c)->getIsTransformed())
        check(crava_hip_grid_op(h, CRAVA_HIP_OP_ADD, seeds[c], sph.mu(c)->devicePtr(), s_.nxp, s_.nyp, s_.nzp, 0,
                                0.0f),
              "add mean");
  }
  for (int c = 0; c < 3; ++c) real[c]->setIsTransformed(false);
  return nfail;
}

void Crava::correctAlphaBetaRho(const double priorVar0[3][3], const double postVar0[3][3], FFTGrid *post[3],
                                FFTGrid *mean2[3], const double *G) {
  const int nt = s_.ntheta;
  if ((int)s_.localNoiseScale.size() != nt) throw Error(CRAVA_HIP_E_INVALID, "one noise-scale table per angle");
  for (const auto &t : s_.localNoiseScale)
    if (t.size() != (size_t)s_.nx * s_.ny) throw Error(CRAVA_HIP_E_INVALID, "noise-scale tables are [ny*nx]");
  for (int c = 0; c < 3; ++c)
    if (post[c]->getIsTransformed() || mean2[c]->getIsTransformed())
      throw Error(CRAVA_HIP_E_STATE, "correctAlphaBetaRho works on time-domain grids");
  double sm[9], smd[9];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      sm[r * 3 + c] = priorVar0[r][c];
      smd[r * 3 + c] = postVar0[r][c];
    }
  std::vector<double> g((size_t)nt * 3);
  if (G)
    std::copy(G, G + g.size(), g.begin());
  else
    check(crava_hip_local_noise_g(nt, errThetaCov_.data(), sm, smd, g.data()), "computeG");
  if (!traceCov_) {
    // storage for ny*nx 3x3 double matrices: rows of 18 doubles hold two
    const int ntr = s_.nx * s_.ny;
    check(crava_hip_grid_alloc(ctx_->handle(), 16, 1, (ntr + 1) / 2, CRAVA_HIP_F64, &traceCov_), "alloc");
  }
  std::vector<const double *> tabs(nt);
  for (int a = 0; a < nt; ++a) tabs[a] = s_.localNoiseScale[a].data();
  crava_hip_local_noise p{};
  p.nx = s_.nx;
  p.ny = s_.ny;
  p.nz = s_.nz;
  p.nxp = s_.nxp;
  p.nyp = s_.nyp;
  p.nzp = s_.nzp;
  p.ntheta = nt;
  p.G = g.data();
  p.errThetaCov = errThetaCov_.data();
  p.priorVar0 = sm;
  p.postVar0 = smd;
  p.noiseScale = tabs.data();
  void *pg[3];
  const void *mg[3];
  for (int c = 0; c < 3; ++c) {
    pg[c] = post[c]->devicePtr();
    mg[c] = mean2[c]->devicePtr();
  }
  check(crava_hip_local_noise_correct(ctx_->handle(), &p, pg, mg, static_cast<double *>(traceCov_)),
        "correctAlphaBetaRho");
}

void Crava::computeSyntSeismic(FFTGrid *alpha, FFTGrid *beta, FFTGrid *rho, const std::vector<FFTGrid *> &syn) {
  const int nt = s_.ntheta;
  if ((int)syn.size() != nt) throw Error(CRAVA_HIP_E_INVALID, "computeSyntSeismic needs one grid per angle");
  const bool fftDomain = alpha->getIsTransformed();
  if (fftDomain) {
    alpha->invFFTInPlace();
    beta->invFFTInPlace();
    rho->invFFTInPlace();
  }
  const size_t ntr = (size_t)s_.nx * s_.ny;
  auto table = [&](const std::vector<std::vector<float>> &t, std::vector<const float *> &ptrs) -> const float *const * {
    if (t.empty()) return nullptr;
    if ((int)t.size() != nt) throw Error(CRAVA_HIP_E_INVALID, "local wavelet tables need one entry per angle");
    ptrs.assign(nt, nullptr);
    for (int l = 0; l < nt; ++l) {
      if (t[l].empty()) continue;
      if (t[l].size() != ntr) throw Error(CRAVA_HIP_E_INVALID, "local wavelet table of the wrong size");
      ptrs[l] = t[l].data();
    }
    return ptrs.data();
  };
  std::vector<const float *> pshift, pgain, pstretch;
  crava_hip_synt_params p{};
  p.nx = s_.nx;
  p.ny = s_.ny;
  p.nz = s_.nz;
  p.nxp = s_.nxp;
  p.nyp = s_.nyp;
  p.nzp = s_.nzp;
  p.ntheta = nt;
  p.dz = s_.dz;
  p.A = s_.A.data();
  p.wavelets = s_.wavelets.data();
  p.shift = table(s_.localShift, pshift);
  p.gain = table(s_.localGain, pgain);
  p.stretch = table(s_.stretch, pstretch);
  std::vector<void *> out(nt);
  for (int l = 0; l < nt; ++l) out[l] = syn[l]->devicePtr();
  check(crava_hip_synthetic_seismic(ctx_->handle(), &p, alpha->devicePtr(), beta->devicePtr(), rho->devicePtr(),
                                    out.data()),
        "synthetic seismic");
  for (int l = 0; l < nt; ++l) syn[l]->setIsTransformed(false);
  if (fftDomain) {
    alpha->fftInPlace();
    beta->fftInPlace();
    rho->fftInPlace();
  }
}

void Crava::divideDataByScaleWavelet(const SeismicParametersHolder &sph) {
  const int nt = s_.ntheta;
  if ((long long)s_.relThick.size() != (long long)s_.nxp * s_.nyp)
    throw Error(CRAVA_HIP_E_INVALID, "divideDataByScaleWavelet needs relThick for every padded trace");
  for (int l = 0; l < nt; ++l)
    if (seisData_[l]->getIsTransformed()) throw Error(CRAVA_HIP_E_STATE, "seismic grids must be in the time domain");
  const std::vector<float> corrT = sph.getPriorCorrTFiltered(s_.nz, s_.nzp);
  std::vector<float> errVar(nt);
  for (int l = 0; l < nt; ++l) errVar[l] = (float)errThetaCov_[(size_t)l * nt + l];
  const size_t ntr = (size_t)s_.nx * s_.ny;
  auto table = [&](const std::vector<std::vector<float>> &t, std::vector<const float *> &ptrs) -> const float *const * {
    if (t.empty()) return nullptr;
    if ((int)t.size() != nt) throw Error(CRAVA_HIP_E_INVALID, "local wavelet tables need one entry per angle");
    ptrs.assign(nt, nullptr);
    for (int l = 0; l < nt; ++l) {
      if (t[l].empty()) continue;
      if (t[l].size() != ntr) throw Error(CRAVA_HIP_E_INVALID, "local wavelet table of the wrong size");
      ptrs[l] = t[l].data();
    }
    return ptrs.data();
  };
  std::vector<const float *> pshift, pgain;
  crava_hip_swdiv_params p{};
  p.nx = s_.nx;
  p.ny = s_.ny;
  p.nz = s_.nz;
  p.nxp = s_.nxp;
  p.nyp = s_.nyp;
  p.nzp = s_.nzp;
  p.ntheta = nt;
  p.wavelets = s_.wavelets.data();
  p.wavelet_norm = s_.waveletNorm.data();
  p.A = s_.A.data();
  p.prior_var0 = &sph.getPriorVar0()[0][0];
  p.corr_t = corrT.data();
  p.err_var = errVar.data();
  p.rel_thick = s_.relThick.data();
  p.shift = table(s_.localShift, pshift);
  p.gain = table(s_.localGain, pgain);
  p.dz = s_.dz;
  p.lz = (double)s_.nz * s_.dz;
  p.low_cut = s_.lowCut;
  p.high_cut = s_.highCut;
  p.wnc = s_.wnc;
  std::vector<void *> grids(nt);
  for (int l = 0; l < nt; ++l) grids[l] = seisData_[l]->devicePtr();
  check(crava_hip_divide_by_scale_wavelet(ctx_->handle(), &p, grids.data()), "divideDataByScaleWavelet");
  for (int l = 0; l < nt; ++l) seisData_[l]->interpolateSeismic(s_.energyThreshold);
}

}  // namespace crava_hip
