// C++ host mirror of the reference's FFTGrid / SeismicParametersHolder /
// Crava interface on the hot path, with device-resident grids behind the C
// ABI of include/crava_hip.h (no HIP or torch type appears here).
//
//   FFTGrid                  src/fftgrid.h:22-295
//   SeismicParametersHolder  src/seismicparametersholder.h
//   Crava (posterior path)   src/crava.h; src/crava.cpp:322-473,887-1295,1416-1599
//
// Names, argument meaning and call order follow the reference, so code
// written against it ports by changing the namespace.  Differences that
// follow from the storage living in HBM:
//   * A grid's values live on the GPU.  As for the reference's disk-backed
//     FFTFileGrid, setAccessMode(READ / READANDWRITE / RANDOMACCESS) brings
//     them into a host copy for the cursor and (i,j,k) accessors, and
//     endAccess() writes the host copy back after WRITE / READANDWRITE /
//     RANDOMACCESS.  Transforms, fills and element-wise operators run on the
//     device and need no access mode.
//   * Errors are reported as exceptions (crava_hip::Error) carrying the C
//     ABI's status and message, where the reference asserts or exits.
#pragma once

#include <complex>
#include <cstdint>
#include <functional>
#include <stdexcept>
#include <string>
#include <vector>

#include "../../include/crava_hip.h"

namespace crava_hip {

class Error : public std::runtime_error {
 public:
  Error(int status, const std::string &what) : std::runtime_error(what), status_(status) {}
  int status() const { return status_; }

 private:
  int status_;
};

// Throw Error if rc is not CRAVA_HIP_OK.
void check(int rc, const char *what);

struct fftw_complex_f {   // layout of the reference's fftw_complex (float)
  float re, im;
};

// A crava_hip_ctx: one stream and the transform plans.
class Context {
 public:
  explicit Context(void *stream = nullptr);
  ~Context();
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  crava_hip_ctx *handle() const { return ctx_; }
  void synchronize();

 private:
  crava_hip_ctx *ctx_ = nullptr;
};

class FFTGrid {
 public:
  enum gridTypes { CTMISSING, DATA, PARAMETER, COVARIANCE, VELOCITY };   // src/fftgrid.h:167
  enum accessMode { NONE, READ, WRITE, READANDWRITE, RANDOMACCESS };     // src/fftgrid.h:168

  FFTGrid(Context *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp);
  virtual ~FFTGrid();
  FFTGrid(const FFTGrid &) = delete;
  FFTGrid &operator=(const FFTGrid &) = delete;

  void createRealGrid(bool add = true);
  void createComplexGrid();
  void setType(int cubeType) { cubetype_ = cubeType; }
  int getType() const { return cubetype_; }
  // CRAVA_HIP_F32 (default) or CRAVA_HIP_MIXED: arithmetic of this grid's
  // transforms (float storage either way).
  void setTransformPrecision(int precision) { precision_ = precision; }
  int getTransformPrecision() const { return precision_; }

  int fftInPlace();      // src/fftgrid.cpp:1804-1829
  int invFFTInPlace();   // src/fftgrid.cpp:1832-1862
  bool getIsTransformed() const { return istransformed_; }
  void setIsTransformed(bool v) { istransformed_ = v; }

  void fillInConstant(float value);
  void fillInParamCorr(const std::vector<float> &corrXY, const float *circCorrT, float gradI, float gradJ);
  void fillInErrCorr(const std::vector<float> &corrXY, float gradI, float gradJ);
  // src/fftgrid.cpp:1214-1263 with the reference's RandomGen stream (host,
  // then uploaded); randomGenSeed is the generator state, advanced as the
  // reference advances it.  The grid becomes a transformed PARAMETER grid.
  void fillInComplexNoise(uint32_t &randomGenSeed);

  void multiplyByScalar(float scalar);
  void addScalar(float scalar);
  void add(FFTGrid *g);
  void subtract(FFTGrid *g);
  void multiply(FFTGrid *g);
  void conjugate();
  void changeSign();

  virtual void setAccessMode(int mode);
  virtual void endAccess();
  float getNextReal();
  int setNextReal(float value);
  fftw_complex_f getNextComplex();
  int setNextComplex(fftw_complex_f value);
  float getRealValue(int i, int j, int k, bool extSimbox = false) const;
  // values k0 <= k < k0 + nk of trace (i, j) of the real grid, read from the
  // device without an access mode (crava_hip_grid_read_trace)
  std::vector<float> readTrace(int i, int j, int k0, int nk) const;
  int setRealValue(int i, int j, int k, float value, bool extSimbox = false);
  fftw_complex_f getComplexValue(int i, int j, int k, bool extSimbox = false) const;

  int getNx() const { return nx_; }
  int getNy() const { return ny_; }
  int getNz() const { return nz_; }
  int getNxp() const { return nxp_; }
  int getNyp() const { return nyp_; }
  int getNzp() const { return nzp_; }
  int getRNxp() const { return rnxp_; }
  int getCNxp() const { return cnxp_; }
  long long getrsize() const { return rsize_; }
  long long getcsize() const { return csize_; }

  void *devicePtr() const { return dev_; }
  Context *context() const { return ctx_; }

  // CRAVA binary grid file (src/fftgrid.cpp:2399-2487): a text line
  // "crava_fftgrid_binary", the simbox geometry below, rnxp, nyp, nzp and
  // the rsize padded real values, all big-endian.
  struct CravaHeader {
    double x0 = 0.0, y0 = 0.0, dx = 1.0, dy = 1.0;
    int nx = 0, ny = 0;
    double il0 = 0.0, xl0 = 0.0, ilStepX = 1.0, ilStepY = 0.0, xlStepX = 0.0, xlStepY = 1.0, angle = 0.0;
  };
  // writes fileName + ".crava" (IO::SuffixCrava); throws Error on I/O failure
  void writeCravaFile(const std::string &fileName, const CravaHeader &h) const;
  // reads a file written by writeCravaFile; a failure (unreadable file,
  // other grid dimensions) is appended to errText and leaves the grid as it was
  void readCravaFile(const std::string &fileName, std::string &errText, bool nopadding = false);

  // STORM grid file (src/fftgrid.cpp:2118-2181, header of
  // Simbox::getStormHeader, src/simbox.cpp:408-427): geometry of the header.
  struct StormHeader {
    double xmin = 0.0, lx = 1.0, ymin = 0.0, ly = 1.0, lz = 1.0, angle = 0.0;  // angle in radians
    std::string topName, botName;  // surfaces (flat == false)
  };
  // writes fileName + ".storm" (binary, big-endian floats) or ".txt" (ascii);
  // padding: the padded nxp x nyp x nzp cube instead of nx x ny x nz
  void writeStormFile(const std::string &fileName, const StormHeader &h, bool ascii, bool padding = false,
                      bool flat = false, bool scientific = false) const;

  // Simbox geometry of the SEG-Y and SGRI writers (the Simbox accessors of
  // src/simbox.h): lateral grid and rotation (getx0, gety0, getdx, getdy,
  // getAngle in radians), GetLZ and getdz, getMinMaxZ, the inline /
  // crossline transform (getIL0 ... getXLStepY) and the top and base time
  // surfaces at the trace centres (getTop / getBot, RMISSING where missing),
  // [ny*nx] with i fastest.
  struct SimboxGeometry {
    double xmin = 0.0, ymin = 0.0, dx = 1.0, dy = 1.0, angle = 0.0;
    double lz = 1.0, dz = 1.0, zmin = 0.0, zmax = 1.0;
    double il0 = 0.0, xl0 = 0.0, ilStepX = 1.0, ilStepY = 0.0, xlStepX = 0.0, xlStepY = 1.0;
    std::vector<double> top, bot;
  };
  // SEG-Y file fileName + ".segy" of the volume (FFTGrid::writeSegyFile,
  // src/fftgrid.cpp:2189-2264, with NRLib::SegY): the reference's standard
  // textual header, its binary header, and one trace per (i, j) with a
  // defined top, resampled to z0 + k dz (dz the simbox's rounded to whole
  // ms), sorted by inline and crossline, IBM floats.  traceHeaderFormat:
  // SEISWORKS 0, IESX 1, SIP 2, CHARISMA 3, SIPX 4 (NRLib::TraceHeaderFormat).
  void writeSegyFile(const std::string &fileName, const SimboxGeometry &g, float z0, int traceHeaderFormat = 0) const;
  // SGRI files fileName + ".Sgrh" (text header) and ".Sgri" (big-endian
  // floats) of the volume on a regular time axis from zmin
  // (FFTGrid::writeSgriFile, src/fftgrid.cpp:2325-2397).
  void writeSgriFile(const std::string &fileName, const SimboxGeometry &g, const std::string &label) const;

  // FFTGrid::interpolateSeismic (src/fftgrid.cpp:2564-2655) on a real DATA
  // grid: traces with energy at most energyTreshold times the mean are
  // rebuilt from their neighbours, the traces outside the box of the others
  // extrapolated.  In an access mode the host copy is written to the device
  // first (unless READ) and refreshed after.  Returns the traces flagged.
  int interpolateSeismic(float energyTreshold);

 private:
  int abiCubeType() const;
  void op(int code, const FFTGrid *other, float scalar);

  Context *ctx_;
  int nx_, ny_, nz_, nxp_, nyp_, nzp_, rnxp_, cnxp_;
  long long rsize_, csize_;
  int cubetype_ = CTMISSING;
  int precision_ = CRAVA_HIP_F32;
  bool istransformed_ = false;
  void *dev_ = nullptr;
  int mode_ = NONE;
  std::vector<float> host_;
  long long counterForGet_ = 0, counterForSet_ = 0;
};

class SeismicParametersHolder {
 public:
  explicit SeismicParametersHolder(Context *ctx) : ctx_(ctx) {}
  ~SeismicParametersHolder();

  void allocateGrids(int nx, int ny, int nz, int nxPad, int nyPad, int nzPad);   // :227-260
  void setPriorVar0(const double priorVar0[3][3]);
  const double (&getPriorVar0() const)[3][3] { return priorVar0_; }
  // :192-225: fill the six covariance grids from the lateral correlation
  // corrXY [nyp*nxp] and the temporal correlation priorCorrT, times Sigma0_ab.
  void initializeCorrelations(const std::vector<float> &priorCorrXY, const std::vector<float> &priorCorrT,
                              float corrGradI, float corrGradJ, int lowIntCut, int nzp);
  // :431-499 (host, crava_hip_circ_corr_t).
  std::vector<float> computeCircCorrT(const std::vector<float> &priorCorrT, int minIntFq, int nzp) const;
  // :748-768: the cyclic temporal correlation, the real covAlpha grid at
  // (0, 0, k) over its value at the origin, k < nzp.
  std::vector<float> extractParamCorrFromCovAlpha(int nzp) const;
  // :504-530: extractParamCorrFromCovAlpha, zero where min(k, nzp - k) >= nz.
  std::vector<float> getPriorCorrTFiltered(int nz, int nzp) const;
  // :771-806: the prior variances from the covariance grids' values at
  // the origin (time domain), and the posterior covariance's z profile at
  // lateral lag (0, 0) over the volume's nz samples.
  float getOrigin(FFTGrid *grid) const;
  void updatePriorVar();
  std::vector<float> createPostCov00(FFTGrid *postCov) const;
  void makeCircCorrTPosDef(std::vector<float> &circCorrT, int minIntFq, int nzp) const;

  void FFTAllGrids();       // :282-298, one launch per pass for each set
  void invFFTAllGrids();    // :262-278
  void FFTCovGrids();
  void invFFTCovGrids();
  // Arithmetic of the covariance grids' forward transforms (CRAVA_HIP_F32 or
  // CRAVA_HIP_MIXED): their spectra span many decades.
  void setCovTransformPrecision(int precision);

  FFTGrid *GetMuAlpha() const { return mu_[0]; }
  FFTGrid *GetMuBeta() const { return mu_[1]; }
  FFTGrid *GetMuRho() const { return mu_[2]; }
  FFTGrid *GetCovAlpha() const { return cov_[0]; }
  FFTGrid *GetCovBeta() const { return cov_[1]; }
  FFTGrid *GetCovRho() const { return cov_[2]; }
  FFTGrid *GetCrCovAlphaBeta() const { return cov_[3]; }
  FFTGrid *GetCrCovAlphaRho() const { return cov_[4]; }
  FFTGrid *GetCrCovBetaRho() const { return cov_[5]; }
  FFTGrid *mu(int c) const { return mu_[c]; }
  FFTGrid *cov(int q) const { return cov_[q]; }

 private:
  void transformSet(FFTGrid *const *grids, int n, bool inverse);
  Context *ctx_;
  FFTGrid *mu_[3] = {nullptr, nullptr, nullptr};
  FFTGrid *cov_[6] = {nullptr, nullptr, nullptr, nullptr, nullptr, nullptr};
  double priorVar0_[3][3] = {{0}};
};

// Static and dynamic parts of a time-lapse prior (State4D, src/state4d.h),
// transformed grids owned by the caller:
//   muStatic[3], muDynamic[3]
//   sigmaStaticStatic[6], sigmaDynamicDynamic[6]  order aa, ab, ar, bb, br, rr
//   sigmaStaticDynamic[9]                         [3a+b] = Cov(static a, dynamic b)
class State4D {
 public:
  State4D(Context *ctx, std::vector<FFTGrid *> muStatic, std::vector<FFTGrid *> muDynamic,
          std::vector<FFTGrid *> sigmaStaticStatic, std::vector<FFTGrid *> sigmaDynamicDynamic,
          std::vector<FFTGrid *> sigmaStaticDynamic);
  // src/state4d.cpp:90-159: the current state's means and covariances
  void merge(SeismicParametersHolder &current);
  // src/state4d.cpp:161-236: sigma in the order aa, ab, ar, bb, br, rr
  void mergeCov(const std::vector<FFTGrid *> &sigma);

 private:
  Context *ctx_;
  std::vector<FFTGrid *> ms_, md_, ss_, dd_, sd_;
};

// The posterior core of Crava on the constant-thickness path.
class Crava {
 public:
  struct Settings {
    int ntheta = 0;
    int nx = 0, ny = 0, nz = 0, nxp = 0, nyp = 0, nzp = 0;
    float dz = 4.0f;                    // time sample (ms): lz = nz * dz
    float lowCut = 5.0f, highCut = 55.0f;
    float wnc = 0.1f;                   // white-noise component
    std::vector<float> thetaDeg;        // [ntheta]
    std::vector<float> snr;             // empirical S/N per angle
    std::vector<float> A;               // reflection matrix [ntheta][3]
    std::vector<float> wavelets;        // [ntheta][nzp] scaled wavelets in FFT order
    std::vector<float> waveletNorm;     // [ntheta]
    std::vector<float> corrXY;          // lateral correlation [nyp*nxp] (error correlation)
    float corrGradI = 0.0f, corrGradJ = 0.0f;
    int precision = CRAVA_HIP_F32;      // posterior solve: F32 or MIXED
    double mixedThreshold = 0.1;
    // forward modelling: per angle, empty or [ny*nx] (i fastest) local
    // wavelet time shifts (ms), gains and stretches relThick*localStretch
    std::vector<std::vector<float>> localShift, localGain, stretch;
    // non-constant thickness: Simbox::getRelThick of every padded trace
    // [nyp*nxp] (empty: constant thickness), Simbox::getMinRelThick (0: the
    // smallest relThick) and the energy threshold of interpolateSeismic
    // (Crava::energyTreshold_)
    std::vector<double> relThick;
    double minRelThick = 0.0;
    // energy matching per angle (ModelAVODynamic::getMatchEnergies; empty: none)
    std::vector<unsigned char> matchEnergies;
    // local noise: per angle getLocalNoiseScale(angle) as [ny*nx] (i fastest)
    std::vector<std::vector<double>> localNoiseScale;
    float energyThreshold = 0.0f;
  };

  Crava(Context *ctx, const Settings &s, std::vector<FFTGrid *> seisData);
  Crava(const Crava &) = delete;             // owns device storage
  Crava &operator=(const Crava &) = delete;
  ~Crava();

  void computeDataVariance();                                               // :322-363
  void setupErrorCorrelation(const std::function<double(float)> &angularCorr);  // :366-401
  // :404-466: computeDataVariance and setupErrorCorrelation, then per angle
  // the model variance (the differenced wavelet under corrT, times
  // A Sigma0 A^T), the signal variance and the theoretical S/N; where
  // Settings::matchEnergies is set, the wavelet is scaled by the
  // energy-matching gain (Wavelet::scale, which leaves the norm as it was).
  // corrT: the holder's extractParamCorrFromCovAlpha(nzp).
  void computeVariances(const std::vector<float> &corrT, const SeismicParametersHolder &sph,
                        const std::function<double(float)> &angularCorr);
  // :887-1204: forward transforms, per-point posterior, inverse transforms
  // of the mean and the residuals (the covariance stays transformed).  With
  // Settings::relThick set (a simbox not of constant thickness) the data
  // are expected divided by the scale wavelet (divideDataByScaleWavelet)
  // and the per-k vectors are those of K = D A (:1058-1083).
  // Returns the number of failed factorisations.
  long long computePostMeanResidAndFFTCov(SeismicParametersHolder &sph);
  // :1416-1599 one realisation: noise (seed, stream) -> real grids real[3].
  long long simulate(SeismicParametersHolder &sph, unsigned long long seed, unsigned long long stream,
                     FFTGrid *real[3]);
  // The same realisation with the reference's RandomGen stream for the three
  // noise grids (seed0, seed1, seed2 in turn, :1466-1468): equal to the
  // reference's realisation for the same generator state.
  long long simulate(SeismicParametersHolder &sph, uint32_t &randomGenSeed, FFTGrid *real[3]);
  // :2340-2470 correctAlphaBetaRho: post[3] (time-domain posterior means)
  // corrected per trace towards mean2[3] (the prior means) with
  // Settings::localNoiseScale; keeps the realisation factors (sigmamdnew_)
  // that both simulate overloads then apply (:1547-1566).  For the
  // reference's flow post[c] is sph.mu(c), the grids the realisations add
  // (postAlpha_ etc., corrected in place, :1568).  G: the reference's
  // computeG [ntheta][3] (its LAPACK eigenvector signs), or nullptr for
  // crava_hip_local_noise_g's.
  void correctAlphaBetaRho(const double priorVar0[3][3], const double postVar0[3][3], FFTGrid *post[3],
                           FFTGrid *mean2[3], const double *G = nullptr);
  // :1667-1759 forward modelling: synthetic seismic of (alpha, beta, rho)
  // for every angle into syn[ntheta] (real grids).  Parameters held in the
  // Fourier domain are transformed back and forth, as the reference does.
  void computeSyntSeismic(FFTGrid *alpha, FFTGrid *beta, FFTGrid *rho, const std::vector<FFTGrid *> &syn);
  // :592-714 (with computeAdjustmentFactor, :726-817): divides every padded
  // trace of the real seismic grids by its local scale wavelet, then
  // interpolates low-energy traces (interpolateSeismic) per angle.  Needs the
  // error correlation (errThetaCov) and the prior covariance grids in the
  // time domain.  The local wavelet shifts and gains are localShift /
  // localGain; the stretch is relThick.
  void divideDataByScaleWavelet(const SeismicParametersHolder &sph);

  const std::vector<float> &getDataVariance() const { return dataVariance_; }
  const std::vector<float> &getErrorVariance() const { return errorVariance_; }
  const std::vector<float> &getModelVariance() const { return modelVariance_; }
  const std::vector<float> &getSignalVariance() const { return signalVariance_; }
  const std::vector<float> &getTheoSNRatio() const { return theoSNRatio_; }
  const std::vector<float> &getWavelets() const { return s_.wavelets; }
  const std::vector<float> &getWaveletNorm() const { return s_.waveletNorm; }
  const std::vector<double> &getErrThetaCov() const { return errThetaCov_; }
  long long getFP64Points() const { return nFP64_; }

 private:
  Context *ctx_;
  Settings s_;
  std::vector<FFTGrid *> seisData_;
  std::vector<float> dataVariance_, errorVariance_, modelVariance_, signalVariance_, theoSNRatio_;
  std::vector<double> errThetaCov_;   // [ntheta][ntheta]
  long long nFP64_ = 0;
  void *traceCov_ = nullptr;          // device [ny][nx][9] realisation factors
  long long realise(SeismicParametersHolder &sph, FFTGrid *real[3]);
};

}  // namespace crava_hip
