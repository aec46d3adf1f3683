// FFTGrid / Context of the C++ host mirror (see crava_host.hpp).
#include "crava_host.hpp"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <fstream>
#include <iomanip>
#include <sstream>
#include <stdexcept>

namespace crava_hip {

namespace {
constexpr float kRMissing = -99999.0f;  // RMISSING, src/definitions.h:35
}

void check(int rc, const char *what) {
  if (rc != CRAVA_HIP_OK) {
    const char *m = crava_hip_last_error();
    throw Error(rc, std::string(what) + ": " + (m ? m : ""));
  }
}

Context::Context(void *stream) { check(crava_hip_ctx_create(&ctx_, stream), "crava_hip_ctx_create"); }

Context::~Context() {
  if (ctx_) crava_hip_ctx_destroy(ctx_);
}

void Context::synchronize() { check(crava_hip_synchronize(ctx_), "crava_hip_synchronize"); }

FFTGrid::FFTGrid(Context *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp)
    : ctx_(ctx), nx_(nx), ny_(ny), nz_(nz), nxp_(nxp), nyp_(nyp), nzp_(nzp) {
  // src/fftgrid.cpp:37-67
  cnxp_ = nxp_ / 2 + 1;
  rnxp_ = 2 * cnxp_;
  csize_ = (long long)cnxp_ * nyp_ * nzp_;
  rsize_ = (long long)rnxp_ * nyp_ * nzp_;
}

FFTGrid::~FFTGrid() {
  if (dev_) crava_hip_grid_free(dev_);
}

void FFTGrid::createRealGrid(bool) {
  if (!dev_) check(crava_hip_grid_alloc(ctx_->handle(), nxp_, nyp_, nzp_, CRAVA_HIP_F32, &dev_), "createRealGrid");
  istransformed_ = false;
}

void FFTGrid::createComplexGrid() {
  if (!dev_) check(crava_hip_grid_alloc(ctx_->handle(), nxp_, nyp_, nzp_, CRAVA_HIP_F32, &dev_), "createComplexGrid");
  istransformed_ = true;
}

int FFTGrid::abiCubeType() const {
  switch (cubetype_) {
    case COVARIANCE: return CRAVA_HIP_COVARIANCE;
    case PARAMETER: return CRAVA_HIP_PARAMETER;
    default: return CRAVA_HIP_DATA;   // DATA, VELOCITY, CTMISSING: 1/sqrt(N)
  }
}

int FFTGrid::fftInPlace() {
  check(crava_hip_fft_r2c(ctx_->handle(), dev_, nxp_, nyp_, nzp_, abiCubeType(), precision_), "fftInPlace");
  istransformed_ = true;
  return 0;
}

int FFTGrid::invFFTInPlace() {
  check(crava_hip_fft_c2r(ctx_->handle(), dev_, nxp_, nyp_, nzp_, abiCubeType(), CRAVA_HIP_F32), "invFFTInPlace");
  istransformed_ = false;
  return 0;
}

void FFTGrid::op(int code, const FFTGrid *other, float scalar) {
  if (other && (other->nxp_ != nxp_ || other->nyp_ != nyp_ || other->nzp_ != nzp_))
    throw Error(CRAVA_HIP_E_INVALID, "grids of different size");
  check(crava_hip_grid_op(ctx_->handle(), code, dev_, other ? other->dev_ : nullptr, nxp_, nyp_, nzp_,
                          istransformed_ ? 1 : 0, scalar),
        "FFTGrid element-wise operator");
}

void FFTGrid::fillInConstant(float value) { op(CRAVA_HIP_OP_FILL_CONSTANT, nullptr, value); }
void FFTGrid::multiplyByScalar(float scalar) { op(CRAVA_HIP_OP_MULTIPLY_SCALAR, nullptr, scalar); }
void FFTGrid::addScalar(float scalar) { op(CRAVA_HIP_OP_ADD_SCALAR, nullptr, scalar); }
void FFTGrid::add(FFTGrid *g) { op(CRAVA_HIP_OP_ADD, g, 0.0f); }
void FFTGrid::subtract(FFTGrid *g) { op(CRAVA_HIP_OP_SUBTRACT, g, 0.0f); }
void FFTGrid::multiply(FFTGrid *g) { op(CRAVA_HIP_OP_MULTIPLY, g, 0.0f); }
void FFTGrid::conjugate() { op(CRAVA_HIP_OP_CONJUGATE, nullptr, 0.0f); }
void FFTGrid::changeSign() { op(CRAVA_HIP_OP_CHANGE_SIGN, nullptr, 0.0f); }

// Upload small host tables to a temporary device grid-sized buffer: the fill
// kernels read corrXY [nyp*nxp] and circCorrT [nzp] from device memory.
namespace {
struct DeviceTable {
  crava_hip_ctx *ctx;
  void *p = nullptr;
  DeviceTable(crava_hip_ctx *c, const float *host, size_t n) : ctx(c) {
    // one grid of (2*n) x 1 x 1 floats holds n values
    check(crava_hip_grid_alloc(ctx, (int)(2 * n), 1, 1, CRAVA_HIP_F32, &p), "table alloc");
    check(crava_hip_grid_upload(ctx, p, host, n * sizeof(float)), "table upload");
  }
  ~DeviceTable() {
    if (p) crava_hip_grid_free(p);
  }
};
}  // namespace

void FFTGrid::fillInParamCorr(const std::vector<float> &corrXY, const float *circCorrT, float gradI, float gradJ) {
  if ((long long)corrXY.size() < (long long)nxp_ * nyp_) throw Error(CRAVA_HIP_E_INVALID, "corrXY too small");
  DeviceTable cxy(ctx_->handle(), corrXY.data(), (size_t)nxp_ * nyp_);
  DeviceTable ct(ctx_->handle(), circCorrT, (size_t)nzp_);
  const float one = 1.0f;
  void *g = dev_;
  check(crava_hip_fill_param_corr(ctx_->handle(), &g, 1, &one, nxp_, nyp_, nzp_, static_cast<const float *>(cxy.p),
                                  static_cast<const float *>(ct.p), gradI, gradJ),
        "fillInParamCorr");
  ctx_->synchronize();
  istransformed_ = false;
}

void FFTGrid::fillInErrCorr(const std::vector<float> &corrXY, float gradI, float gradJ) {
  if ((long long)corrXY.size() < (long long)nxp_ * nyp_) throw Error(CRAVA_HIP_E_INVALID, "corrXY too small");
  DeviceTable cxy(ctx_->handle(), corrXY.data(), (size_t)nxp_ * nyp_);
  check(crava_hip_fill_err_corr(ctx_->handle(), dev_, nxp_, nyp_, nzp_, static_cast<const float *>(cxy.p), gradI,
                                gradJ),
        "fillInErrCorr");
  ctx_->synchronize();
  istransformed_ = false;
}

void FFTGrid::setAccessMode(int mode) {
  if (mode < NONE || mode > RANDOMACCESS) throw Error(CRAVA_HIP_E_INVALID, "invalid access mode");
  mode_ = mode;
  counterForGet_ = counterForSet_ = 0;
  if (mode == NONE) return;
  host_.resize((size_t)rsize_);
  if (mode != WRITE) check(crava_hip_grid_download(ctx_->handle(), host_.data(), dev_, sizeof(float) * rsize_),
                           "setAccessMode");
}

int FFTGrid::interpolateSeismic(float energyTreshold) {
  if (cubetype_ != DATA || istransformed_) throw Error(CRAVA_HIP_E_STATE, "interpolateSeismic needs a real DATA grid");
  const bool hostCopy = mode_ != NONE;
  if (hostCopy && mode_ != READ)
    check(crava_hip_grid_upload(ctx_->handle(), dev_, host_.data(), sizeof(float) * rsize_), "interpolateSeismic");
  int flagged = 0, zero = 0;
  check(crava_hip_interpolate_seismic(ctx_->handle(), nx_, ny_, nz_, nxp_, nyp_, nzp_, energyTreshold, dev_, &flagged,
                                      &zero),
        "interpolateSeismic");
  if (hostCopy)
    check(crava_hip_grid_download(ctx_->handle(), host_.data(), dev_, sizeof(float) * rsize_), "interpolateSeismic");
  return flagged;
}

void FFTGrid::endAccess() {
  if (mode_ == WRITE || mode_ == READANDWRITE || mode_ == RANDOMACCESS)
    check(crava_hip_grid_upload(ctx_->handle(), dev_, host_.data(), sizeof(float) * rsize_), "endAccess");
  mode_ = NONE;
  counterForGet_ = counterForSet_ = 0;   // src/fftgrid.h:173
}

// Cursor access (src/fftgrid.cpp:1419-1453,1604-1653): memory order, i fastest,
// wrapping to the start after the last value.
float FFTGrid::getNextReal() {
  if (mode_ == NONE || istransformed_) throw Error(CRAVA_HIP_E_STATE, "getNextReal needs a real grid in access mode");
  const float r = host_[(size_t)counterForGet_];
  if (++counterForGet_ == rsize_) counterForGet_ = 0;
  return r;
}

int FFTGrid::setNextReal(float value) {
  if (mode_ == NONE || istransformed_) throw Error(CRAVA_HIP_E_STATE, "setNextReal needs a real grid in access mode");
  host_[(size_t)counterForSet_] = value;
  if (++counterForSet_ == rsize_) counterForSet_ = 0;
  return 0;
}

fftw_complex_f FFTGrid::getNextComplex() {
  if (mode_ == NONE || !istransformed_) throw Error(CRAVA_HIP_E_STATE, "getNextComplex needs a transformed grid");
  fftw_complex_f v{host_[2 * (size_t)counterForGet_], host_[2 * (size_t)counterForGet_ + 1]};
  if (++counterForGet_ == csize_) counterForGet_ = 0;
  return v;
}

int FFTGrid::setNextComplex(fftw_complex_f value) {
  if (mode_ == NONE || !istransformed_) throw Error(CRAVA_HIP_E_STATE, "setNextComplex needs a transformed grid");
  host_[2 * (size_t)counterForSet_] = value.re;
  host_[2 * (size_t)counterForSet_ + 1] = value.im;
  if (++counterForSet_ == csize_) counterForSet_ = 0;
  return 0;
}

// src/fftgrid.cpp getRealValue: RMISSING outside the (extended) simbox.
float FFTGrid::getRealValue(int i, int j, int k, bool extSimbox) const {
  if (mode_ == NONE || mode_ == WRITE) throw Error(CRAVA_HIP_E_STATE, "getRealValue needs READ or RANDOMACCESS");
  const bool in = extSimbox ? (i < nxp_ && j < nyp_ && k < nzp_) : (i < nx_ && j < ny_ && k < nz_);
  if (!in || i < 0 || j < 0 || k < 0) return kRMissing;
  return host_[(size_t)i + (size_t)rnxp_ * ((size_t)j + (size_t)nyp_ * k)];
}

int FFTGrid::setRealValue(int i, int j, int k, float value, bool extSimbox) {
  if (mode_ == NONE || mode_ == READ) throw Error(CRAVA_HIP_E_STATE, "setRealValue needs WRITE or RANDOMACCESS");
  const bool in = extSimbox ? (i < nxp_ && j < nyp_ && k < nzp_) : (i < nx_ && j < ny_ && k < nz_);
  if (!in || i < 0 || j < 0 || k < 0) return 1;
  host_[(size_t)i + (size_t)rnxp_ * ((size_t)j + (size_t)nyp_ * k)] = value;
  return 0;
}

fftw_complex_f FFTGrid::getComplexValue(int i, int j, int k, bool extSimbox) const {
  if (mode_ == NONE || mode_ == WRITE) throw Error(CRAVA_HIP_E_STATE, "getComplexValue needs READ or RANDOMACCESS");
  const bool in = extSimbox ? (i < cnxp_ && j < nyp_ && k < nzp_) : (i < nx_ && j < ny_ && k < nz_);
  if (!in || i < 0 || j < 0 || k < 0) return fftw_complex_f{kRMissing, 0.0f};
  const size_t o = (size_t)i + (size_t)cnxp_ * ((size_t)j + (size_t)nyp_ * k);
  return fftw_complex_f{host_[2 * o], host_[2 * o + 1]};
}

// ---------------------------------------------------------------------------
// CRAVA binary grid files (src/fftgrid.cpp:2399-2487, NRLib big-endian
// binary numbers)
namespace {

template <typename T, typename U>
void put_be(std::ostream &os, T v) {
  U u;
  std::memcpy(&u, &v, sizeof(T));
  unsigned char b[sizeof(T)];
  for (size_t i = 0; i < sizeof(T); ++i) b[i] = (unsigned char)(u >> (8 * (sizeof(T) - 1 - i)));
  os.write(reinterpret_cast<const char *>(b), sizeof(T));
}

template <typename T, typename U>
T get_be(std::istream &is) {
  unsigned char b[sizeof(T)];
  if (!is.read(reinterpret_cast<char *>(b), sizeof(T))) throw std::runtime_error("unexpected end of file");
  U u = 0;
  for (size_t i = 0; i < sizeof(T); ++i) u = (U)((u << 8) | b[i]);
  T v;
  std::memcpy(&v, &u, sizeof(T));
  return v;
}

}  // namespace

void FFTGrid::writeCravaFile(const std::string &fileName, const CravaHeader &h) const {
  const std::string fname = fileName + ".crava";
  std::vector<float> v((size_t)rsize_);
  check(crava_hip_grid_download(ctx_->handle(), v.data(), dev_, sizeof(float) * rsize_), "writeCravaFile");
  std::ofstream os(fname, std::ios::out | std::ios::binary);
  if (!os) throw Error(CRAVA_HIP_E_INVALID, "cannot open " + fname + " for writing");
  os << "crava_fftgrid_binary" << "\n";
  for (double d : {h.x0, h.y0, h.dx, h.dy}) put_be<double, uint64_t>(os, d);
  put_be<int32_t, uint32_t>(os, h.nx);
  put_be<int32_t, uint32_t>(os, h.ny);
  for (double d : {h.il0, h.xl0, h.ilStepX, h.ilStepY, h.xlStepX, h.xlStepY, h.angle}) put_be<double, uint64_t>(os, d);
  put_be<int32_t, uint32_t>(os, rnxp_);
  put_be<int32_t, uint32_t>(os, nyp_);
  put_be<int32_t, uint32_t>(os, nzp_);
  std::vector<unsigned char> buf(v.size() * 4);
  for (size_t i = 0; i < v.size(); ++i) {
    uint32_t u;
    std::memcpy(&u, &v[i], 4);
    buf[4 * i] = (unsigned char)(u >> 24);
    buf[4 * i + 1] = (unsigned char)(u >> 16);
    buf[4 * i + 2] = (unsigned char)(u >> 8);
    buf[4 * i + 3] = (unsigned char)u;
  }
  os.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)buf.size());
  if (!os) throw Error(CRAVA_HIP_E_INVALID, "write failed on " + fname);
}

void FFTGrid::readCravaFile(const std::string &fileName, std::string &errText, bool) {
  try {
    std::ifstream is(fileName, std::ios::in | std::ios::binary);
    if (!is) throw std::runtime_error("cannot open '" + fileName + "'");
    std::string type;
    std::getline(is, type);
    for (int i = 0; i < 4; ++i) (void)get_be<double, uint64_t>(is);
    for (int i = 0; i < 2; ++i) (void)get_be<int32_t, uint32_t>(is);
    for (int i = 0; i < 7; ++i) (void)get_be<double, uint64_t>(is);
    const int rnxp = get_be<int32_t, uint32_t>(is);
    const int nyp = get_be<int32_t, uint32_t>(is);
    const int nzp = get_be<int32_t, uint32_t>(is);
    if (rnxp != rnxp_ || nyp != nyp_ || nzp != nzp_)
      throw std::runtime_error("Grid dimension is wrong for file '" + fileName + "'.");
    std::vector<unsigned char> buf((size_t)rsize_ * 4);
    if (!is.read(reinterpret_cast<char *>(buf.data()), (std::streamsize)buf.size()))
      throw std::runtime_error("unexpected end of file '" + fileName + "'");
    std::vector<float> v((size_t)rsize_);
    for (size_t i = 0; i < v.size(); ++i) {
      const uint32_t u = ((uint32_t)buf[4 * i] << 24) | ((uint32_t)buf[4 * i + 1] << 16) |
                         ((uint32_t)buf[4 * i + 2] << 8) | (uint32_t)buf[4 * i + 3];
      std::memcpy(&v[i], &u, 4);
    }
    createRealGrid();
    check(crava_hip_grid_upload(ctx_->handle(), dev_, v.data(), sizeof(float) * rsize_), "readCravaFile");
    istransformed_ = false;
  } catch (const std::exception &e) {
    errText += std::string("Error: ") + e.what() + "\n";
  }
}

// ---------------------------------------------------------------------------
// STORM grid files (src/fftgrid.cpp:2118-2181; header src/simbox.cpp:408-427;
// numbers formatted as NRLib::ToString(x, 6): fixed, 6 decimals)
namespace {
std::string fixed6(double v) {
  std::ostringstream o;
  o << std::fixed << std::setprecision(6) << v;
  return o.str();
}
}  // namespace

void FFTGrid::writeStormFile(const std::string &fileName, const StormHeader &h, bool ascii, bool padding, bool flat,
                             bool scientific) const {
  const int nx = padding ? nxp_ : nx_, ny = padding ? nyp_ : ny_, nz = padding ? nzp_ : nz_;
  if (!flat && h.topName.empty()) throw Error(CRAVA_HIP_E_INVALID, "writeStormFile needs surface names unless flat");
  std::string header = ascii ? "storm_petro_ascii\n" : "storm_petro_binary\n";
  header += "0 " + std::to_string(cubetype_) + " " + fixed6(-99999.0) + "\n";   // RMISSING
  header += "FFTGrid\n";
  if (!flat)
    header += fixed6(h.xmin) + " " + fixed6(h.lx) + " " + fixed6(h.ymin) + " " + fixed6(h.ly) + " " + h.topName + " " +
              h.botName + " 0.0 0.0\n";
  else
    header += fixed6(h.xmin) + " " + fixed6(h.lx) + " " + fixed6(h.ymin) + " " + fixed6(h.ly) + " 0.0 " + fixed6(h.lz) +
              " 0.0 0.0\n";
  header += fixed6(h.lz) + " " + fixed6(h.angle * 180 / 3.14159265358979323846) + "\n\n";
  header += std::to_string(nx) + " " + std::to_string(ny) + " " + std::to_string(nz) + "\n";

  std::vector<float> v((size_t)rsize_);
  check(crava_hip_grid_download(ctx_->handle(), v.data(), dev_, sizeof(float) * rsize_), "writeStormFile");
  const auto at = [&](int i, int j, int k) { return v[(size_t)i + (size_t)rnxp_ * ((size_t)j + (size_t)nyp_ * k)]; };
  const std::string fname = fileName + (ascii ? ".txt" : ".storm");
  std::ofstream os(fname, ascii ? std::ios::out : (std::ios::out | std::ios::binary));
  if (!os) throw Error(CRAVA_HIP_E_INVALID, "cannot open " + fname + " for writing");
  os << header;
  if (!ascii) {
    std::vector<unsigned char> buf((size_t)nx * ny * nz * 4);
    size_t o = 0;
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j)
        for (int i = 0; i < nx; ++i) {
          const float f = at(i, j, k);
          uint32_t u;
          std::memcpy(&u, &f, 4);
          buf[o++] = (unsigned char)(u >> 24);
          buf[o++] = (unsigned char)(u >> 16);
          buf[o++] = (unsigned char)(u >> 8);
          buf[o++] = (unsigned char)u;
        }
    os.write(reinterpret_cast<const char *>(buf.data()), (std::streamsize)buf.size());
  } else {
    if (scientific) os << std::scientific;
    else os << std::fixed << std::setprecision(6);
    for (int k = 0; k < nz; ++k)
      for (int j = 0; j < ny; ++j) {
        for (int i = 0; i < nx - 1; ++i) os << at(i, j, k) << " ";
        os << at(nx - 1, j, k) << "\n";
      }
  }
  os << "0\n";
  if (!os) throw Error(CRAVA_HIP_E_INVALID, "write failed on " + fname);
}

}  // namespace crava_hip
