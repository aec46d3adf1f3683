// FFTGrid / Context of the C++ host mirror (see crava_host.hpp).
#include "crava_host.hpp"

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <map>
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

void FFTGrid::fillInComplexNoise(uint32_t &randomGenSeed) {
  std::vector<float> v((size_t)csize_ * 2);
  check(crava_hip_complex_noise_reference(nxp_, nyp_, nzp_, &randomGenSeed, v.data()), "fillInComplexNoise");
  check(crava_hip_grid_upload(ctx_->handle(), dev_, v.data(), sizeof(float) * v.size()), "fillInComplexNoise");
  istransformed_ = true;
  cubetype_ = PARAMETER;
}

std::vector<float> FFTGrid::readTrace(int i, int j, int k0, int nk) const {
  if (istransformed_) throw Error(CRAVA_HIP_E_STATE, "readTrace needs a real grid");
  std::vector<float> v((size_t)(nk > 0 ? nk : 0));
  check(crava_hip_grid_read_trace(ctx_->handle(), dev_, nxp_, nyp_, nzp_, i, j, k0, nk, v.data()), "readTrace");
  return v;
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

// ---------------------------------------------------------------------------
// SEG-Y (src/fftgrid.cpp:2189-2264 with NRLib::SegY, TextualHeader,
// BinaryHeader, TraceHeader, SegyGeometry) and SGRI (src/fftgrid.cpp:2325-2397)
namespace {

constexpr double kRMissingD = -99999.000;    // src/definitions.h:35
constexpr double kWellMissingD = -999.00000;  // src/definitions.h:39

// EBCDIC of the characters of the textual header (code page 037)
unsigned char ebcdic(char c) {
  if (c >= '0' && c <= '9') return (unsigned char)(0xF0 + (c - '0'));
  if (c >= 'A' && c <= 'I') return (unsigned char)(0xC1 + (c - 'A'));
  if (c >= 'J' && c <= 'R') return (unsigned char)(0xD1 + (c - 'J'));
  if (c >= 'S' && c <= 'Z') return (unsigned char)(0xE2 + (c - 'S'));
  return 0x40;   // space
}

// The reference's standard SEG-Y textual header (TextualHeader::standardHeader):
// 40 card images of 80 columns, "C n" in the first columns.
std::string textual_header() {
  static const char *const body[40] = {
      "", "", "", "", "", "", "", "", "", "", "", "         THIS SEGY FILE WAS WRITTEN BY", "",
      "         THE CRAVA PROGRAM", "", "         DEVELOPED BY", "", "         RAGNAR HAUGE", "",
      "         ANNE RANDI SYVERSVEEN           AND", "", "         ODD KOLBJORNSEN", "", "         AT", "",
      "         THE NORWEGIAN COMPUTING CENTER", "", "         2003", "", "", "", "", "", "", "", "", "", "", "",
      " END EBCDIC"};
  std::string out;
  for (int l = 0; l < 40; ++l) {
    char card[81];
    std::snprintf(card, sizeof(card), "C%2d%-77s", l + 1, body[l]);
    for (int c = 0; c < 80; ++c) out.push_back((char)ebcdic(card[c]));
  }
  return out;
}

void be16(std::string &b, short v) {
  b.push_back((char)((unsigned short)v >> 8));
  b.push_back((char)((unsigned short)v & 0xFF));
}
void be32(std::string &b, int v) {
  for (int s = 24; s >= 0; s -= 8) b.push_back((char)(((unsigned int)v >> s) & 0xFF));
}

// IEEE single -> IBM single: value = (M / 2^24) 2^(e - 126), M the 24-bit
// significand with its leading bit, written as F 16^(E - 64) / 2^24 with
// E = ceil((e - 126) / 4) + 64 and F = M >> (4 (E - 64) - (e - 126)),
// truncated; zero (either sign) -> 0.  The exponent field is used as is for
// every bit pattern (no special cases), as NRLib's writer does.
uint32_t ibm_float(float f) {
  uint32_t u;
  std::memcpy(&u, &f, 4);
  if ((u & 0x7fffffffu) == 0) return 0;
  const int e = (int)((u >> 23) & 0xff);
  const uint32_t m = (u & 0x7fffffu) | 0x800000u;
  const int t = e - 126;
  const int q = (t >= 0) ? (t + 3) / 4 : -((-t) / 4);   // ceil(t / 4)
  const int shift = 4 * q - t;
  return (u & 0x80000000u) | ((uint32_t)(q + 64) << 24) | (m >> shift);
}

// TraceHeaderFormat locations (1-based): scalco, utm x, utm y, inline, crossline
struct Thf {
  int scalco, utmx, utmy, il, xl;
};
Thf trace_header_format(int code) {
  switch (code) {
    case 0: return {71, 73, 77, 9, 21};      // SeisWorks
    case 1: return {71, 73, 77, 221, 21};    // IESX
    case 2: return {71, 181, 185, 189, 193}; // SIP
    case 3: return {71, 73, 77, 5, 21};      // Charisma
    case 4: return {71, 73, 77, 181, 185};   // SIPX
    default: throw Error(CRAVA_HIP_E_INVALID, "unknown trace header format");
  }
}

// the 240 bytes of a trace header (TraceHeader::Write)
std::string trace_header(const Thf &f, float utmx, float utmy, int il, int xl, short ns, short dt) {
  std::string b;
  int i = 0;
  while (i < 240) {
    if (i == f.scalco - 1) { be16(b, 0); i += 2; }
    else if (i == f.utmx - 1) { be32(b, (int)utmx); i += 4; }
    else if (i == f.utmy - 1) { be32(b, (int)utmy); i += 4; }
    else if (i == 114) { be16(b, ns); i += 2; }
    else if (i == 116) { be16(b, dt); i += 2; }
    else if (i == f.il - 1) { be32(b, il); i += 4; }
    else if (i == f.xl - 1) { be32(b, xl); i += 4; }
    else { b.push_back(0); b.push_back(0); i += 2; }
  }
  return b;
}

struct Surfaces {
  const FFTGrid::SimboxGeometry &g;
  int nx;
  void xy(int i, int j, double &x, double &y) const {   // Simbox::getXYCoord
    const double rx = ((double)i + 0.5) * g.dx, ry = ((double)j + 0.5) * g.dy;
    x = rx * std::cos(g.angle) - ry * std::sin(g.angle) + g.xmin;
    y = rx * std::sin(g.angle) + ry * std::cos(g.angle) + g.ymin;
  }
  double top(int i, int j) const { return g.top[(size_t)j * nx + i]; }
  double bot(int i, int j) const { return g.bot[(size_t)j * nx + i]; }
  double rel_thick(int i, int j) const {                 // Simbox::getRelThick
    const double t = top(i, j), b = bot(i, j);
    return (t != kRMissingD && b != kRMissingD) ? (b - t) / g.lz : 1.0;
  }
};

}  // namespace

void FFTGrid::writeSegyFile(const std::string &fileName, const SimboxGeometry &g, float z0,
                            int traceHeaderFormat) const {
  if (istransformed_) throw Error(CRAVA_HIP_E_STATE, "writeSegyFile needs a real grid");
  if ((long long)g.top.size() != (long long)nx_ * ny_ || (long long)g.bot.size() != (long long)nx_ * ny_)
    throw Error(CRAVA_HIP_E_INVALID, "writeSegyFile needs top and base surfaces for every trace");
  const Thf thf = trace_header_format(traceHeaderFormat);
  std::vector<float> v((size_t)rsize_);
  check(crava_hip_grid_download(ctx_->handle(), v.data(), dev_, sizeof(float) * rsize_), "writeSegyFile");
  const auto value = [&](int i, int j, int k) -> float {   // getRealValue inside the volume
    if (i < 0 || j < 0 || k < 0 || i >= nx_ || j >= ny_ || k >= nz_) return (float)kRMissingD;
    return v[(size_t)i + (size_t)rnxp_ * ((size_t)j + (size_t)nyp_ * k)];
  };
  const Surfaces sb{g, nx_};
  float dz = (float)std::floor(g.dz + 0.5);
  if (dz == 0) dz = 1.0f;
  const int segynz = (int)std::ceil((g.zmax - z0) / dz);
  if (segynz < 1) throw Error(CRAVA_HIP_E_INVALID, "writeSegyFile: z0 must lie above zmax");
  // traces by their index in the SEG-Y geometry (SegyGeometry::FindIndex);
  // a later trace at the same index replaces the earlier one
  struct Trace {
    float x, y;
    int il, xl;
    std::vector<float> s;
  };
  std::map<size_t, Trace> traces;
  const double cr = std::cos(g.angle), sr = std::sin(g.angle);
  std::vector<float> tr(segynz > 0 ? segynz : 0);
  for (int j = 0; j < ny_; ++j)
    for (int i = 0; i < nx_; ++i) {
      double x, y;
      sb.xy(i, j, x, y);
      const double z = sb.top(i, j);
      if (z == kRMissingD || z == kWellMissingD) continue;
      const double gdz = g.dz * sb.rel_thick(i, j);
      // a top below zmax would start past the trace (the reference writes out
      // of bounds there): clamp
      const int firstData = std::min((int)std::floor(0.5 + (z - z0) / dz), segynz);
      int endData = (int)std::floor(0.5 + ((z - z0) + nz_ * gdz) / dz);
      if (endData > segynz) endData = segynz;
      int k = 0;
      for (; k < firstData; ++k) tr[k] = 0;
      for (; k < endData; ++k) {
        // getRegularZInterpolatedRealValue (src/fftgrid.cpp:2491-2522)
        const float zz = (float)((double)z0 + (double)dz * k);
        float t = (float)((zz - z) / gdz);
        int kg = int(t);
        t -= kg;
        float val;
        if (kg < -1 || kg > nz_ - 1) {
          val = (float)kRMissingD;
        } else {
          if (kg < 0) { kg = 0; t = 0; }
          else if (kg > nz_ - 2) { kg = nz_ - 2; t = 1; }
          const float v1 = value(i, j, kg), v2 = value(i, j, kg + 1);
          val = (1 - t) * v1 + t * v2;
        }
        tr[k] = val;
      }
      for (; k < segynz; ++k) tr[k] = 0;
      const float xf = (float)x, yf = (float)y;
      // SegyGeometry::IsInside / FindILXL / FindIndex
      const float xi = (float)(((xf - g.xmin) * cr + (yf - g.ymin) * sr) / g.dx);
      const float yi = (float)((-(xf - g.xmin) * sr + (yf - g.ymin) * cr) / g.dy);
      if (!(xi >= 0.0f && xi < nx_ && yi >= 0.0f && yi < ny_))
        throw Error(CRAVA_HIP_E_INVALID, "writeSegyFile: coordinates are outside grid");
      const int il = (int)(0.5 + g.il0 + (xf - g.xmin) * g.ilStepX + (yf - g.ymin) * g.ilStepY);
      const int xl = (int)(0.5 + g.xl0 + (xf - g.xmin) * g.xlStepX + (yf - g.ymin) * g.xlStepY);
      const int ii = (int)(((xf - g.xmin) * cr + (yf - g.ymin) * sr) / g.dx);
      const int jj = (int)((-(xf - g.xmin) * sr + (yf - g.ymin) * cr) / g.dy);
      if (ii < 0 || jj < 0 || ii >= nx_ || jj >= ny_) throw Error(CRAVA_HIP_E_INVALID, "writeSegyFile: invalid index");
      traces[(size_t)ii + (size_t)nx_ * jj] = Trace{xf, yf, il, xl, tr};
    }
  // SegY::WriteAllTracesToFile: by inline, crossline, then index
  std::vector<std::pair<size_t, const Trace *>> order;
  for (const auto &kv : traces) order.push_back({kv.first, &kv.second});
  std::sort(order.begin(), order.end(), [](const std::pair<size_t, const Trace *> &a,
                                           const std::pair<size_t, const Trace *> &b) {
    if (a.second->il != b.second->il) return a.second->il < b.second->il;
    if (a.second->xl != b.second->xl) return a.second->xl < b.second->xl;
    return a.first < b.first;
  });
  std::string out = textual_header();
  // BinaryHeader::Write (written before the geometry is set: 0 traces per line)
  be32(out, 9999);
  be32(out, 1);
  be32(out, 1);
  be16(out, 0);
  be16(out, 0);
  be16(out, (short)(1000 * (double)dz));
  be16(out, 0);
  be16(out, (short)segynz);
  be16(out, 0);
  be16(out, 1);
  be16(out, 1);
  be16(out, 4);
  for (int q = 0; q < 12; ++q) be16(out, 0);
  be16(out, 1);
  out.append(344, '\0');
  const short dt = (short)(unsigned short)(dz * 1000);
  for (const auto &o : order) {
    out += trace_header(thf, o.second->x, o.second->y, o.second->il, o.second->xl, (short)segynz, dt);
    for (float f : o.second->s) be32(out, (int)ibm_float(f));
  }
  const std::string fname = fileName + ".segy";
  std::ofstream os(fname, std::ios::out | std::ios::binary);
  if (!os) throw Error(CRAVA_HIP_E_INVALID, "cannot open " + fname + " for writing");
  os.write(out.data(), (std::streamsize)out.size());
  if (!os) throw Error(CRAVA_HIP_E_INVALID, "write failed on " + fname);
}

void FFTGrid::writeSgriFile(const std::string &fileName, const SimboxGeometry &g, const std::string &label) const {
  if (istransformed_) throw Error(CRAVA_HIP_E_STATE, "writeSgriFile needs a real grid");
  if ((long long)g.top.size() != (long long)nx_ * ny_ || (long long)g.bot.size() != (long long)nx_ * ny_)
    throw Error(CRAVA_HIP_E_INVALID, "writeSgriFile needs top and base surfaces for every trace");
  std::vector<float> v((size_t)rsize_);
  check(crava_hip_grid_download(ctx_->handle(), v.data(), dev_, sizeof(float) * rsize_), "writeSgriFile");
  const Surfaces sb{g, nx_};
  const std::string hname = fileName + ".Sgrh", bname = fileName + ".Sgri";
  std::ofstream h(hname);
  if (!h) throw Error(CRAVA_HIP_E_INVALID, "cannot open " + hname + " for writing");
  h << "NORSAR General Grid Format v1.0\n3\nX (km)\nY (km)\nT (s)\nFFT-grid\n1\n" << label << "\n1 1 1\n";
  const float dz = (float)g.dz;
  const int nz = (int)std::ceil((g.zmax - g.zmin) / dz);
  h << nx_ << " " << ny_ << " " << nz << "\n";
  h << std::setprecision(10);
  h << g.dx * 0.001 << " " << g.dy * 0.001 << " " << dz * 0.001 << "\n";
  h << (g.xmin + 0.5 * g.dx) * 0.001 << " " << (g.ymin + 0.5 * g.dy) * 0.001 << " " << (g.zmin + 0.5 * dz) * 0.001
    << "\n";
  h << g.angle << " 0\n" << kRMissingD << "\n" << bname << "\n0\n";
  if (!h) throw Error(CRAVA_HIP_E_INVALID, "write failed on " + hname);
  std::string out;
  out.reserve((size_t)4 * nx_ * ny_ * (nz > 0 ? nz : 0));
  for (int k = 0; k < nz; ++k)
    for (int j = 0; j < ny_; ++j)
      for (int i = 0; i < nx_; ++i) {
        const double zt = sb.top(i, j), zb = sb.bot(i, j);
        const double z = g.zmin + k * dz;
        float val = (float)kRMissingD;
        if (!(z < zt || z > zb)) {
          const int sk = (int)((z - zt) / (g.dz * sb.rel_thick(i, j)) + 0.5);
          if (sk >= 0 && sk < nz_) val = v[(size_t)i + (size_t)rnxp_ * ((size_t)j + (size_t)nyp_ * sk)];
        }
        uint32_t u;
        std::memcpy(&u, &val, 4);
        be32(out, (int)u);
      }
  std::ofstream b(bname, std::ios::out | std::ios::binary);
  if (!b) throw Error(CRAVA_HIP_E_INVALID, "cannot open " + bname + " for writing");
  b.write(out.data(), (std::streamsize)out.size());
  if (!b) throw Error(CRAVA_HIP_E_INVALID, "write failed on " + bname);
}

}  // namespace crava_hip
