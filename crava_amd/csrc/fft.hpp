// 3-D real <-> half-complex transforms on the FFTGrid layout, as hand-written
// Stockham mixed-radix (2/3/4/5/7/8) kernels for gfx950.
//
// Layout (src/fftgrid.cpp:51-66): real [nzp][nyp][rnxp] with
// rnxp = 2*(nxp/2+1); complex [nzp][nyp][cnxp] aliasing the same memory.
// A forward transform is three passes over HBM: x rows (r2c, two real rows
// packed into one complex pencil), y pencils, z pencils; the inverse runs
// z, y, then the c2r rows.  Each pass stages whole pencils in LDS.
#pragma once
#include "common.hpp"

namespace crava {

struct PassPlan {
  int n = 1;          // transform length
  int nrad = 0;       // number of radix passes
  int rad[24] = {0};  // radices in execution order
  int rmin = 1;       // smallest radix (sets threads per pencil = n / rmin)
  int tx = 1;         // pencils per workgroup
  int ldsp = 1;       // LDS pitch of one pencil, in complex elements
  int threads = 64;   // workgroup size
  size_t lds_bytes = 0;
};

// Radix sequence for n (7-smooth), preferring radix 8 and keeping the
// smallest radix >= 4 for powers of two >= 4.  Returns false if n has a
// prime factor > 7.
bool factor_radices(int n, PassPlan &p);

// Convert ncomplex values between float and double storage (FftIo code).
int launch_convert(int io, const void *src, void *dst, long long ncomplex, hipStream_t s);

class Fft3D {
 public:
  // prec: 0 float, 1 float storage with double arithmetic, 2 double.
  Fft3D(int nxp, int nyp, int nzp, int prec);
  ~Fft3D();
  int init();  // CRAVA_HIP_OK or an error code
  // In-place forward r2c; the result is multiplied by `scale`.
  int r2c(void *grid, hipStream_t s, double scale);
  // In-place inverse c2r; the result is multiplied by `scale`.
  int c2r(void *grid, hipStream_t s, double scale);
  // Individual passes (exposed for profiling and the batch executor).
  int pass_x(void *grid, hipStream_t s, bool inverse, double scale);
  int pass_y(void *grid, hipStream_t s, bool inverse);
  int pass_z(void *grid, hipStream_t s, bool inverse);

  int nxp, nyp, nzp;
  int prec;
  bool dbl;  // double arithmetic (LDS, twiddles)
  PassPlan px, py, pz;
  void *twx = nullptr, *twy = nullptr, *twz = nullptr;  // device twiddle tables
  void *scratch = nullptr;  // double intermediates of a float-storage/double-arithmetic plan
  // Plans built with a caller-chosen pencils-per-workgroup (generic kernels).
  int set_tx(int tx_x, int tx_y, int tx_z);
  // Specialised-kernel launch variant per axis (x, y, z); -1 = generic kernel.
  int var[3] = {-1, -1, -1};
  int set_variant(int axis, int v);
  int num_variants(int axis) const;
  // Time all variants on `grid` (overwritten) and keep the fastest per axis.
  int autotune(void *grid, hipStream_t s, int reps, float *ms_out);

 private:
  int build_plan(int n, PassPlan &p, int want_tx);
};

}  // namespace crava
