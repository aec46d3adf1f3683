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
#include "fft_static.hpp"

#include <cmath>

namespace crava {

// One radix pass of a runtime-length plan, with its index constants
// precomputed on the host: ns_mul / nb_mul give j / Ns and b / NB as
// __umulhi(x, mul) (exact for the index ranges of a workgroup).
struct PassStep {
  int R = 2;           // radix
  int Ns = 1;          // product of the radices before this pass
  int NB = 1;          // butterflies per pencil, n / R
  int tstep = 1;       // twiddle stride n / (Ns R)
  unsigned ns_mul = 0, nb_mul = 0;
};

struct PassPlan {
  int n = 1;          // transform length
  int nrad = 0;       // number of radix passes
  int rad[24] = {0};  // radices in execution order
  int rmin = 1;       // smallest radix (sets the butterflies of the widest pass)
  int tx = 1;         // pencils per workgroup
  int ldsp = 1;       // LDS pitch of one pencil (padded), in complex elements
  int threads = 64;   // workgroup size
  unsigned tx_mul = 0, n_mul = 0, cn_mul = 0;  // division by tx, n, n/2+1
  size_t lds_bytes = 0;
  PassStep step[24];
};

// Multiplier m with x / d == __umulhi(x, m) for 0 <= x < 2^32 / d (d >= 2);
// 0 for d == 1.
inline unsigned div_mul(int d) {
  return d <= 1 ? 0u : (unsigned)(((1ull << 32) + (unsigned long long)d - 1) / (unsigned long long)d);
}

// Radix sequence for n (7-smooth), preferring radix 8 and keeping the
// smallest radix >= 4 for powers of two >= 4; order 1 reverses it.  Returns
// false if n has a prime factor > 7.
bool factor_radices(int n, PassPlan &p, int order);

// Scale of a transform (src/fftgrid.cpp:1817,1848-1859): forward 1/sqrt(N)
// except covariance grids (unscaled); inverse 1/sqrt(N), covariance 1/N.
// Float transforms round the factor to float as FFTGrid does; transforms with
// double arithmetic (f64) keep it in double, so that a forward + inverse pair
// of DATA transforms scales by 1/N to double accuracy.
inline double grid_fft_scale(bool covariance, bool f64, long long N, bool inverse) {
  if (!inverse) {
    if (covariance) return 1.0;
    return f64 ? 1.0 / std::sqrt((double)N) : (double)(1.0f / std::sqrt((float)N));
  }
  if (covariance) return f64 ? 1.0 / (double)N : (double)(float)(1.0 / (double)N);
  return f64 ? 1.0 / std::sqrt((double)N) : (double)(float)(1.0 / std::sqrt((float)N));
}

// Variants of the generic (runtime-length) kernels: radix order x pencils per workgroup.
constexpr int kNumGenericVariants = 6;

class Fft3D {
 public:
  // prec: 0 float, 1 float storage with double arithmetic inside each pass
  // (every pass rounds its outputs once to float), 2 double.
  Fft3D(int nxp, int nyp, int nzp, int prec);
  ~Fft3D();
  int init();  // CRAVA_HIP_OK or an error code
  // In-place forward r2c of `ng` (<= kMaxGrids) grids of this size in one
  // launch per pass; grid i's result is multiplied by scale[i] (1 if NULL).
  // Mean deflation of the grids whose bit is set in dcmask (dc: device
  // memory for kMaxGrids doubles, private to this call's stream): the
  // transform runs on x - c, c = the grid's first value, and c N dscale is
  // added to the DC bin in double; a float transform then keeps the
  // accuracy of the grid's variation instead of its level.  dscale: the
  // exact (double) scale of each grid for the level (scale if NULL), so that
  // the level survives a forward + inverse pair exactly while the kernels
  // apply `scale` as their storage type rounds it.
  int r2c(void *const *grids, int ng, const double *scale, hipStream_t s, double *dc = nullptr,
          unsigned dcmask = 0, const double *dscale = nullptr);
  // In-place inverse c2r of a set of grids, scaled as r2c.  Mean deflation:
  // the DC bin v is zeroed before the passes and v * dscale is added to
  // every output value.
  int c2r(void *const *grids, int ng, const double *scale, hipStream_t s, double *dc = nullptr,
          unsigned dcmask = 0, const double *dscale = nullptr);
  int r2c(void *grid, hipStream_t s, double scale) { return r2c(&grid, 1, &scale, s); }
  int c2r(void *grid, hipStream_t s, double scale) { return c2r(&grid, 1, &scale, s); }
  // One pass (axis 0 x rows, 1 y, 2 z) over a set; `scale` applies on the x
  // pass, and so do dc / dcmask (as in r2c / c2r).
  int pass(int axis, void *const *grids, int ng, bool inverse, const double *scale, hipStream_t s,
           const double *dc = nullptr, unsigned dcmask = 0);
  long long complex_elems() const;

  int nxp, nyp, nzp;
  int prec;
  bool dbl;  // double arithmetic (LDS, twiddles)
  PassPlan px, py, pz;
  void *twx = nullptr, *twy = nullptr, *twz = nullptr;  // device twiddle tables
  void *stw[3] = {nullptr, nullptr, nullptr};  // per-pass tables of the specialised kernels, per axis
  // Launch variant per axis (x, y, z): v >= 0 specialised kernel variant v,
  // v < 0 generic kernel variant -1 - v.
  int var[3] = {-1, -1, -1};
  bool gdio[3] = {false, false, false};  // generic pass with direct first-pass loads (variants -7..-12)
  int set_variant(int axis, int v);
  int num_variants(int axis) const;  // specialised variants available
  // Time all variants (forward + inverse pass) on the set `grids` (overwritten)
  // and keep the fastest per axis.
  int autotune(void *const *grids, int ng, hipStream_t s, int reps, float *ms_out);

 private:
  int build_plan(int n, PassPlan &p, int generic_variant);
  int make_set(int axis, bool inverse, void *const *grids, int ng, const double *scale, GridSet &gs, int &io) const;
  int dc_kernel(int which, void *const *grids, int ng, const double *dscale, double *dc, unsigned dcmask,
                hipStream_t s) const;
};

}  // namespace crava
