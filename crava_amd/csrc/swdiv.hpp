// Non-constant-thickness data preparation kernels (see swdiv.hip): the
// wavelet division of every trace and the interpolation / extrapolation of
// low-energy traces.
#pragma once
#include <vector>

#include "common.hpp"

namespace crava {

constexpr int kSwdMaxTheta = 8;

struct SwdArgs {
  float *g[kSwdMaxTheta];            // real padded grids [nzp][nyp][rnxp], divided in place
  const cx<float> *W;                // [nt][cn] global wavelet half spectra
  const float *t3;                   // [nt][cn] (2 cos t - 2) sin t / |E3|_band, t = 2 pi m / nzp
  const float *tau2;                 // [nt][cn] |Re| of the reflection-coefficient covariance spectrum
  const float *wg2;                  // [nt][cn] |W|^2
  const float *shift[kSwdMaxTheta];  // per angle [ny][nx] local time shift, or nullptr (0)
  const float *gain[kSwdMaxTheta];   // per angle [ny][nx] local gain, or nullptr (1)
  const double *rel;                 // [nyp][nxp] relative thickness of every padded trace
  float mod_w[kSwdMaxTheta];         // |w|^2 maxfrequency / highCut
  float ev[kSwdMaxTheta];            // error variance of the angle
  float iwn2[kSwdMaxTheta];          // 1 / (|w| * |w|)
  float dz, wnc, low_cut, high_cut;
  float sq;                          // sqrtf(nzp)
  float inv_n;                       // 1 / nzp (the layout's slack traces)
  double lz;
  int nt, nx, ny, nz, nxp, nyp, nzp, rnxp;
};

// Host tables of one angle (cn = nzp/2 + 1 bins): W {re, im}, t3, tau2, wg2
// and the scalar mod_w.
void swd_angle_tables(int nz, int nzp, const float *wavelet, float wnorm, const float *A3, const double *var0,
                      const float *corr_t, float dz, double lz, float low_cut, float high_cut, float *W_re_im,
                      float *t3, float *tau2, float *wg2, float *mod_w);

int launch_swd_prepare(const SwdArgs &a, hipStream_t s);
int launch_swd_filter(const SwdArgs &a, hipStream_t s);
int launch_swd_finalize(const SwdArgs &a, hipStream_t s);
// One fused kernel for the whole division when nzp has a specialised
// transform; tw is the plan's per-pass twiddle table of length nzp (float).
bool swd_fused_available(int nzp);
int launch_swd_fused(const SwdArgs &a, const void *tw, hipStream_t s);

// interpolateSeismic
struct IsqOp {
  int trace;  // j * nx + i
  int mask;   // neighbours averaged: 1 left, 2 right, 4 up, 8 down
};

// The interpolation schedule of FFTGrid::interpolateSeismic, derived from
// the per-trace energies alone: the traces written, in the reference's
// order, each tagged with the neighbours it averages, grouped into levels
// whose operations are independent of each other.
struct IsqSchedule {
  std::vector<IsqOp> ops;         // sorted by level
  std::vector<int> level_start;   // ops of level v: [level_start[v], level_start[v + 1])
  int imin, imax, jmin, jmax;     // box of the traces above the threshold
  int n_flagged, n_zero;
};
void isq_schedule(const float *energy, int nx, int ny, float threshold, IsqSchedule *out);

int launch_isq_energy(const float *g, int nx, int ny, int nz, int nyp, int rnxp, float *energy, hipStream_t s);
int launch_isq_interp(float *g, const IsqOp *ops, int nops, int nx, int nyp, int nzp, int rnxp, hipStream_t s);
int launch_isq_extrapolate(float *g, int nx, int ny, int nz, int nxp, int nyp, int nzp, int rnxp, int imin,
                           int imax, int jmin, int jmax, hipStream_t s);

}  // namespace crava
