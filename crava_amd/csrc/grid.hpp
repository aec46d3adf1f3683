// Grid-wide kernels of FFTGrid and its callers (see grid.hip).
#pragma once
#include "common.hpp"
#include "fft_static.hpp"

namespace crava {

int launch_fill_param_corr(const GridSet &gs, int nxp, int nyp, int nzp, const float *corr_xy, const float *circ_t,
                           float grad_i, float grad_j, hipStream_t s);
int launch_fill_err_corr(float *grid, int nxp, int nyp, int nzp, const float *corr_xy, float grad_i, float grad_j,
                         hipStream_t s);
// Plane k = 0 of the error-correlation grid with zero gradients, as float or
// double, into the real-layout plane [nyp][2 (nxp / 2 + 1)].
int launch_fill_err_corr_plane(void *plane, bool f64, int nxp, int nyp, int nzp, const float *corr_xy,
                               hipStream_t s);
// Planes 0 .. nzp - 1 of a separable correlation spectrum from the 2-D
// spectrum of corrXY in plane 0 and the 1-D transform of circ_t (device).
int launch_corr_outer(void *grid, bool f64, int nxp, int nyp, int nzp, const float *circ_t, hipStream_t s);
// n floats (real domain) or n complex values (transformed).
int launch_grid_op(int op, float *a, const float *b, long long n, bool transformed, float scalar, hipStream_t s);
// Time-lapse merge (grid.hip): n complex points per grid.
struct Merge4D {
  const cx<float> *ms[3], *md[3];            // static / dynamic means
  cx<float> *mu[3];                          // merged means
  const cx<float> *ss[6], *dd[6], *sd[9];    // static-static, dynamic-dynamic, static-dynamic
  cx<float> *sig[6];                         // merged covariance (upper triangle order)
  long long n;
};
int launch_merge_4d(const Merge4D &a, bool cov, hipStream_t s);
// Synchronous: out2 (host) = {sum of squares, count}; `work` device doubles of
// data_variance_work_doubles(nyp, nzp).
int data_variance(const float *grid, int nx, int ny, int nz, int nxp, int nyp, int nzp, double *work,
                  double *out2, hipStream_t s);
size_t data_variance_work_doubles(int nyp, int nzp);
int launch_facies_geomodel(const GridSet &probs, const float *undef, int nfac, int nx, int ny, int nz,
                           hipStream_t s);

}  // namespace crava
