// Grid-wide kernels of FFTGrid and its callers (see grid.hip).
#pragma once
#include "common.hpp"
#include "fft_static.hpp"

namespace crava {

int launch_fill_param_corr(const GridSet &gs, int nxp, int nyp, int nzp, const float *corr_xy, const float *circ_t,
                           float grad_i, float grad_j, hipStream_t s);
int launch_fill_err_corr(float *grid, int nxp, int nyp, int nzp, const float *corr_xy, float grad_i, float grad_j,
                         hipStream_t s);
// n floats (real domain) or n complex values (transformed).
int launch_grid_op(int op, float *a, const float *b, long long n, bool transformed, float scalar, hipStream_t s);
// Synchronous: out2 (host) = {sum of squares, count}; `work` device doubles of
// data_variance_work_doubles(nyp, nzp).
int data_variance(const float *grid, int nx, int ny, int nz, int nxp, int nyp, int nzp, double *work,
                  double *out2, hipStream_t s);
size_t data_variance_work_doubles(int nyp, int nzp);
int launch_facies_geomodel(const GridSet &probs, const float *undef, int nfac, int nx, int ny, int nz,
                           hipStream_t s);

}  // namespace crava
