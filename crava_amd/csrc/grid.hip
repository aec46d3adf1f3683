// Grid-wide kernels of FFTGrid and its callers on the padded FFTGrid layout:
// the prior and error correlation fills, the element-wise operators, the
// data-variance reduction and the facies geomodel update.  Each reproduces
// the reference's float arithmetic operation by operation (this file is
// compiled with contraction off), so results are bit-identical.
//
// Every kernel is one pass over HBM: a grid-stride loop over the rsize (or
// csize) values of the grid, 256 threads per workgroup.
#include "grid.hpp"
#include "../../include/crava_hip.h"

#include <algorithm>
#include <cmath>
#include <vector>

namespace crava {

constexpr float kRMissing = -99999.0f;  // RMISSING, src/definitions.h:35

static unsigned blocks_for(long long n) {
  return (unsigned)std::max<long long>(1, std::min<long long>((n + 255) / 256, 1 << 20));
}

// FFTGrid::fillInParamCorr (src/fftgrid.cpp:1119-1164): the correlation at
// (i, j, k), i < rnxp.
__device__ __forceinline__ float param_corr_value(int i, int j, int k, int nxp, int nyp, int nzp,
                                                  const float *__restrict__ corr_xy, const float *__restrict__ circ_t,
                                                  float grad_i, float grad_j) {
  if (i >= nxp) return kRMissing;
  int cycle_i = i - nxp;
  if (i < -cycle_i) cycle_i = i;
  int cycle_j = j - nyp;
  if (j < -cycle_j) cycle_j = j;
  float sub_k = (float)k + (float)cycle_i * grad_i + (float)cycle_j * grad_j;
  int base_k = (int)floorf(sub_k);
  sub_k -= (float)base_k;
  while (base_k < 0) base_k += nzp;
  while (base_k >= nzp) base_k -= nzp;
  float value = (1.0f - sub_k) * circ_t[base_k];
  if (base_k != nzp - 1) value += sub_k * circ_t[base_k + 1];
  else value += sub_k * circ_t[0];
  value *= corr_xy[i + nxp * j];
  return value;
}

// FFTGrid::fillInErrCorr (src/fftgrid.cpp:1065-1116), range 1: the value at
// (i, j, k), i < rnxp.
__device__ __forceinline__ float err_corr_value(int i, int j, int k, int nxp, int nyp, int nzp,
                                                const float *__restrict__ corr_xy, float grad_i, float grad_j) {
  int cycle_i = i - nxp;
  if (i < -cycle_i) cycle_i = i;
  int cycle_j = j - nyp;
  if (j < -cycle_j) cycle_j = j;
  float sub_k = (float)k + (float)cycle_i * grad_i + (float)cycle_j * grad_j;
  if (i >= nxp) return kRMissing;
  if (fabsf(sub_k) < 1.0f || fabsf(sub_k - (float)nzp) < 1.0f) {
    int base_k = (int)sub_k;
    sub_k -= (float)base_k;
    while (base_k < -1) base_k += nzp;
    while (base_k >= 1) base_k -= nzp;
    if (base_k < 0) {
      base_k = -1 - base_k;
      sub_k = 1.0f - sub_k;
    }
    return (1.0f - sub_k) * corr_xy[i + nxp * j];
  }
  return 0.0f;
}

// The fills write two consecutive x values per thread as one float2 (rnxp
// is even, so a pair never crosses a row): consecutive threads take
// consecutive pairs, and the pair index is split into (x, y, z) with 32-bit
// arithmetic (IDX = unsigned) whenever the grid has fewer than 2^32 pairs,
// 64-bit otherwise.

// fillInParamCorr followed by multiplyByScalar(scale[g])
// (src/seismicparametersholder.cpp:210-215) for each grid g of the set: the
// same correlation written to up to 16 grids.
template <typename IDX>
__global__ void __launch_bounds__(256) k_fill_param_corr(GridSet gs, int nxp, int nyp, int nzp,
                                                         const float *__restrict__ corr_xy,
                                                         const float *__restrict__ circ_t, float grad_i,
                                                         float grad_j) {
  const IDX cn = (IDX)(nxp / 2 + 1);
  const IDX npair = cn * (IDX)nyp * (IDX)nzp;
  for (IDX q = (IDX)blockIdx.x * 256u + threadIdx.x; q < npair; q += (IDX)gridDim.x * 256u) {
    const IDX row = q / cn;
    const int h = (int)(q - row * cn);
    const int j = (int)(row % (IDX)nyp), k = (int)(row / (IDX)nyp);
    const float v0 = param_corr_value(2 * h, j, k, nxp, nyp, nzp, corr_xy, circ_t, grad_i, grad_j);
    const float v1 = param_corr_value(2 * h + 1, j, k, nxp, nyp, nzp, corr_xy, circ_t, grad_i, grad_j);
    for (int g = 0; g < gs.n; ++g) {
      const float sc = (float)gs.scale[g];
      reinterpret_cast<float2 *>(gs.dst[g])[q] = make_float2(v0 * sc, v1 * sc);
    }
  }
}

template <typename IDX>
__global__ void __launch_bounds__(256) k_fill_err_corr(float *grid, int nxp, int nyp, int nzp,
                                                       const float *__restrict__ corr_xy, float grad_i,
                                                       float grad_j) {
  const IDX cn = (IDX)(nxp / 2 + 1);
  const IDX npair = cn * (IDX)nyp * (IDX)nzp;
  for (IDX q = (IDX)blockIdx.x * 256u + threadIdx.x; q < npair; q += (IDX)gridDim.x * 256u) {
    const IDX row = q / cn;
    const int h = (int)(q - row * cn);
    const int j = (int)(row % (IDX)nyp), k = (int)(row / (IDX)nyp);
    const float v0 = err_corr_value(2 * h, j, k, nxp, nyp, nzp, corr_xy, grad_i, grad_j);
    const float v1 = err_corr_value(2 * h + 1, j, k, nxp, nyp, nzp, corr_xy, grad_i, grad_j);
    reinterpret_cast<float2 *>(grid)[q] = make_float2(v0, v1);
  }
}

// Plane k = 0 of fillInErrCorr with zero gradients (the only plane that is
// not zero then), stored as T: rnxp / 2 pairs of each of the nyp rows.
template <typename T>
__global__ void __launch_bounds__(256) k_fill_err_corr_plane(T *plane, int nxp, int nyp, int nzp,
                                                             const float *__restrict__ corr_xy) {
  const int cn = nxp / 2 + 1;
  const int npair = cn * nyp;
  for (int q = (int)(blockIdx.x * 256u + threadIdx.x); q < npair; q += (int)(gridDim.x * 256u)) {
    const int j = q / cn, h = q - j * cn;
    plane[2 * q] = (T)err_corr_value(2 * h, j, 0, nxp, nyp, nzp, corr_xy, 0.0f, 0.0f);
    plane[2 * q + 1] = (T)err_corr_value(2 * h + 1, j, 0, nxp, nyp, nzp, corr_xy, 0.0f, 0.0f);
  }
}

// Spectrum of a separable correlation grid corrXY(i, j) circT(k) (the grid
// fillInParamCorr writes with zero gradients): plane 0 holds the 2-D
// spectrum C of corrXY; plane kz of the output is C T(kz), with
// T(kz) = sum_m circT[m] exp(-2 pi i kz m / nzp) summed in double by each
// workgroup (fixed order) and the product rounded once.  Planes kz0 ..
// kz0 + gridDim.y - 1; a launch that writes plane 0 reads and writes each
// point in one thread, so it runs after the launch over the other planes.
template <typename T>
__global__ void __launch_bounds__(256) k_corr_outer(T *grid, long long plane, int nzp, int kz0,
                                                    const float *__restrict__ circ_t) {
  __shared__ double red[2][256];
  const int kz = kz0 + (int)blockIdx.y;
  const int tid = (int)threadIdx.x;
  double sr = 0.0, si = 0.0;
  for (int m = tid; m < nzp; m += 256) {
    const long long r = ((long long)kz * m) % nzp;
    double sn, cs;
    sincospi(-2.0 * (double)r / (double)nzp, &sn, &cs);
    sr += (double)circ_t[m] * cs;
    si += (double)circ_t[m] * sn;
  }
  red[0][tid] = sr;
  red[1][tid] = si;
  __syncthreads();
  for (int h = 128; h > 0; h >>= 1) {
    if (tid < h) {
      red[0][tid] += red[0][tid + h];
      red[1][tid] += red[1][tid + h];
    }
    __syncthreads();
  }
  const double tr = red[0][0], ti = red[1][0];
  for (long long p = (long long)blockIdx.x * 256 + tid; p < plane; p += (long long)gridDim.x * 256) {
    const double cr = (double)grid[2 * p], ci = (double)grid[2 * p + 1];
    const long long o = (long long)kz * plane + p;
    grid[2 * o] = (T)(cr * tr - ci * ti);
    grid[2 * o + 1] = (T)(cr * ti + ci * tr);
  }
}

// Element-wise operators (src/fftgrid.cpp:1877-1999), over n floats (real
// domain) or n complex values (transformed).
__global__ void k_grid_op(int op, float *a, const float *__restrict__ b, long long n, int transformed, float s) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
    if (transformed) {
      float2 *ac = reinterpret_cast<float2 *>(a);
      const float2 *bc = reinterpret_cast<const float2 *>(b);
      float2 v = ac[p];
      switch (op) {
        case CRAVA_HIP_OP_ADD: { const float2 w = bc[p]; v.x += w.x; v.y += w.y; break; }
        case CRAVA_HIP_OP_SUBTRACT: { const float2 w = bc[p]; v.x -= w.x; v.y -= w.y; break; }
        case CRAVA_HIP_OP_MULTIPLY: {
          const float2 w = bc[p], t = v;
          v.x = w.x * t.x - w.y * t.y;
          v.y = w.y * t.x + w.x * t.y;
          break;
        }
        case CRAVA_HIP_OP_CHANGE_SIGN: v.x = -v.x; v.y = -v.y; break;
        case CRAVA_HIP_OP_CONJUGATE: v.y = -v.y; break;
        default: break;
      }
      ac[p] = v;
    } else {
      float v = a[p];
      switch (op) {
        case CRAVA_HIP_OP_ADD: v += b[p]; break;
        case CRAVA_HIP_OP_SUBTRACT: v -= b[p]; break;
        case CRAVA_HIP_OP_MULTIPLY: v *= b[p]; break;
        case CRAVA_HIP_OP_CHANGE_SIGN: v = -v; break;
        case CRAVA_HIP_OP_MULTIPLY_SCALAR: v *= s; break;
        case CRAVA_HIP_OP_ADD_SCALAR: v += s; break;
        case CRAVA_HIP_OP_FILL_CONSTANT: v = s; break;
        default: break;
      }
      a[p] = v;
    }
  }
}

// Copy of a whole grid (rsize floats, an even count), as the FFTGrid copy
// constructor copies rvalue_ (src/fftgrid.cpp:69-100): 8-byte vector loads
// and stores, the plain device copy the HBM counters are calibrated with.
__global__ void k_grid_copy(float2 *__restrict__ dst, const float2 *__restrict__ src, long long n2) {
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n2; p += (long long)gridDim.x * blockDim.x)
    dst[p] = src[p];
}

// Crava::computeDataVariance (src/crava.cpp:322-363) in the reference's
// summation order: per (j,k) row a double sum over i (one thread per row),
// per k a sum over rows, then over k.
__global__ void k_row_sumsq(const float *grid, int nx, int ny, int nz, int rnxp, int nyp, int nzp, double *rowsum,
                            long long *rowcnt) {
  const long long nrows = (long long)nyp * nzp;
  for (long long r = (long long)blockIdx.x * blockDim.x + threadIdx.x; r < nrows;
       r += (long long)gridDim.x * blockDim.x) {
    const int j = (int)(r % nyp), k = (int)(r / nyp);
    double s = 0.0;
    long long c = 0;
    if (k < nz && j < ny) {
      const float *row = grid + r * rnxp;
      for (int i = 0; i < nx; ++i) {
        const float t = row[i];
        if (t != 0.0f) {
          s += (double)(t * t);
          ++c;
        }
      }
    }
    rowsum[r] = s;
    rowcnt[r] = c;
  }
}

__global__ void k_plane_sum(const double *rowsum, const long long *rowcnt, int nyp, int nzp, double *planesum,
                            long long *planecnt) {
  const int k = blockIdx.x * blockDim.x + threadIdx.x;
  if (k >= nzp) return;
  double s = 0.0;
  long long c = 0;
  for (int j = 0; j < nyp; ++j) {
    s += rowsum[(long long)k * nyp + j];
    c += rowcnt[(long long)k * nyp + j];
  }
  planesum[k] = s;
  planecnt[k] = c;
}

__global__ void k_total_sum(const double *planesum, const long long *planecnt, int nzp, double *out) {
  if (blockIdx.x != 0 || threadIdx.x != 0) return;
  double s = 0.0;
  long long c = 0;
  for (int k = 0; k < nzp; ++k) {
    s += planesum[k];
    c += planecnt[k];
  }
  out[0] = s;
  out[1] = (double)c;
}

// FaciesProb::calculateFaciesProbGeomodel (src/faciesprob.cpp:2286-2330),
// constant prior probabilities (probs.scale[l]): prob_l <- prior_l * undef +
// prob_l inside the volume, RMISSING in the x padding, on grids of
// nz x ny x 2*(nx/2+1).
__global__ void k_facies_geomodel(GridSet probs, const float *__restrict__ undef, int nfac, int nx, int ny, int nz) {
  const int rnx = 2 * (nx / 2 + 1);
  const long long n = (long long)rnx * ny * nz;
  for (long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x; p < n; p += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(p % rnx);
    const float u = undef[p];
    for (int l = 0; l < nfac; ++l) {
      float *g = static_cast<float *>(probs.dst[l]);
      const float v = (float)probs.scale[l] * u + g[p];
      g[p] = i < nx ? v : kRMissing;
    }
  }
}

// Time-lapse merge of the static and dynamic parts into the current state,
// transformed grids (State4D::merge / mergeCov, src/state4d.cpp:90-236).
// Mean: mu_l = mu_static_l + mu_dynamic_l.  Covariance, l <= m:
//   Sigma_lm = ((F[l][m] + F[l+3][m+3]) + F[l+3][m]) + F[l][m+3],
// F the 6x6 joint prior: F[a][b] = ss(a,b), F[3+a][3+b] = dd(a,b),
// F[a][3+b] = sd[3a+b], F[3+b][a] = conj(sd[3a+b]); float, in the
// reference's order.  One thread per complex point.
__global__ void k_merge_mean_4d(Merge4D a) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= a.n) return;
#pragma unroll
  for (int l = 0; l < 3; ++l) {
    const cx<float> s = a.ms[l][p], d = a.md[l][p];
    a.mu[l][p] = mk(s.x + d.x, s.y + d.y);
  }
}

__global__ void k_merge_cov_4d(Merge4D a) {
  const long long p = (long long)blockIdx.x * 256 + threadIdx.x;
  if (p >= a.n) return;
  // upper-triangle order of the six grids: (0,0) (0,1) (0,2) (1,1) (1,2) (2,2)
  const int ql[6] = {0, 0, 0, 1, 1, 2}, qm[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    const int l = ql[q], m = qm[q];
    const cx<float> f0 = a.ss[q][p], f1 = a.dd[q][p];
    const cx<float> g = a.sd[3 * m + l][p];          // F[m][l+3]; F[l+3][m] is its conjugate
    const cx<float> f3 = a.sd[3 * l + m][p];         // F[l][m+3]
    float re = f0.x;
    re += f1.x;
    re += g.x;
    re += f3.x;
    float im = f0.y;
    im += f1.y;
    im += -g.y;
    im += f3.y;
    a.sig[q][p] = mk(re, im);
  }
}

int launch_merge_4d(const Merge4D &a, bool cov, hipStream_t s) {
  const unsigned blocks = (unsigned)((a.n + 255) / 256);
  if (cov) hipLaunchKernelGGL(k_merge_cov_4d, dim3(blocks), dim3(256), 0, s, a);
  else hipLaunchKernelGGL(k_merge_mean_4d, dim3(blocks), dim3(256), 0, s, a);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

// Value pairs of a padded grid (32-bit indices below 2^32 - 2^28 pairs: the
// grid-stride increment, at most 2^20 workgroups of 256, cannot overflow).
static long long fill_pairs(int nxp, int nyp, int nzp) { return (long long)(nxp / 2 + 1) * nyp * nzp; }

int launch_fill_param_corr(const GridSet &gs, int nxp, int nyp, int nzp, const float *corr_xy, const float *circ_t,
                           float grad_i, float grad_j, hipStream_t s) {
  const long long n = fill_pairs(nxp, nyp, nzp);
  if (n < (1LL << 32) - 256LL * 1048576LL)
    hipLaunchKernelGGL(k_fill_param_corr<unsigned>, dim3(blocks_for(n)), dim3(256), 0, s, gs, nxp, nyp, nzp, corr_xy,
                       circ_t, grad_i, grad_j);
  else
    hipLaunchKernelGGL(k_fill_param_corr<unsigned long long>, dim3(blocks_for(n)), dim3(256), 0, s, gs, nxp, nyp, nzp,
                       corr_xy, circ_t, grad_i, grad_j);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_fill_err_corr(float *grid, int nxp, int nyp, int nzp, const float *corr_xy, float grad_i, float grad_j,
                         hipStream_t s) {
  const long long n = fill_pairs(nxp, nyp, nzp);
  if (n < (1LL << 32) - 256LL * 1048576LL)
    hipLaunchKernelGGL(k_fill_err_corr<unsigned>, dim3(blocks_for(n)), dim3(256), 0, s, grid, nxp, nyp, nzp, corr_xy,
                       grad_i, grad_j);
  else
    hipLaunchKernelGGL(k_fill_err_corr<unsigned long long>, dim3(blocks_for(n)), dim3(256), 0, s, grid, nxp, nyp, nzp,
                       corr_xy, grad_i, grad_j);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_fill_err_corr_plane(void *plane, bool f64, int nxp, int nyp, int nzp, const float *corr_xy,
                               hipStream_t s) {
  const long long n = fill_pairs(nxp, nyp, 1);
  if (f64)
    hipLaunchKernelGGL(k_fill_err_corr_plane<double>, dim3(blocks_for(n)), dim3(256), 0, s,
                       static_cast<double *>(plane), nxp, nyp, nzp, corr_xy);
  else
    hipLaunchKernelGGL(k_fill_err_corr_plane<float>, dim3(blocks_for(n)), dim3(256), 0, s,
                       static_cast<float *>(plane), nxp, nyp, nzp, corr_xy);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_corr_outer(void *grid, bool f64, int nxp, int nyp, int nzp, const float *circ_t, hipStream_t s) {
  const long long plane = (long long)(nxp / 2 + 1) * nyp;
  const unsigned gx = (unsigned)std::min<long long>((plane + 255) / 256, 64);
  // planes 1 .. nzp - 1 from plane 0, then plane 0 in place
  for (int pass = 0; pass < 2; ++pass) {
    const int kz0 = pass == 0 ? 1 : 0, ny = pass == 0 ? nzp - 1 : 1;
    if (ny < 1) continue;
    if (f64)
      hipLaunchKernelGGL(k_corr_outer<double>, dim3(gx, (unsigned)ny), dim3(256), 0, s, static_cast<double *>(grid),
                         plane, nzp, kz0, circ_t);
    else
      hipLaunchKernelGGL(k_corr_outer<float>, dim3(gx, (unsigned)ny), dim3(256), 0, s, static_cast<float *>(grid),
                         plane, nzp, kz0, circ_t);
    CRAVA_HIP_CHECK(hipGetLastError());
  }
  return CRAVA_HIP_OK;
}

int launch_grid_op(int op, float *a, const float *b, long long n, bool transformed, float scalar, hipStream_t s) {
  if (op == CRAVA_HIP_OP_COPY) {
    const long long n2 = transformed ? n : n / 2;   // float2 elements
    hipLaunchKernelGGL(k_grid_copy, dim3(blocks_for(n2)), dim3(256), 0, s, reinterpret_cast<float2 *>(a),
                       reinterpret_cast<const float2 *>(b), n2);
    CRAVA_HIP_CHECK(hipGetLastError());
    return CRAVA_HIP_OK;
  }
  hipLaunchKernelGGL(k_grid_op, dim3(blocks_for(n)), dim3(256), 0, s, op, a, b, n, transformed ? 1 : 0, scalar);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int data_variance(const float *grid, int nx, int ny, int nz, int nxp, int nyp, int nzp, double *work,
                  double *out2, hipStream_t s) {
  const long long nrows = (long long)nyp * nzp;
  double *rowsum = work;
  long long *rowcnt = reinterpret_cast<long long *>(work + nrows);
  double *planesum = work + 2 * nrows;
  long long *planecnt = reinterpret_cast<long long *>(work + 2 * nrows + nzp);
  double *tot = work + 2 * nrows + 2 * nzp;
  const int rnxp = 2 * (nxp / 2 + 1);
  hipLaunchKernelGGL(k_row_sumsq, dim3(blocks_for(nrows)), dim3(256), 0, s, grid, nx, ny, nz, rnxp, nyp, nzp, rowsum,
                     rowcnt);
  hipLaunchKernelGGL(k_plane_sum, dim3((nzp + 255) / 256), dim3(256), 0, s, rowsum, rowcnt, nyp, nzp, planesum,
                     planecnt);
  hipLaunchKernelGGL(k_total_sum, dim3(1), dim3(64), 0, s, planesum, planecnt, nzp, tot);
  CRAVA_HIP_CHECK(hipGetLastError());
  CRAVA_HIP_CHECK(hipMemcpyAsync(out2, tot, 2 * sizeof(double), hipMemcpyDeviceToHost, s));
  CRAVA_HIP_CHECK(hipStreamSynchronize(s));
  return CRAVA_HIP_OK;
}

size_t data_variance_work_doubles(int nyp, int nzp) { return 2 * (size_t)nyp * nzp + 2 * (size_t)nzp + 2; }

int launch_facies_geomodel(const GridSet &probs, const float *undef, int nfac, int nx, int ny, int nz,
                           hipStream_t s) {
  const long long n = 2LL * (nx / 2 + 1) * ny * nz;
  hipLaunchKernelGGL(k_facies_geomodel, dim3(blocks_for(n)), dim3(256), 0, s, probs, undef, nfac, nx, ny, nz);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

}  // namespace crava
