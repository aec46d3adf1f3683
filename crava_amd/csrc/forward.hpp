// Forward modelling kernels (see forward.hip).
#pragma once
#include "common.hpp"

namespace crava {

constexpr int kFwdMaxTheta = 8;

struct FwdArgs {
  const float *alpha, *beta, *rho;    // real padded grids [nzp][nyp][rnxp]
  float *out[kFwdMaxTheta];            // one output grid per angle, same layout
  float A[kFwdMaxTheta][3];            // reflection matrix rows
  const cx<float> *W;                  // [nt][nzp/2+1] wavelet half spectra
  const float *sf[kFwdMaxTheta];       // per angle [ny][nx] stretch, or nullptr (1)
  const float *shift[kFwdMaxTheta];    // per angle [ny][nx] time shift (same unit as dz), or nullptr (0)
  const float *gain[kFwdMaxTheta];     // per angle [ny][nx] gain, or nullptr (1)
  float dz;
  float inv_n;                         // 1 / nzp
  int nt, nx, ny, nz, nyp, nzp, rnxp;
};

int launch_fm_reflect(const FwdArgs &a, hipStream_t s);
int launch_fm_filter(const FwdArgs &a, hipStream_t s);
int launch_fm_padding(const FwdArgs &a, hipStream_t s);
// One fused kernel for the whole model when nzp has a specialised transform;
// tw is the plan's per-pass twiddle table of length nzp (float).
bool fm_fused_available(int nzp);
int launch_fm_fused(const FwdArgs &a, const void *tw, hipStream_t s);

}  // namespace crava
