// Facies-probability grid kernel (see facies.hip).
#pragma once
#include "common.hpp"

namespace crava {
constexpr int kMaxFacies = 16;

struct FaciesCall {
  int nx, ny, nz, nxp, nyp, nzp;
  const float *alpha, *beta, *rho;  // device, padded real grids
  int nfac;
  const float *tables;              // device, nfac tables [tn3][tn2][2*(tn1/2+1)]
  int tn1, tn2, tn3;
  const double *vol;                // host: x0, y0, dx, dy, ztop, zbot
  const int *vn;                    // host: volume nx, ny, nz
  const float *prior;               // host, nfac
  float p_undefined;
  float *const *prob;               // host array of nfac+1 device grids
  int *argmax;                      // device [nz][ny][nx] or null
  unsigned long long *counts;       // device [nfac] or null
};

int launch_facies(const FaciesCall &c, hipStream_t s);
}  // namespace crava
