// Facies-probability grid kernel (see facies.hip).
#pragma once
#include "common.hpp"

namespace crava {
constexpr int kMaxFacies = 16;

constexpr int kMaxDensAngles = 8;   // densdim = 2^nAng <= 256

// The lattice of one density table (the Simbox of FindDensity).
struct FaciesVolume {
  double x0, y0, dx, dy, ztop, zbot;
  int n0, n1, n2, pad;
};

struct FaciesCall {
  int nx, ny, nz, nxp, nyp, nzp;
  const float *alpha, *beta, *rho;  // device, padded real grids
  int nfac;
  const float *tables;              // device, ndens x nfac tables [tn3][tn2][2*(tn1/2+1)], index d*nfac + l
  int tn1, tn2, tn3;
  const double *vol;                // host: x0, y0, dx, dy, ztop, zbot of density 0
  const int *vn;                    // host: volume nx, ny, nz of density 0
  const float *prior;               // host, nfac
  float p_undefined;
  float *const *prob;               // host array of nfac+1 device grids
  int *argmax;                      // device [nz][ny][nx] or null
  unsigned long long *counts;       // device [nfac] or null
  // local-noise density dimension (ndens = 2^nang, nang = 0: one density)
  int ndens = 1, nang = 0;
  const FaciesVolume *vols = nullptr;  // device [ndens]
  const float *tgrid = nullptr;        // device [nang][ny][nx] angle weights t
};

int launch_facies(const FaciesCall &c, hipStream_t s);
}  // namespace crava
