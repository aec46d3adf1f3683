// Local-noise correction of the posterior (see localnoise.hip).
#pragma once
#include "common.hpp"

namespace crava {

constexpr int kLncMaxTheta = 8;

// Trace-independent inputs of the per-trace matrices.
struct LncConst {
  double g[kLncMaxTheta][3];           // G of Crava::computeG
  double e[kLncMaxTheta][kLncMaxTheta];   // errThetaCov
  double sm[3][3];                  // priorVar0
  double sdi[3][3];                 // (priorVar0 - postVar0)^(-1/2), eigenvalues <= 1e-7 dropped
  double old[3][3];                 // postVar0^(-1/2), eigenvalues <= 0 dropped
  double min_scale[kLncMaxTheta];      // minimum of each angle's noise-scale table
};

// Symmetric matrix function V diag(f(l)) V^T on the host (f = 1/sqrt with
// the given threshold: l > thr kept).
void lnc_inv_sqrt_host(const double in[3][3], double thr, double out[3][3]);

// Crava::computeG on the host.  The eigenvectors of
// Sm^(-1/2) (Sm - Smd)+ Sm^(-1/2) come from the restatement of LAPACK
// DGEEV's path (dgeev3) where it handles the matrix, else from Jacobi
// rotations signed so that each one's component of largest magnitude is
// positive.  E: [ntheta][ntheta], Sm, Smd: [3][3], G: [ntheta][3],
// all row-major.
int lnc_compute_g(int ntheta, const double *E, const double *Sm, const double *Smd, double *G);

// LncConst from G [ntheta][3], errThetaCov [ntheta][ntheta], priorVar0 and
// postVar0 (row-major); min_scale set to 1.
void lnc_const_host(int ntheta, const double *G, const double *E, const double *Sm, const double *Smd, LncConst &c);
// Crava::newPosteriorCovPointwise's sigmanew (row-major 3x3) for one scale
// vector [ntheta] (D = diag(sqrt(scale))), on the host.
int lnc_sigmanew_host(const LncConst &c, int ntheta, const double *scale, double *sigmanew);

// scales: device [ntheta][ny][nx]; sigmanew, trace_cov: device [ny][nx][9]
// (trace_cov may be nullptr).
int launch_lnc_trace(const LncConst &c, int ntheta, const double *scales, int nx, int ny, double *sigmanew,
                     double *trace_cov, hipStream_t s);
int launch_lnc_apply(float *const post[3], const float *const mean2[3], const double *sigmanew, int nx, int ny,
                     int nz, int nyp, int rnxp, hipStream_t s);
int launch_lnc_realisation(float *const seed[3], const double *trace_cov, int nx, int ny, int nz, int nyp, int rnxp,
                           hipStream_t s);

}  // namespace crava
