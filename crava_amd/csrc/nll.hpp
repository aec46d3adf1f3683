// Negative log marginal likelihood and its adjoint (see nll.hip).
#pragma once
#include "common.hpp"
#include "posterior.hpp"

namespace crava {

// Per-k constants of the likelihood at the current parameters.
struct NllTab {
  cx<double> kwd[kMaxTheta];  // s_l kD conj(W_l)
  cx<double> Ec[kTri];        // c0 Sigma_e,lm g_lm (diagonal real)
  cx<double> g[kTri];         // e1_l conj(e1_m) + e2_l conj(e2_m)
  double Ed[kMaxTheta];       // wnc Sigma_e,ll e3_l^2
  double wd[kMaxTheta];       // wnc e3_l^2
  double c0;                  // (1 - wnc) / 2
  double pad;
};

struct NllCall {
  int nt, nxp, nyp;
  bool dbl;
  const void *mu[3];
  const void *data[kMaxTheta];
  const void *cov0;
  const void *errcorr;
  bool ec_plane;  // errcorr is one Fourier plane, the same on every kz plane
  const NllTab *tab;
  const int *klist;
  int n_inband;
  double A[kMaxTheta * 3];
  double sig0[6];
  double base00;
  double *partial;
  int blocks;
};

int nll_accumulator_count(int nt);
int launch_nll(const NllCall &c, hipStream_t s);

}  // namespace crava
