// Per-Fourier-point Bayesian AVO posterior (Crava::computePostMeanResidAndFFTCov,
// src/crava.cpp:1042-1160) as one batched kernel.
#pragma once
#include "common.hpp"

namespace crava {

constexpr int kMaxTheta = 8;
constexpr int kTri = kMaxTheta * (kMaxTheta + 1) / 2;

// Everything that is constant over one k-plane (src/crava.cpp:1042-1083),
// pre-combined on the host so that the per-point noise matrix is
//   E(p) = lambda(p) * Ec + diag(Ed)       (src/crava.cpp:1298-1334)
// and the per-point signal matrix is K S K^H = P o (A S A^T) with
//   P_lm = kWD_l conj(kWD_m), K_lc = kWD_l A_lc.
// Triangles are packed row-major over l >= m: q = l(l+1)/2 + m.
template <typename T>
struct PerK {
  cx<T> kwd[kMaxTheta];
  cx<T> P[kTri];
  cx<T> Ec[kTri];
  T Ed[kMaxTheta];
  T trEc;
  T trEd;
  int inv;
  int pad;
};

// Packed Gaussian-posterior problem description (host side).
struct PostProblem {
  int nxp = 0, nyp = 0, nzp = 0, ntheta = 0;
  float wnc = 0.1f;
  double sigma0[9] = {0};
  float A[kMaxTheta * 3] = {0};
  double err_theta_cov[kMaxTheta * kMaxTheta] = {0};
  double mixed_threshold = 1e-3;
};

// Build the per-k tables from the reference-level per-k vectors
// (invert, kWD, e1, e2, e3 as in src/crava.cpp:1044-1057).
void build_perk(const PostProblem &pp, const unsigned char *invert, const float *kWD, const float *e1,
                const float *e2, const float *e3, PerK<float> *tf, PerK<double> *td);

struct PostDevice {
  PostProblem pp;
  PerK<float> *tabf = nullptr;   // device, nzp entries
  PerK<double> *tabd = nullptr;  // device, nzp entries
};

// precision: 0 = FP32 (grids complex64), 1 = mixed (complex64 grids, FP64 per
// point below the noise/signal threshold), 2 = FP64 (complex128 grids).
int launch_posterior(const PostDevice &pd, int precision, void *const *mu, void *const *data, void *const *cov,
                     const void *errcorr, unsigned long long *counters, hipStream_t s);

}  // namespace crava
