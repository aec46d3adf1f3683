// Per-Fourier-point Bayesian AVO posterior (Crava::computePostMeanResidAndFFTCov,
// src/crava.cpp:1042-1160) as one batched kernel.
#pragma once
#include "common.hpp"

#include <vector>

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

// Queue of the mixed form: the 64-point waves that the FP32 kernel leaves to
// the FP64 kernel, as {k-plane, first point of the wave}, and their count.
struct DeferQueue {
  int2 *list = nullptr;        // device, cap entries
  unsigned *count = nullptr;   // device
  long long cap = 0;
};

// Entries a mixed posterior of this problem can queue: every in-band wave.
long long defer_capacity(const PostProblem &pp, const unsigned char *invert);

struct PostDevice {
  PostProblem pp;
  PerK<float> *tabf = nullptr;   // device, nzp entries
  PerK<double> *tabd = nullptr;  // device, nzp entries
  long long defer_need = 0;      // defer_capacity of the problem
  // device, nzp entries: the in-band k-planes in increasing order, then the
  // out-of-band ones (or NULL: one launch over every plane); nk_in in band
  int *kmap = nullptr;
  int nk_in = 0;
  std::vector<unsigned char> invert;  // host copy of the frequency gate, nzp entries
  // the error-correlation spectrum is one Fourier plane [nyp][nxp/2 + 1],
  // the same on every kz plane (crava_hip_err_corr_plane)
  bool ec_plane = false;
};

// kmap of a problem from its frequency gate (invert[k] != 0: in band).
std::vector<int> posterior_kmap(int nzp, const unsigned char *invert, int *nk_in);

// precision: 0 = FP32 (grids complex64, the reference's float algorithm),
// 1 = mixed (complex64 grids; FP32 square-root information solve, and the
// FP64 reference-form solve for every wave that holds a point below the
// noise/signal threshold or whose FP32 factorisations fail; `defer` must
// hold defer_need entries), 2 = FP64 (complex128 grids).  Launches: FP32 /
// FP64: one; mixed: queue reset, FP32 kernel, FP64 kernel over the queue.
// corr: NULL, or the one Fourier-domain correlation grid of a prior
// Sigma0 (x) corr (the cov grids are then outputs only; see PostArgs::corr).
int launch_posterior(const PostDevice &pd, int precision, void *const *mu, void *const *data, void *const *cov,
                     const void *errcorr, unsigned long long *counters, const DeferQueue *defer, hipStream_t s,
                     const void *corr = nullptr);

// One volume of a batched posterior: its problem and device grids.
struct PostVolume {
  const PostDevice *pd = nullptr;
  void *mu[3] = {nullptr, nullptr, nullptr};
  void *data[kMaxTheta] = {nullptr};
  void *cov[6] = {nullptr};
  const void *errcorr = nullptr;
  const void *corr = nullptr;               // separable prior: its correlation grid (else NULL)
  unsigned long long *counters = nullptr;  // device [2]: failed Cholesky, FP64 points
};

// Launch forms of a batched posterior: one persistent launch over every
// tile (registers of the largest ntheta), one persistent launch per ntheta
// class, or one launch of the single-volume kernel per volume.
enum PostForm { kFormSingle = 0, kFormPerClass = 1, kFormPerVolume = 2 };

// The posterior of many volumes (uneven sizes and ntheta) over a tile list:
// ppb points of one k-plane per tile, `blocks` persistent workgroups
// striding over the tiles (0: one workgroup per tile).  Tiles are ordered by
// ntheta, so each class is one contiguous segment.  launch() serves the
// first two forms; the per-volume form is launched by the batch executor.
struct PostBatch {
  int precision = 0, nvol = 0, ppb = 0, blocks = 0;
  int form = kFormPerClass;
  void *args = nullptr;       // device PostArgs per volume
  void *queues = nullptr;     // mixed form: storage of every volume's queue
  int4 *tiles = nullptr;      // device {volume, k, first point, ntheta}
  long long ntiles = 0;
  // in-band tiles of ntheta c: [seg[c-1], seg[c]); the out-of-band tiles of
  // every volume (the prior passes through): [seg[kMaxTheta], ntiles)
  long long seg[kMaxTheta + 1] = {0};
  std::vector<std::vector<unsigned char>> inv;  // each volume's frequency gate
  std::vector<int> nt, nzp;
  std::vector<long long> plane;
  ~PostBatch();
  int init(int precision, const PostVolume *vols, int nvol);
  int set_launch(int points_per_block, int blocks);
  int num_classes() const;
  int launch(const hipStream_t *streams, int nstreams) const;
  // Mixed form around launch(): clear every volume's queue before it, and run
  // the FP64 kernel over every queue after it.
  int reset_queues(hipStream_t s) const;
  int launch_queued(hipStream_t s) const;
  // Mixed form: the queue of each volume (each volume has its own, even
  // when volumes share a problem).
  std::vector<DeferQueue> vq;
};

}  // namespace crava
