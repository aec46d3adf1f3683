// Many inversion volumes of uneven size and ntheta as one captured sequence:
// per volume the forward transforms of (data, error correlation) and of
// (means, covariances) as two grid sets, the posterior, then the inverse
// transforms of (residuals, covariances) and of the means.
// The sequence runs eagerly or as a hipGraph captured once and replayed.
#pragma once
#include "common.hpp"
#include "context.hpp"
#include "posterior.hpp"

#include <vector>

namespace crava {

struct BatchVolume {
  const PostDevice *pd = nullptr;
  void *mu[3] = {nullptr, nullptr, nullptr};
  void *data[kMaxTheta] = {nullptr};
  void *cov[6] = {nullptr};
  void *errcorr = nullptr;
  void *corr = nullptr;   // separable prior: its correlation grid (cov[] are outputs)
};

class BatchExec {
 public:
  static constexpr int kStreams = 4;
  BatchExec(Context *ctx, int precision) : ctx_(ctx), precision_(precision) {}
  ~BatchExec();
  int init(const BatchVolume *vols, int n);
  // One pass of the whole sequence on `s`; graph: capture on first use, replay after.
  int run(hipStream_t s, bool graph);
  // Choose FFT variants per grid shape and the posterior launch (form x
  // points per block x workgroups, see PostForm) by timing; the registered
  // grids are restored.
  int autotune(hipStream_t s, int *ppb, int *blocks, int *form, float *post_ms);
  int set_launch(int ppb, int blocks, int form);
  // Per-volume counters of the last run (synchronises `s`).
  int counts(hipStream_t s, long long *n_fail, long long *n_fp64);
  // One eager pass with events between its phases: ms[0] forward
  // transforms, ms[1] posterior, ms[2] inverse transforms (synchronises).
  int phase_times(hipStream_t s, float *ms);
  int nvol() const { return (int)vols_.size(); }

 private:
  struct Set {
    Fft3D *plan = nullptr;
    int ng = 0;
    void *grids[kMaxGrids] = {nullptr};
    double scale[kMaxGrids] = {0};
    double exact[kMaxGrids] = {0};  // level scales of the mean-deflated grids
    unsigned dcmask = 0;             // grids transformed with mean deflation
    double *dc = nullptr;            // device, kMaxGrids levels
  };
  struct Work {
    Set fwd[2], inv[2];   // fwd: [0] data, error correlation, means; [1] covariances;
                          // inv: [0] residuals, covariances; [1] means
    int stream[2] = {0, 0};
  };
  int build_sets();
  int enqueue(hipStream_t origin, bool transforms_only, bool forward_only, hipEvent_t *marks = nullptr);
  int launch_post(hipStream_t origin);
  int fork(hipStream_t origin);
  int join(hipStream_t origin);
  int backup_grids(hipStream_t s, bool restore);
  int check_ec_form() const;
  void drop_graph();

  Context *ctx_;
  int precision_;
  std::vector<BatchVolume> vols_;
  // per volume: its error correlation is one Fourier plane (taken from its
  // problem at init; not transformed here)
  std::vector<unsigned char> ec_plane_;
  std::vector<Work> work_;
  PostBatch post_;
  unsigned long long *counters_ = nullptr;  // [nvol][2]
  double *dcbufs_ = nullptr;                 // [nvol][4][kMaxGrids] deflation levels
  hipStream_t side_[kStreams] = {nullptr};
  hipEvent_t fork_ev_ = nullptr;
  hipEvent_t join_ev_[kStreams] = {nullptr};
  hipStream_t cap_ = nullptr;
  hipGraph_t graph_ = nullptr;
  hipGraphExec_t exec_ = nullptr;
  void *backup_ = nullptr;
  size_t backup_bytes_ = 0;
};

}  // namespace crava
