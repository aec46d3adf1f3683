// Batch executor (see batch.hpp).  No reference counterpart: the reference
// inverts one volume per Crava object (main.cpp:279-308 runs vintages one
// after another); each volume here follows the sequence of
// Crava::computePostMeanResidAndFFTCov (src/crava.cpp:887-1204).
//
// Volumes are spread over a few side streams forked from the origin stream,
// so small volumes overlap.  Within a volume the two transform sets of each
// direction go to two different streams: in the mixed form the covariance
// transforms run double arithmetic inside each pass while the float
// transforms of the other set stream HBM, so the two overlap.  Volumes of
// one grid shape share transform plans, so each kind of set of a shape
// always goes to the same stream.
#include "batch.hpp"
#include "../../include/crava_hip.h"

#include <algorithm>
#include <map>
#include <tuple>

namespace crava {

__global__ void k_zero_u64(unsigned long long *p, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = 0ULL;
}

static size_t real_grid_bytes(const PostProblem &pp, bool f64) {
  return (size_t)2 * (pp.nxp / 2 + 1) * pp.nyp * pp.nzp * (f64 ? 8 : 4);
}

// Bytes of a volume's error-correlation buffer: a grid, or one Fourier plane.
static size_t errcorr_bytes(const PostProblem &pp, bool f64, bool plane) {
  return plane ? (size_t)2 * (pp.nxp / 2 + 1) * pp.nyp * (f64 ? 8 : 4) : real_grid_bytes(pp, f64);
}

BatchExec::~BatchExec() {
  drop_graph();
  for (int i = 0; i < kStreams; ++i) {
    if (side_[i]) (void)hipStreamDestroy(side_[i]);
    if (join_ev_[i]) (void)hipEventDestroy(join_ev_[i]);
  }
  if (fork_ev_) (void)hipEventDestroy(fork_ev_);
  if (cap_) (void)hipStreamDestroy(cap_);
  if (counters_) (void)hipFree(counters_);
  if (dcbufs_) (void)hipFree(dcbufs_);
  if (backup_) (void)hipFree(backup_);
}

void BatchExec::drop_graph() {
  if (exec_) (void)hipGraphExecDestroy(exec_);
  if (graph_) (void)hipGraphDestroy(graph_);
  exec_ = nullptr;
  graph_ = nullptr;
}

int BatchExec::init(const BatchVolume *vols, int n) {
  if (!vols || n < 1) {
    set_error("a batch needs at least one volume");
    return CRAVA_HIP_E_INVALID;
  }
  if (precision_ < CRAVA_HIP_F32 || precision_ > CRAVA_HIP_F64) {
    set_error("invalid batch precision");
    return CRAVA_HIP_E_INVALID;
  }
  vols_.assign(vols, vols + n);
  for (int v = 0; v < n; ++v) {
    const BatchVolume &bv = vols_[v];
    if (!bv.pd || !bv.errcorr) {
      set_error("volume " + std::to_string(v) + ": NULL problem or error-correlation grid");
      return CRAVA_HIP_E_INVALID;
    }
    for (int c = 0; c < 3; ++c)
      if (!bv.mu[c]) { set_error("volume " + std::to_string(v) + ": NULL mean grid"); return CRAVA_HIP_E_INVALID; }
    for (int q = 0; q < 6; ++q)
      if (!bv.cov[q]) { set_error("volume " + std::to_string(v) + ": NULL covariance grid"); return CRAVA_HIP_E_INVALID; }
    for (int l = 0; l < bv.pd->pp.ntheta; ++l)
      if (!bv.data[l]) { set_error("volume " + std::to_string(v) + ": NULL data grid"); return CRAVA_HIP_E_INVALID; }
  }
  ec_plane_.assign(n, 0);
  for (int v = 0; v < n; ++v) ec_plane_[v] = vols_[v].pd->ec_plane ? 1 : 0;
  CRAVA_HIP_CHECK(hipMalloc(&counters_, sizeof(unsigned long long) * 2 * n));
  CRAVA_HIP_CHECK(hipMalloc(&dcbufs_, sizeof(double) * 4 * kMaxGrids * (size_t)n));
  int rc;
  if ((rc = build_sets())) return rc;
  std::vector<PostVolume> pv(n);
  for (int v = 0; v < n; ++v) {
    pv[v].pd = vols_[v].pd;
    for (int c = 0; c < 3; ++c) pv[v].mu[c] = vols_[v].mu[c];
    for (int l = 0; l < kMaxTheta; ++l) pv[v].data[l] = vols_[v].data[l];
    for (int q = 0; q < 6; ++q) pv[v].cov[q] = vols_[v].cov[q];
    pv[v].errcorr = vols_[v].errcorr;
    pv[v].corr = vols_[v].corr;
    pv[v].counters = counters_ + 2 * v;
  }
  if ((rc = post_.init(precision_, pv.data(), n))) return rc;
  for (int i = 0; i < kStreams; ++i) {
    CRAVA_HIP_CHECK(hipStreamCreateWithFlags(&side_[i], hipStreamNonBlocking));
    CRAVA_HIP_CHECK(hipEventCreateWithFlags(&join_ev_[i], hipEventDisableTiming));
  }
  CRAVA_HIP_CHECK(hipEventCreateWithFlags(&fork_ev_, hipEventDisableTiming));
  CRAVA_HIP_CHECK(hipStreamCreateWithFlags(&cap_, hipStreamNonBlocking));
  return CRAVA_HIP_OK;
}

// Transform sets per volume and the stream of each volume.  Forward: [0] the
// data (DATA scaling), the error correlation (COVARIANCE) and the means
// (DATA); [1] the covariances (COVARIANCE).  Inverse: [0] the residuals and
// the covariances, [1] the means.  Plan precision: the float form uses float
// transforms throughout; the mixed form transforms the covariances forward
// with double arithmetic inside each pass (float storage: the small prior
// eigenvalues need it), the means in float with mean deflation (the
// posterior mean update needs the accuracy of the means' variation, not of
// their level) and the rest in float; the double form uses double
// transforms.
int BatchExec::build_sets() {
  const bool f64 = precision_ == CRAVA_HIP_F64;
  const int pa = f64 ? 2 : 0;
  const int pc = f64 ? 2 : (precision_ == CRAVA_HIP_MIXED ? 1 : 0);
  const bool dc = pc != 0;   // the covariance set keeps double scale factors
  work_.assign(vols_.size(), Work{});
  std::map<std::tuple<int, int, int>, std::vector<int>> by_shape;
  for (size_t v = 0; v < vols_.size(); ++v) {
    const BatchVolume &bv = vols_[v];
    const PostProblem &pp = bv.pd->pp;
    const long long N = (long long)pp.nxp * pp.nyp * pp.nzp;
    Work &w = work_[v];
    int rc;
    Set *fa = &w.fwd[0], *fc = &w.fwd[1], *ia = &w.inv[0], *ic = &w.inv[1];
    if ((rc = ctx_->plan(pp.nxp, pp.nyp, pp.nzp, pa, &fa->plan))) return rc;
    if ((rc = ctx_->plan(pp.nxp, pp.nyp, pp.nzp, pc, &fc->plan))) return rc;
    ia->plan = fa->plan;
    ic->plan = fa->plan;
    for (int i = 0; i < 4; ++i) (i < 2 ? w.fwd[i] : w.inv[i - 2]).dc = dcbufs_ + ((size_t)v * 4 + i) * kMaxGrids;
    const bool deflate = precision_ == CRAVA_HIP_MIXED;
    const double sf = grid_fft_scale(false, f64, N, false), si = grid_fft_scale(false, f64, N, true);
    const double xf = grid_fft_scale(false, true, N, false), xi = grid_fft_scale(false, true, N, true);
    const double cf = grid_fft_scale(true, dc, N, false);
    const double ef = grid_fft_scale(true, f64, N, false), ei = grid_fft_scale(true, f64, N, true);
    for (int l = 0; l < pp.ntheta; ++l) {
      fa->grids[fa->ng] = bv.data[l];
      fa->scale[fa->ng++] = sf;
      ia->grids[ia->ng] = bv.data[l];
      ia->scale[ia->ng++] = si;
    }
    if (!ec_plane_[v]) {  // a one-plane spectrum is already transformed
      fa->grids[fa->ng] = bv.errcorr;
      fa->scale[fa->ng++] = ef;
    }
    for (int c = 0; c < 3; ++c) {
      if (deflate) {
        fa->dcmask |= 1u << fa->ng;
        ic->dcmask |= 1u << ic->ng;
      }
      fa->exact[fa->ng] = xf;
      fa->grids[fa->ng] = bv.mu[c];
      fa->scale[fa->ng++] = sf;
      ic->exact[ic->ng] = xi;
      ic->grids[ic->ng] = bv.mu[c];
      ic->scale[ic->ng++] = si;
    }
    if (bv.corr) {  // separable prior: one correlation transform
      fc->grids[fc->ng] = bv.corr;
      fc->scale[fc->ng++] = cf;
    }
    for (int q = 0; q < 6; ++q) {
      if (!bv.corr) {
        fc->grids[fc->ng] = bv.cov[q];
        fc->scale[fc->ng++] = cf;
      }
      ia->grids[ia->ng] = bv.cov[q];
      ia->scale[ia->ng++] = ei;
    }
    by_shape[std::make_tuple(pp.nxp, pp.nyp, pp.nzp)].push_back((int)v);
  }
  // greedy balance of grid bytes over the side streams, one stream per shape
  std::vector<std::pair<size_t, std::vector<int>>> groups;
  for (auto &kv : by_shape) {
    size_t bytes = 0;
    for (int v : kv.second) {
      const PostProblem &pp = vols_[v].pd->pp;
      bytes += real_grid_bytes(pp, f64) * (size_t)(10 + pp.ntheta);  // every grid of the volume
    }
    groups.emplace_back(bytes, kv.second);
  }
  std::sort(groups.begin(), groups.end(), [](const auto &a, const auto &b) { return a.first > b.first; });
  size_t load[kStreams] = {0};
  for (auto &g : groups) {
    const int s = (int)(std::min_element(load, load + kStreams) - load);
    load[s] += g.first;
    for (int v : g.second) {
      work_[v].stream[0] = s;
      work_[v].stream[1] = (s + kStreams / 2) % kStreams;
    }
  }
  return CRAVA_HIP_OK;
}

int BatchExec::fork(hipStream_t origin) {
  CRAVA_HIP_CHECK(hipEventRecord(fork_ev_, origin));
  for (int i = 0; i < kStreams; ++i) CRAVA_HIP_CHECK(hipStreamWaitEvent(side_[i], fork_ev_, 0));
  return CRAVA_HIP_OK;
}

int BatchExec::join(hipStream_t origin) {
  for (int i = 0; i < kStreams; ++i) {
    CRAVA_HIP_CHECK(hipEventRecord(join_ev_[i], side_[i]));
    CRAVA_HIP_CHECK(hipStreamWaitEvent(origin, join_ev_[i], 0));
  }
  return CRAVA_HIP_OK;
}

int BatchExec::enqueue(hipStream_t origin, bool transforms_only, bool forward_only, hipEvent_t *marks) {
  int rc;
  const int n = (int)vols_.size();
  if (!transforms_only) {
    hipLaunchKernelGGL(k_zero_u64, dim3((2 * n + 255) / 256), dim3(256), 0, origin, counters_, 2 * n);
    CRAVA_HIP_CHECK(hipGetLastError());
  }
  if (marks) CRAVA_HIP_CHECK(hipEventRecord(marks[0], origin));
  if ((rc = fork(origin))) return rc;
  for (int v = 0; v < n; ++v)
    for (int i = 0; i < 2; ++i) {
      const Set &st = work_[v].fwd[i];
      if ((rc = st.plan->r2c(st.grids, st.ng, st.scale, side_[work_[v].stream[i]], st.dc, st.dcmask, st.exact)))
        return rc;
    }
  if ((rc = join(origin))) return rc;
  if (forward_only) return CRAVA_HIP_OK;
  if (marks) CRAVA_HIP_CHECK(hipEventRecord(marks[1], origin));
  if (!transforms_only && (rc = launch_post(origin))) return rc;
  if (marks) CRAVA_HIP_CHECK(hipEventRecord(marks[2], origin));
  if ((rc = fork(origin))) return rc;
  for (int v = 0; v < n; ++v)
    for (int i = 0; i < 2; ++i) {
      const Set &st = work_[v].inv[i];
      if ((rc = st.plan->c2r(st.grids, st.ng, st.scale, side_[work_[v].stream[i]], st.dc, st.dcmask, st.exact)))
        return rc;
    }
  if ((rc = join(origin))) return rc;
  if (marks) CRAVA_HIP_CHECK(hipEventRecord(marks[3], origin));
  return CRAVA_HIP_OK;
}

int BatchExec::phase_times(hipStream_t s, float *ms) {
  hipEvent_t ev[4] = {nullptr, nullptr, nullptr, nullptr};
  int rc = check_ec_form();
  for (int i = 0; i < 4 && rc == CRAVA_HIP_OK; ++i)
    if (hipEventCreate(&ev[i]) != hipSuccess) rc = CRAVA_HIP_E_HIP;
  if (rc == CRAVA_HIP_OK) rc = enqueue(s, false, false, ev);
  if (rc == CRAVA_HIP_OK && hipEventSynchronize(ev[3]) != hipSuccess) rc = CRAVA_HIP_E_HIP;
  for (int i = 0; i < 3 && rc == CRAVA_HIP_OK; ++i)
    if (hipEventElapsedTime(&ms[i], ev[i], ev[i + 1]) != hipSuccess) rc = CRAVA_HIP_E_HIP;
  for (int i = 0; i < 4; ++i)
    if (ev[i]) (void)hipEventDestroy(ev[i]);
  if (rc == CRAVA_HIP_E_HIP) set_error("batch phase timing: event failure");
  return rc;
}

int BatchExec::launch_post(hipStream_t origin) {
  int rc;
  if (post_.form == kFormPerVolume) {
    if ((rc = fork(origin))) return rc;
    for (size_t v = 0; v < vols_.size(); ++v) {
      const BatchVolume &bv = vols_[v];
      const DeferQueue *q = post_.vq.empty() ? nullptr : &post_.vq[v];
      if ((rc = launch_posterior(*bv.pd, precision_, bv.mu, bv.data, bv.cov, bv.errcorr, counters_ + 2 * v, q,
                                 side_[work_[v].stream[0]], bv.corr)))
        return rc;
    }
    return join(origin);
  }
  if ((rc = post_.reset_queues(origin))) return rc;
  if (post_.form == kFormSingle || post_.num_classes() < 2) {
    if ((rc = post_.launch(&origin, 1))) return rc;
  } else {
    if ((rc = fork(origin))) return rc;
    if ((rc = post_.launch(side_, kStreams))) return rc;
    if ((rc = join(origin))) return rc;
  }
  return post_.launch_queued(origin);
}

// The problems' error-correlation forms must be those the batch was built with.
int BatchExec::check_ec_form() const {
  for (size_t v = 0; v < vols_.size(); ++v)
    if (vols_[v].pd->ec_plane != (ec_plane_[v] != 0)) {
      set_error("volume " + std::to_string(v) + ": error-correlation form changed after the batch was created");
      return CRAVA_HIP_E_INVALID;
    }
  return CRAVA_HIP_OK;
}

int BatchExec::run(hipStream_t s, bool graph) {
  if (int rc = check_ec_form()) return rc;
  if (!graph) return enqueue(s, false, false);
  if (!exec_) {
    CRAVA_HIP_CHECK(hipStreamBeginCapture(cap_, hipStreamCaptureModeRelaxed));
    int rc = enqueue(cap_, false, false);
    hipGraph_t g = nullptr;
    const hipError_t e = hipStreamEndCapture(cap_, &g);
    if (rc) {
      if (g) (void)hipGraphDestroy(g);
      return rc;
    }
    if (e != hipSuccess) {
      set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
      return CRAVA_HIP_E_HIP;
    }
    graph_ = g;
    const hipError_t ei = hipGraphInstantiate(&exec_, graph_, nullptr, nullptr, 0);
    if (ei != hipSuccess) {
      drop_graph();
      set_error(std::string("hipGraphInstantiate: ") + hipGetErrorString(ei));
      return CRAVA_HIP_E_HIP;
    }
  }
  CRAVA_HIP_CHECK(hipGraphLaunch(exec_, s));
  return CRAVA_HIP_OK;
}

int BatchExec::set_launch(int ppb, int blocks, int form) {
  if (form < kFormSingle || form > kFormPerVolume) {
    set_error("posterior launch form must be 0, 1 or 2");
    return CRAVA_HIP_E_INVALID;
  }
  const int rc = post_.set_launch(ppb, blocks);
  if (rc == CRAVA_HIP_OK) {
    post_.form = form;
    drop_graph();
  }
  return rc;
}

// Save (restore = false) or restore every registered grid to / from one
// device buffer.
int BatchExec::backup_grids(hipStream_t s, bool restore) {
  const bool f64 = precision_ == CRAVA_HIP_F64;
  size_t total = 0;
  for (const BatchVolume &bv : vols_)
    total += real_grid_bytes(bv.pd->pp, f64) * (size_t)(9 + bv.pd->pp.ntheta + (bv.corr ? 1 : 0)) +
             errcorr_bytes(bv.pd->pp, f64, bv.pd->ec_plane);
  if (!restore && total > backup_bytes_) {
    if (backup_) (void)hipFree(backup_);
    backup_ = nullptr;
    backup_bytes_ = 0;
    CRAVA_HIP_CHECK(hipMalloc(&backup_, total));
    backup_bytes_ = total;
  }
  unsigned char *b = static_cast<unsigned char *>(backup_);
  for (const BatchVolume &bv : vols_) {
    const size_t gb = real_grid_bytes(bv.pd->pp, f64);
    std::vector<void *> gl;
    for (int c = 0; c < 3; ++c) gl.push_back(bv.mu[c]);
    for (int l = 0; l < bv.pd->pp.ntheta; ++l) gl.push_back(bv.data[l]);
    for (int q = 0; q < 6; ++q) gl.push_back(bv.cov[q]);
    if (bv.corr) gl.push_back(bv.corr);
    gl.push_back(bv.errcorr);  // last: a grid, or one Fourier plane
    for (size_t i = 0; i < gl.size(); ++i) {
      const size_t nb = i + 1 < gl.size() ? gb : errcorr_bytes(bv.pd->pp, f64, bv.pd->ec_plane);
      if (restore) CRAVA_HIP_CHECK(hipMemcpyAsync(gl[i], b, nb, hipMemcpyDeviceToDevice, s));
      else CRAVA_HIP_CHECK(hipMemcpyAsync(b, gl[i], nb, hipMemcpyDeviceToDevice, s));
      b += nb;
    }
  }
  return CRAVA_HIP_OK;
}

int BatchExec::autotune(hipStream_t s, int *ppb_out, int *blocks_out, int *form_out, float *ms_out) {
  int rc;
  if ((rc = check_ec_form())) return rc;
  drop_graph();
  if ((rc = backup_grids(s, false))) return rc;
  // transforms: each plan once, on the largest set that uses it
  std::map<Fft3D *, const Set *> tune;
  for (const Work &w : work_)
    for (const Set *st : {&w.fwd[0], &w.fwd[1], &w.inv[0], &w.inv[1]}) {
      auto it = tune.find(st->plan);
      if (it == tune.end() || it->second->ng < st->ng) tune[st->plan] = st;
    }
  for (auto &kv : tune)
    if ((rc = kv.first->autotune(kv.second->grids, kv.second->ng, s, 3, nullptr))) return rc;

  // posterior launch: form x points per block x persistent workgroups (the
  // per-volume form has one geometry)
  hipEvent_t e0, e1;
  CRAVA_HIP_CHECK(hipEventCreate(&e0));
  CRAVA_HIP_CHECK(hipEventCreate(&e1));
  const int ppbs[3] = {64, 128, 256};
  const int blks[5] = {0, 1024, 2048, 4096, 8192};
  float best = 1e30f;
  int bp = 256, bb = 0, bf = kFormPerClass;
  bool warm = false;
  for (int m = 0; m < 3 && rc == CRAVA_HIP_OK; ++m)
    for (int a = 0; a < 3 && rc == CRAVA_HIP_OK; ++a)
      for (int b = 0; b < 5 && rc == CRAVA_HIP_OK; ++b) {
        if (m == kFormPerVolume && (a > 0 || b > 0)) continue;
        if ((rc = post_.set_launch(ppbs[a], blks[b]))) break;
        post_.form = m;
        for (int rep = warm ? 0 : 1; rep >= 0 && rc == CRAVA_HIP_OK; --rep) {
          if ((rc = backup_grids(s, true))) break;
          if ((rc = enqueue(s, true, true))) break;
          if (hipEventRecord(e0, s) != hipSuccess) { rc = CRAVA_HIP_E_HIP; break; }
          if ((rc = launch_post(s))) break;
          if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = CRAVA_HIP_E_HIP; break; }
          if (rep == 0) {
            float ms = 0.0f;
            (void)hipEventElapsedTime(&ms, e0, e1);
            if (ms < best) { best = ms; bp = ppbs[a]; bb = blks[b]; bf = m; }
          }
        }
        warm = true;
      }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) return rc;
  if ((rc = set_launch(bp, bb, bf))) return rc;
  if ((rc = backup_grids(s, true))) return rc;
  CRAVA_HIP_CHECK(hipStreamSynchronize(s));
  if (backup_) (void)hipFree(backup_);
  backup_ = nullptr;
  backup_bytes_ = 0;
  if (ppb_out) *ppb_out = bp;
  if (blocks_out) *blocks_out = bb;
  if (form_out) *form_out = bf;
  if (ms_out) *ms_out = best;
  set_error("");
  return CRAVA_HIP_OK;
}

int BatchExec::counts(hipStream_t s, long long *n_fail, long long *n_fp64) {
  const int n = (int)vols_.size();
  std::vector<unsigned long long> h(2 * (size_t)n);
  CRAVA_HIP_CHECK(hipMemcpyAsync(h.data(), counters_, sizeof(unsigned long long) * 2 * n, hipMemcpyDeviceToHost, s));
  CRAVA_HIP_CHECK(hipStreamSynchronize(s));
  for (int v = 0; v < n; ++v) {
    if (n_fail) n_fail[v] = (long long)h[2 * v];
    if (n_fp64) n_fp64[v] = (long long)h[2 * v + 1];
  }
  return CRAVA_HIP_OK;
}

}  // namespace crava
