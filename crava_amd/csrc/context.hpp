// Execution context behind crava_hip_ctx.
#pragma once
#include "common.hpp"
#include "fft.hpp"

#include <map>
#include <memory>
#include <tuple>

namespace crava {
struct Context {
  hipStream_t stream = nullptr;
  unsigned long long *counters = nullptr;  // small device counters
  double *dcbuf = nullptr;                 // levels of a mean-deflated transform set (kMaxGrids)
  std::map<std::tuple<int, int, int, int>, std::unique_ptr<Fft3D>> plans;
  // prec: 0 float, 1 float storage / double arithmetic, 2 double
  int plan(int nxp, int nyp, int nzp, int prec, Fft3D **out);
  // Device workspace of at least `bytes`, kept and grown across calls.
  void *work = nullptr;
  size_t work_bytes = 0;
  int workspace(size_t bytes, void **out);
  ~Context();
};
}  // namespace crava

struct crava_hip_ctx : crava::Context {};
