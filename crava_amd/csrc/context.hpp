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
  std::map<std::tuple<int, int, int, bool>, std::unique_ptr<Fft3D>> plans;
  int plan(int nxp, int nyp, int nzp, bool dbl, Fft3D **out);
  ~Context();
};
}  // namespace crava

struct crava_hip_ctx : crava::Context {};
