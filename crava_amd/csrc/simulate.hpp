// Posterior simulation kernels (see simulate.hip).
#pragma once
#include "common.hpp"

namespace crava {
int launch_simulate_points(long long npts, const void *const *cov, void *const *seed, unsigned long long *fails,
                           hipStream_t s);
int launch_add_real(float *dst, const float *src, long long n, hipStream_t s);
int launch_fill_noise(void *grid, int nxp, int nyp, int nzp, uint64_t seed, uint64_t stream, hipStream_t s);
}  // namespace crava
