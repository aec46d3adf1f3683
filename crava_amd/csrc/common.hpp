// Shared device/host helpers for the CRAVA MI355X kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <cstdint>
#include <string>

namespace crava {

// Complex value with the memory layout of fftw_complex ({re, im}).
template <typename T>
struct alignas(2 * sizeof(T)) cx {
  T x, y;
};

template <typename T>
__host__ __device__ __forceinline__ cx<T> mk(T a, T b) { cx<T> r; r.x = a; r.y = b; return r; }
template <typename T>
__host__ __device__ __forceinline__ cx<T> add(cx<T> a, cx<T> b) { return mk(a.x + b.x, a.y + b.y); }
template <typename T>
__host__ __device__ __forceinline__ cx<T> sub(cx<T> a, cx<T> b) { return mk(a.x - b.x, a.y - b.y); }
template <typename T>
__host__ __device__ __forceinline__ cx<T> mul(cx<T> a, cx<T> b) { return mk(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x); }
// a * conj(b)
template <typename T>
__host__ __device__ __forceinline__ cx<T> mulc(cx<T> a, cx<T> b) { return mk(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y); }
// conj(a) * b
template <typename T>
__host__ __device__ __forceinline__ cx<T> cmul(cx<T> a, cx<T> b) { return mk(a.x * b.x + a.y * b.y, a.x * b.y - a.y * b.x); }
template <typename T>
__host__ __device__ __forceinline__ cx<T> scl(cx<T> a, T s) { return mk(a.x * s, a.y * s); }
template <typename T>
__host__ __device__ __forceinline__ T norm2(cx<T> a) { return a.x * a.x + a.y * a.y; }

// Workgroups are dispatched round-robin over the 8 XCDs, each with its own
// L2.  Map dispatch-order id L (of N workgroups) to a work id so that every
// XCD receives one contiguous range of work ids: neighbouring tiles, which
// share HBM lines, then meet in the same L2.  A bijection on [0, N).
__device__ __forceinline__ unsigned xcd_work_id(unsigned L, unsigned N) {
  constexpr unsigned kXcd = 8;
  const unsigned x = L % kXcd, i = L / kXcd, q = N / kXcd, r = N % kXcd;
  return x * q + (x < r ? x : r) + i;
}

// Last error message of the calling thread (exposed through the C ABI).
void set_error(const std::string &msg);
const char *get_error();

}  // namespace crava

#define CRAVA_HIP_CHECK(expr)                                                     \
  do {                                                                            \
    hipError_t e_ = (expr);                                                       \
    if (e_ != hipSuccess) {                                                       \
      crava::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));        \
      return CRAVA_HIP_E_HIP;                                                     \
    }                                                                             \
  } while (0)
