// Posterior simulation on the half-complex layout (Crava::simulate,
// src/crava.cpp:1416-1599) and a counter-based complex-noise generator.
//
//   seed(p) <- chol(Sigma_post(p)) seed(p)   (0 where the factorisation fails)
//   seed    <- c2r(seed) (PARAMETER scaling) + posterior mean
//
// The noise generator replaces FFTGrid::fillInComplexNoise
// (src/fftgrid.cpp:1214-1263), whose sequential RandomGen stream cannot run
// in parallel: every point draws from Philox4x32-10 (Salmon, Moraes, Dror,
// Shaw, SC'11) keyed by (seed, stream) with the point's linear index as the
// counter, and a Box-Muller transform.  Points on the self-conjugate x planes
// (i = 0 and, for even nxp, i = nxp/2) take the conjugate of their
// (-j, -k) partner's draw, and self-conjugate points are real N(0,1), so the
// spectrum is that of a real white field.
#include "../../include/crava_hip.h"
#include "common.hpp"
#include "simulate.hpp"

#include <cmath>

namespace crava {

// 3x3 Hermitian Cholesky with the failure rule of lib_matrCholCpx
// (libs/lib/lib_matr.c:78-167); L * z as lib_matrProdCholVec (:453-470).
__device__ __forceinline__ bool chol3_mul(const cx<float> (&c)[6], cx<float> (&z)[3]) {
  // c: aa, bb, rr, ab, ar, br
  const float f = c[0].x;
  if (f <= 0.0f) return false;
  const float tol = 1e-20f * f;
  if (c[1].x <= tol || c[2].x <= tol) return false;
  const float l00 = sqrtf(f);
  const float r0 = 1.0f / l00;
  // row 1: L10 = conj(Sigma01) / l00  (Sigma10 = conj(Sigma01))
  const cx<float> l10 = mk(c[3].x * r0, -c[3].y * r0);
  float d1 = c[1].x - norm2(l10);
  if (d1 <= tol) return false;
  const float l11 = sqrtf(d1);
  const float r1 = 1.0f / l11;
  const cx<float> l20 = mk(c[4].x * r0, -c[4].y * r0);
  // L21 = (Sigma21 - L20 conj(L10)) / l11, Sigma21 = conj(Sigma12)
  const cx<float> s21 = sub(mk(c[5].x, -c[5].y), mulc(l20, l10));
  const cx<float> l21 = scl(s21, r1);
  float d2 = c[2].x - norm2(l20) - norm2(l21);
  if (d2 <= tol) return false;
  const float l22 = sqrtf(d2);
  const cx<float> z0 = z[0], z1 = z[1], z2 = z[2];
  z[0] = scl(z0, l00);
  z[1] = add(mul(l10, z0), scl(z1, l11));
  z[2] = add(add(mul(l20, z0), mul(l21, z1)), scl(z2, l22));
  return true;
}

struct CovPtrs {
  const cx<float> *p[6];
};

__global__ void __launch_bounds__(256) k_simulate_points(long long npts, CovPtrs cov, cx<float> *s0, cx<float> *s1,
                                                         cx<float> *s2, unsigned long long *fails) {
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  bool bad = false;
  if (p < npts) {
    cx<float> c[6];
#pragma unroll
    for (int q = 0; q < 6; ++q) c[q] = cov.p[q][p];
    cx<float> z[3] = {s0[p], s1[p], s2[p]};
    if (!chol3_mul(c, z)) {
      bad = true;
      z[0] = z[1] = z[2] = mk(0.0f, 0.0f);
    }
    s0[p] = z[0];
    s1[p] = z[1];
    s2[p] = z[2];
  }
  if (fails) {
    const unsigned long long b = __ballot(bad);
    if ((threadIdx.x & 63) == 0 && b) atomicAdd(fails, (unsigned long long)__popcll(b));
  }
}

// FFTGrid::add on the real layout (src/fftgrid.cpp:1877-1890)
__global__ void k_add_real(float *__restrict__ dst, const float *__restrict__ src, long long n) {
  long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long stride = (long long)gridDim.x * blockDim.x;
  for (; i < n; i += stride) dst[i] += src[i];
}

// ---- Philox4x32-10 ----------------------------------------------------
struct u4 {
  uint32_t x, y, z, w;
};

__device__ __forceinline__ u4 philox10(u4 c, uint32_t k0, uint32_t k1) {
  const uint32_t M0 = 0xD2511F53u, M1 = 0xCD9E8D57u, W0 = 0x9E3779B9u, W1 = 0xBB67AE85u;
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(M0, c.x), lo0 = M0 * c.x;
    const uint32_t hi1 = __umulhi(M1, c.z), lo1 = M1 * c.z;
    c = u4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += W0;
    k1 += W1;
  }
  return c;
}

__device__ __forceinline__ float u01(uint32_t v) {  // (0, 1]
  return ((float)(v >> 8) + 1.0f) * (1.0f / 16777216.0f);
}

// Two independent N(0,1) from one Philox draw (Box-Muller).
__device__ __forceinline__ void gauss2(uint64_t idx, uint64_t seed, uint64_t stream, float &a, float &b) {
  const u4 c{(uint32_t)idx, (uint32_t)(idx >> 32), (uint32_t)stream, (uint32_t)(stream >> 32)};
  const u4 r = philox10(c, (uint32_t)seed, (uint32_t)(seed >> 32));
  const float rad = sqrtf(-2.0f * logf(u01(r.x)));
  float s, co;
  sincospif(2.0f * u01(r.y), &s, &co);
  a = rad * co;
  b = rad * s;
}

__global__ void __launch_bounds__(256) k_fill_noise(cx<float> *__restrict__ g, int cnxp, int nyp, int nzp, int nxp,
                                                    uint64_t seed, uint64_t stream) {
  const long long n = (long long)cnxp * nyp * nzp;
  const long long p = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (p >= n) return;
  const int i = (int)(p % cnxp);
  const long long jk = p / cnxp;
  const int j = (int)(jk % nyp);
  const int k = (int)(jk / nyp);
  const bool selfplane = (i == 0) || ((nxp % 2 == 0) && i == nxp / 2);
  float a, b;
  if (!selfplane) {
    gauss2((uint64_t)p, seed, stream, a, b);
    g[p] = mk(a * 0.70710678118654752f, b * 0.70710678118654752f);
    return;
  }
  const int jc = j == 0 ? 0 : nyp - j;
  const int kc = k == 0 ? 0 : nzp - k;
  const long long jkc = (long long)kc * nyp + jc;
  if (jkc == jk) {  // self-conjugate point: real N(0,1)
    gauss2((uint64_t)p, seed, stream, a, b);
    g[p] = mk(a, 0.0f);
  } else {
    const long long owner = (jkc < jk ? jkc : jk) * cnxp + i;
    gauss2((uint64_t)owner, seed, stream, a, b);
    const float im = jkc < jk ? -b : b;
    g[p] = mk(a * 0.70710678118654752f, im * 0.70710678118654752f);
  }
}

int launch_simulate_points(long long npts, const void *const *cov, void *const *seed, unsigned long long *fails,
                           hipStream_t s) {
  const unsigned blocks = (unsigned)((npts + 255) / 256);
  CovPtrs cp;
  for (int q = 0; q < 6; ++q) cp.p[q] = reinterpret_cast<const cx<float> *>(cov[q]);
  hipLaunchKernelGGL(k_simulate_points, dim3(blocks), dim3(256), 0, s, npts, cp,
                     reinterpret_cast<cx<float> *>(seed[0]), reinterpret_cast<cx<float> *>(seed[1]),
                     reinterpret_cast<cx<float> *>(seed[2]), fails);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_add_real(float *dst, const float *src, long long n, hipStream_t s) {
  const long long blocks = std::min<long long>((n + 255) / 256, 256LL * 64);
  hipLaunchKernelGGL(k_add_real, dim3((unsigned)blocks), dim3(256), 0, s, dst, src, n);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_fill_noise(void *grid, int nxp, int nyp, int nzp, uint64_t seed, uint64_t stream, hipStream_t s) {
  const int cnxp = nxp / 2 + 1;
  const long long n = (long long)cnxp * nyp * nzp;
  hipLaunchKernelGGL(k_fill_noise, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<cx<float> *>(grid), cnxp, nyp, nzp, nxp, seed, stream);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

}  // namespace crava
