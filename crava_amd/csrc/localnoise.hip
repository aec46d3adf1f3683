// Local-noise correction of the posterior: the per-trace matrices of
// Crava::newPosteriorCovPointwise / correctAlphaBetaRho
// (src/crava.cpp:2213-2470), the per-point correction of the posterior
// grids, and the realisation transform of Crava::simulate (:1547-1566).
//
// Every eigen-decomposition of the per-trace work enters through a matrix
// function V diag(f(l)) V^T (inverse above 1e-8, square root of the
// positive part), which does not depend on the eigenvectors' signs; the
// kernels use cyclic Jacobi rotations on the symmetrised matrix, fully
// unrolled for each angle count so that the matrices stay in registers.
// The pseudo-inverse takes a Cholesky route when h - 1e-8 I is positive
// definite (then no eigenvalue is dropped and it is the inverse).
//
// The per-point kernels follow the reference's arithmetic: float
// differences, double sums from the left, one rounding to float (built with
// -ffp-contract=off).  Each thread owns one trace over a range of k and
// batches its loads four k at a time.
#include <cmath>

#include "../../include/crava_hip.h"
#include "eig3.hpp"
#include "localnoise.hpp"

namespace crava {
namespace {

// Cyclic Jacobi on a symmetric N x N matrix: a becomes diagonal (the
// eigenvalues), v holds the eigenvectors as columns.
template <int N>
__host__ __device__ inline void jacobi(double (&a)[N][N], double (&v)[N][N]) {
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c < N; ++c) v[r][c] = (r == c) ? 1.0 : 0.0;
  for (int sweep = 0; sweep < 40; ++sweep) {
    double off = 0.0, tot = 0.0;
#pragma unroll
    for (int r = 0; r < N; ++r)
#pragma unroll
      for (int c = 0; c < N; ++c) {
        const double x = a[r][c] * a[r][c];
        tot += x;
        if (r != c) off += x;
      }
    if (!(off > 1e-33 * tot)) break;  // also ends on NaN
#pragma unroll
    for (int p = 0; p < N - 1; ++p)
#pragma unroll
      for (int q = p + 1; q < N; ++q) {
        const double apq = a[p][q];
        if (apq == 0.0) continue;
        const double theta = (a[q][q] - a[p][p]) / (2.0 * apq);
        double t;
        if (fabs(theta) > 1e150) {
          t = 0.5 / theta;
        } else {
          t = 1.0 / (fabs(theta) + sqrt(theta * theta + 1.0));
          if (theta < 0.0) t = -t;
        }
        const double c = 1.0 / sqrt(t * t + 1.0), s = t * c;
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double akp = a[k][p], akq = a[k][q];
          a[k][p] = c * akp - s * akq;
          a[k][q] = s * akp + c * akq;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double apk = a[p][k], aqk = a[q][k];
          a[p][k] = c * apk - s * aqk;
          a[q][k] = s * apk + c * aqk;
        }
#pragma unroll
        for (int k = 0; k < N; ++k) {
          const double vkp = v[k][p], vkq = v[k][q];
          v[k][p] = c * vkp - s * vkq;
          v[k][q] = s * vkp + c * vkq;
        }
      }
  }
}

enum FnKind { kInverse, kSqrt, kInvSqrt, kIdentity };

// m <- V diag(f(l)) V^T of the symmetrised m; eigenvalues l > thr are kept,
// the others map to 0.
template <int N, FnKind F>
__host__ __device__ inline void matrix_function(double (&m)[N][N], double thr) {
  double a[N][N], v[N][N];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c < N; ++c) a[r][c] = 0.5 * (m[r][c] + m[c][r]);
  jacobi<N>(a, v);
  double f[N];
#pragma unroll
  for (int k = 0; k < N; ++k) {
    const double l = a[k][k];
    double y = 0.0;
    if (l > thr) y = (F == kInverse) ? 1.0 / l : (F == kSqrt) ? sqrt(l) : (F == kIdentity) ? l : 1.0 / sqrt(l);
    f[k] = y;
  }
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c < N; ++c) {
      double x = 0.0;
#pragma unroll
      for (int k = 0; k < N; ++k) x += v[r][k] * f[k] * v[c][k];
      m[r][c] = x;
    }
}

// Whether every eigenvalue of the symmetrised m exceeds thr: m - thr I has a
// Cholesky factor exactly then.  On success m holds, in its lower triangle,
// the Cholesky factor of m itself (the symmetrised m); on failure m is
// clobbered.
template <int N>
__host__ __device__ inline bool cholesky_above(double (&m)[N][N], double thr) {
  double t[N][N];
#pragma unroll
  for (int r = 0; r < N; ++r)
#pragma unroll
    for (int c = 0; c <= r; ++c) {
      const double v = 0.5 * (m[r][c] + m[c][r]);
      m[r][c] = v;
      t[r][c] = (r == c) ? v - thr : v;
    }
  bool ok = true;
#pragma unroll
  for (int j = 0; j < N; ++j) {
    double dt = t[j][j], dm = m[j][j];
#pragma unroll
    for (int k = 0; k < j; ++k) {
      dt -= t[j][k] * t[j][k];
      dm -= m[j][k] * m[j][k];
    }
    ok = ok && (dt > 0.0) && (dm > 0.0);
    dt = sqrt(fmax(dt, 1e-300));
    dm = sqrt(fmax(dm, 1e-300));
    t[j][j] = dt;
    m[j][j] = dm;
#pragma unroll
    for (int r = j + 1; r < N; ++r) {
      double xt = t[r][j], xm = m[r][j];
#pragma unroll
      for (int k = 0; k < j; ++k) {
        xt -= t[r][k] * t[j][k];
        xm -= m[r][k] * m[j][k];
      }
      t[r][j] = xt / dt;
      m[r][j] = xm / dm;
    }
  }
  return ok;
}

// The per-trace matrices for the angle factors d = sqrt(scale):
// sigmanew = sqrt(deltanew) (Sm - Smd)^(-1/2) and the realisation factor
// sqrt(Sm - deltanew) Smd^(-1/2), with deltanew = Sm G^T (G Sm G^T + D E D)^+ G Sm
// and D = diag(d).
template <int NT>
__host__ __device__ inline void lnc_matrices(const LncConst &c, const double (&d)[NT], double *sn, double *tc) {
  double gs[NT][3];  // G Sm
#pragma unroll
  for (int r = 0; r < NT; ++r)
#pragma unroll
    for (int q = 0; q < 3; ++q) gs[r][q] = c.g[r][0] * c.sm[0][q] + c.g[r][1] * c.sm[1][q] + c.g[r][2] * c.sm[2][q];
  double h[NT][NT];  // G Sm G^T + D E D, then its pseudo-inverse
#pragma unroll
  for (int r = 0; r < NT; ++r)
#pragma unroll
    for (int s = 0; s < NT; ++s)
      h[r][s] = (gs[r][0] * c.g[s][0] + gs[r][1] * c.g[s][1] + gs[r][2] * c.g[s][2]) + (d[r] * c.e[r][s]) * d[s];
  double dn[3][3];
  if (cholesky_above<NT>(h, 0.00000001)) {
    // every eigenvalue of h exceeds the threshold: the pseudo-inverse is the
    // inverse, and deltanew = (G Sm)^T h^-1 (G Sm) = Y^T Y, Y = L^-1 (G Sm)
    double y[NT][3];
#pragma unroll
    for (int r = 0; r < NT; ++r)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double x = gs[r][q];
#pragma unroll
        for (int k = 0; k < r; ++k) x -= h[r][k] * y[k][q];
        y[r][q] = x / h[r][r];
      }
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double x = 0.0;
#pragma unroll
        for (int r = 0; r < NT; ++r) x += y[r][p] * y[r][q];
        dn[p][q] = x;
      }
  } else {
    // rebuild h (cholesky_above overwrote it) and take the eigen route
#pragma unroll
    for (int r = 0; r < NT; ++r)
#pragma unroll
      for (int s = 0; s < NT; ++s)
        h[r][s] = (gs[r][0] * c.g[s][0] + gs[r][1] * c.g[s][1] + gs[r][2] * c.g[s][2]) + (d[r] * c.e[r][s]) * d[s];
    matrix_function<NT, kInverse>(h, 0.00000001);
    // deltanew = (Sm G^T) P (G Sm)
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double x = 0.0;
#pragma unroll
        for (int r = 0; r < NT; ++r) {
          double y = 0.0;
#pragma unroll
          for (int s = 0; s < NT; ++s) y += h[r][s] * gs[s][q];
          x += (c.sm[p][0] * c.g[r][0] + c.sm[p][1] * c.g[r][1] + c.sm[p][2] * c.g[r][2]) * y;
        }
        dn[p][q] = x;
      }
  }
  double pc[3][3];
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) pc[p][q] = c.sm[p][q] - dn[p][q];
  matrix_function<3, kSqrt>(dn, 0.0);
#pragma unroll
  for (int p = 0; p < 3; ++p)
#pragma unroll
    for (int q = 0; q < 3; ++q) sn[p * 3 + q] = dn[p][0] * c.sdi[0][q] + dn[p][1] * c.sdi[1][q] + dn[p][2] * c.sdi[2][q];
  if (tc) {
    matrix_function<3, kSqrt>(pc, 0.0);
    // lib_matr_prod order (libs/lib/lib_matr.c:1218-1224)
#pragma unroll
    for (int p = 0; p < 3; ++p)
#pragma unroll
      for (int q = 0; q < 3; ++q) {
        double x = 0.0;
        x = x + pc[p][0] * c.old[0][q];
        x = x + pc[p][1] * c.old[1][q];
        x = x + pc[p][2] * c.old[2][q];
        tc[p * 3 + q] = x;
      }
  }
}

// One thread per trace, d = sqrt(scale / the angle's smallest scale).
template <int NT>
__global__ __launch_bounds__(64) void k_lnc_trace(const LncConst c, const double *__restrict__ scales, int ntr,
                                                  double *__restrict__ sigmanew, double *__restrict__ trace_cov) {
  const int t = blockIdx.x * 64 + threadIdx.x;
  if (t >= ntr) return;
  double d[NT];
#pragma unroll
  for (int a = 0; a < NT; ++a) d[a] = sqrt(scales[(size_t)a * ntr + t] / c.min_scale[a]);
  lnc_matrices<NT>(c, d, sigmanew + (size_t)t * 9, trace_cov ? trace_cov + (size_t)t * 9 : nullptr);
}

constexpr int kLncThreads = 256;
constexpr int kLncKChunk = 16;
constexpr int kLncBatch = 4;

// post_c <- float(mean_c + s_c0 da + s_c1 db + s_c2 dr), d = float(post - mean)
// (src/crava.cpp:2447-2454).
__global__ __launch_bounds__(kLncThreads) void k_lnc_apply(float *__restrict__ p0, float *__restrict__ p1,
                                                           float *__restrict__ p2, const float *__restrict__ m0,
                                                           const float *__restrict__ m1,
                                                           const float *__restrict__ m2,
                                                           const double *__restrict__ sigmanew, int nx, int ntr,
                                                           int nz, long long plane, int rnxp) {
  const int t = blockIdx.x * kLncThreads + threadIdx.x;
  if (t >= ntr) return;
  const int i = t % nx, j = t / nx;
  const int k0 = blockIdx.y * kLncKChunk;
  const int k1 = min(nz, k0 + kLncKChunk);
  double s[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) s[q] = sigmanew[(size_t)t * 9 + q];
  const size_t base = (size_t)j * rnxp + i;
  for (int kb = k0; kb < k1; kb += kLncBatch) {
    float a[kLncBatch][3], m[kLncBatch][3];
#pragma unroll
    for (int u = 0; u < kLncBatch; ++u) {
      const int k = min(kb + u, k1 - 1);  // clamped: the extra lanes reload the last row and store nothing
      const size_t o = base + (size_t)k * plane;
      a[u][0] = p0[o];
      a[u][1] = p1[o];
      a[u][2] = p2[o];
      m[u][0] = m0[o];
      m[u][1] = m1[o];
      m[u][2] = m2[o];
    }
#pragma unroll
    for (int u = 0; u < kLncBatch; ++u) {
      if (kb + u >= k1) break;
      const size_t o = base + (size_t)(kb + u) * plane;
      const float da = a[u][0] - m[u][0], db = a[u][1] - m[u][1], dr = a[u][2] - m[u][2];
      float out[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        double v = (double)m[u][r] + s[r * 3 + 0] * (double)da;
        v = v + s[r * 3 + 1] * (double)db;
        v = v + s[r * 3 + 2] * (double)dr;
        out[r] = (float)v;
      }
      p0[o] = out[0];
      p1[o] = out[1];
      p2[o] = out[2];
    }
  }
}

// seed_c <- float(M_c0 x0 + M_c1 x1 + M_c2 x2) (src/crava.cpp:1559-1561).
__global__ __launch_bounds__(kLncThreads) void k_lnc_realisation(float *__restrict__ g0, float *__restrict__ g1,
                                                                 float *__restrict__ g2,
                                                                 const double *__restrict__ trace_cov, int nx,
                                                                 int ntr, int nz, long long plane, int rnxp) {
  const int t = blockIdx.x * kLncThreads + threadIdx.x;
  if (t >= ntr) return;
  const int i = t % nx, j = t / nx;
  const int k0 = blockIdx.y * kLncKChunk;
  const int k1 = min(nz, k0 + kLncKChunk);
  double s[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) s[q] = trace_cov[(size_t)t * 9 + q];
  const size_t base = (size_t)j * rnxp + i;
  for (int kb = k0; kb < k1; kb += kLncBatch) {
    float x[kLncBatch][3];
#pragma unroll
    for (int u = 0; u < kLncBatch; ++u) {
      const int k = min(kb + u, k1 - 1);
      const size_t o = base + (size_t)k * plane;
      x[u][0] = g0[o];
      x[u][1] = g1[o];
      x[u][2] = g2[o];
    }
#pragma unroll
    for (int u = 0; u < kLncBatch; ++u) {
      if (kb + u >= k1) break;
      const size_t o = base + (size_t)(kb + u) * plane;
      float out[3];
#pragma unroll
      for (int r = 0; r < 3; ++r) {
        double v = s[r * 3 + 0] * (double)x[u][0];
        v = v + s[r * 3 + 1] * (double)x[u][1];
        v = v + s[r * 3 + 2] * (double)x[u][2];
        out[r] = (float)v;
      }
      g0[o] = out[0];
      g1[o] = out[1];
      g2[o] = out[2];
    }
  }
}

}  // namespace

void lnc_inv_sqrt_host(const double in[3][3], double thr, double out[3][3]) {
  double m[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) m[r][c] = in[r][c];
  matrix_function<3, kInvSqrt>(m, thr);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) out[r][c] = m[r][c];
}

namespace {

template <int NT>
void compute_g_nt(const double *E, const double (&sinv)[3][3], const double (&vec)[3][3], const double (&ev)[3],
                  double *G) {
  double ece[NT][NT];
  for (int r = 0; r < NT; ++r)
    for (int c = 0; c < NT; ++c) ece[r][c] = E[r * NT + c];
  matrix_function<NT, kSqrt>(ece, 0.0);
  double lg[NT][3] = {};
  for (int i = 0; i < (NT < 3 ? NT : 3); ++i) lg[i][i] = std::sqrt(ev[i] / (1.0 - ev[i]));
  // G = ((Ece Lg) V^T) Sinv
  for (int r = 0; r < NT; ++r) {
    double el[3], elv[3];
    for (int q = 0; q < 3; ++q) {
      el[q] = 0.0;
      for (int k = 0; k < NT; ++k) el[q] += ece[r][k] * lg[k][q];
    }
    for (int q = 0; q < 3; ++q) elv[q] = el[0] * vec[q][0] + el[1] * vec[q][1] + el[2] * vec[q][2];
    for (int q = 0; q < 3; ++q) G[r * 3 + q] = elv[0] * sinv[0][q] + elv[1] * sinv[1][q] + elv[2] * sinv[2][q];
  }
}

}  // namespace

int lnc_compute_g(int ntheta, const double *E, const double *Sm, const double *Smd, double *G) {
  double sd[3][3], sinv[3][3], a[3][3], v[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      sd[r][c] = Sm[r * 3 + c] - Smd[r * 3 + c];
      sinv[r][c] = Sm[r * 3 + c];
    }
  matrix_function<3, kIdentity>(sd, 0.0);       // negative eigenvalues dropped
  matrix_function<3, kInvSqrt>(sinv, 0.0000001);
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      double x = 0.0;
      for (int k = 0; k < 3; ++k)
        for (int l = 0; l < 3; ++l) x += sinv[r][k] * sd[k][l] * sinv[l][c];
      a[r][c] = x;
    }
  // LAPACK DGEEV's eigenvectors (signs included) where its restatement
  // handles the matrix; otherwise Jacobi with the largest-component rule
  double wr[3];
  const bool lapack_like = dgeev3(a, wr, v);
  if (lapack_like) {
    for (int k = 0; k < 3; ++k) a[k][k] = wr[k];
  } else {
    for (int r = 0; r < 3; ++r)
      for (int c = 0; c < r; ++c) a[r][c] = a[c][r] = 0.5 * (a[r][c] + a[c][r]);
    jacobi<3>(a, v);
  }
  // NRLib::Sort3x3: descending, the first maximum wins a tie
  double ev[3], vec[3][3];
  int used[3] = {0, 0, 0};
  for (int j = 0; j < 3; ++j) {
    int best = -1;
    double mx = -100000.0;
    for (int i = 0; i < 3; ++i)
      if (!used[i] && a[i][i] > mx) {
        mx = a[i][i];
        best = i;
      }
    if (best < 0) return CRAVA_HIP_E_INVALID;
    used[best] = 1;
    ev[j] = mx;
    // sign: DGEEV's, or the component of largest magnitude positive
    int im = 0;
    for (int k = 1; k < 3; ++k)
      if (std::fabs(v[k][best]) > std::fabs(v[im][best])) im = k;
    const double sg = (!lapack_like && v[im][best] < 0.0) ? -1.0 : 1.0;
    for (int k = 0; k < 3; ++k) vec[k][j] = sg * v[k][best];
  }
  switch (ntheta) {
#define CRAVA_LNC_G(N) \
  case N:              \
    compute_g_nt<N>(E, sinv, vec, ev, G); \
    break;
    CRAVA_LNC_G(1)
    CRAVA_LNC_G(2)
    CRAVA_LNC_G(3)
    CRAVA_LNC_G(4)
    CRAVA_LNC_G(5)
    CRAVA_LNC_G(6)
    CRAVA_LNC_G(7)
    CRAVA_LNC_G(8)
#undef CRAVA_LNC_G
    default:
      return CRAVA_HIP_E_INVALID;
  }
  return CRAVA_HIP_OK;
}

void lnc_const_host(int ntheta, const double *G, const double *E, const double *Sm, const double *Smd, LncConst &c) {
  c = LncConst{};
  for (int r = 0; r < ntheta; ++r) {
    for (int q = 0; q < 3; ++q) c.g[r][q] = G[r * 3 + q];
    for (int q = 0; q < ntheta; ++q) c.e[r][q] = E[r * ntheta + q];
  }
  double sd[3][3], smd[3][3];
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q) {
      c.sm[r][q] = Sm[r * 3 + q];
      smd[r][q] = Smd[r * 3 + q];
      sd[r][q] = c.sm[r][q] - smd[r][q];
    }
  lnc_inv_sqrt_host(sd, 0.0000001, c.sdi);  // src/crava.cpp:2294-2307
  lnc_inv_sqrt_host(smd, 0.0, c.old);       // :2376-2385
  for (int a = 0; a < kLncMaxTheta; ++a) c.min_scale[a] = 1.0;
}

int lnc_sigmanew_host(const LncConst &c, int ntheta, const double *scale, double *sigmanew) {
  switch (ntheta) {
#define CRAVA_LNC_H(N)                                         \
  case N: {                                                    \
    double d[N];                                               \
    for (int a = 0; a < N; ++a) d[a] = std::sqrt(scale[a]);    \
    lnc_matrices<N>(c, d, sigmanew, nullptr);                  \
    break;                                                     \
  }
    CRAVA_LNC_H(1)
    CRAVA_LNC_H(2)
    CRAVA_LNC_H(3)
    CRAVA_LNC_H(4)
    CRAVA_LNC_H(5)
    CRAVA_LNC_H(6)
    CRAVA_LNC_H(7)
    CRAVA_LNC_H(8)
#undef CRAVA_LNC_H
    default:
      return CRAVA_HIP_E_INVALID;
  }
  return CRAVA_HIP_OK;
}

int launch_lnc_trace(const LncConst &c, int ntheta, const double *scales, int nx, int ny, double *sigmanew,
                     double *trace_cov, hipStream_t s) {
  const int ntr = nx * ny;
  const dim3 grid((unsigned)((ntr + 63) / 64));
  switch (ntheta) {
#define CRAVA_LNC_CASE(N)                                                                           \
  case N:                                                                                           \
    hipLaunchKernelGGL(k_lnc_trace<N>, grid, dim3(64), 0, s, c, scales, ntr, sigmanew, trace_cov); \
    break;
    CRAVA_LNC_CASE(1)
    CRAVA_LNC_CASE(2)
    CRAVA_LNC_CASE(3)
    CRAVA_LNC_CASE(4)
    CRAVA_LNC_CASE(5)
    CRAVA_LNC_CASE(6)
    CRAVA_LNC_CASE(7)
    CRAVA_LNC_CASE(8)
#undef CRAVA_LNC_CASE
    default:
      return CRAVA_HIP_E_INVALID;
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_lnc_apply(float *const post[3], const float *const mean2[3], const double *sigmanew, int nx, int ny,
                     int nz, int nyp, int rnxp, hipStream_t s) {
  const int ntr = nx * ny;
  const dim3 grid((unsigned)((ntr + kLncThreads - 1) / kLncThreads), (unsigned)((nz + kLncKChunk - 1) / kLncKChunk));
  hipLaunchKernelGGL(k_lnc_apply, grid, dim3(kLncThreads), 0, s, post[0], post[1], post[2], mean2[0], mean2[1],
                     mean2[2], sigmanew, nx, ntr, nz, (long long)nyp * rnxp, rnxp);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_lnc_realisation(float *const seed[3], const double *trace_cov, int nx, int ny, int nz, int nyp, int rnxp,
                           hipStream_t s) {
  const int ntr = nx * ny;
  const dim3 grid((unsigned)((ntr + kLncThreads - 1) / kLncThreads), (unsigned)((nz + kLncKChunk - 1) / kLncKChunk));
  hipLaunchKernelGGL(k_lnc_realisation, grid, dim3(kLncThreads), 0, s, seed[0], seed[1], seed[2], trace_cov, nx, ntr,
                     nz, (long long)nyp * rnxp, rnxp);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

}  // namespace crava
