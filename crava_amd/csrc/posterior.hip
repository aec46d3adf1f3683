// Batched per-Fourier-point posterior for gfx950.
//
// One lane owns one half-complex Fourier point; a workgroup covers 256
// consecutive points of one k-plane, so every per-k quantity (K, the noise
// matrices, the frequency gate) is wave-uniform and is read through the
// scalar cache.  Per point (src/crava.cpp:1087-1150):
//
//   S  = cleaned prior 3x3 (src/seismicparametersholder.cpp:373-426)
//   M  = K S K^H + E,  K = diag(kWD) A            (:1105-1107)
//   L  = chol(M), failure -> prior pass-through    (:1109-1130, lib_matr.c:78-167)
//   Y  = L^-1 (K S),  z = L^-1 (d - K mu)
//   Sigma_post = S - Y^H Y          (= S - (KS)^H M^-1 (KS), :1112-1115)
//   mu_post    = mu + Y^H z         (= mu + X^H (d - K mu),   :1117-1121)
//   residual   = d - K mu_post                               (:1122-1123)
//
// The Gram form Y^H Y replaces the reference's forward+back substitution and
// product; it is the same matrix algebraically and Hermitian by construction.
// FP32 and FP64 run it.  The mixed form runs a square-root information form
// in FP32 (solve_sqrt) and the Gram form in FP64 for the waves it queues.
#include "posterior.hpp"
#include "../../include/crava_hip.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

namespace crava {

template <typename G>
struct PostArgs {
  cx<G> *mu[3];
  cx<G> *d[kMaxTheta];
  cx<G> *cov[6];
  const cx<G> *ec;
  long long ec_kstride;  // ec of point (k, ip) is ec[k ec_kstride + ip]: the plane size, or 0 for one plane
  // Separable prior Sigma0 (x) corr (SeismicParametersHolder::
  // initializeCorrelations, src/seismicparametersholder.cpp:197-221, fills
  // the six grids with the one correlation times Sigma0_ab): the prior is
  // read from this one grid, S = |Re corr| Sigma0, which is the cleaned
  // prior of the six grids in exact arithmetic; cov[] are then outputs only.
  const cx<G> *corr;
  const PerK<float> *tabf;
  const PerK<double> *tabd;
  unsigned long long *counters;  // [0] failed Cholesky, [1] points solved in FP64
  int2 *defer_list;              // mixed form: queued waves {k, first point}
  unsigned *defer_count;
  long long plane;
  int nt;
  float Af[kMaxTheta][3];
  double Ad[kMaxTheta][3];
  float s0f[6];
  double s0d[6];
  int s0nz[6];
  float thr;
  const int *kmap;  // k-plane of grid row blockIdx.y (with an offset), or NULL
};

template <typename T>
__device__ __forceinline__ const PerK<T> *pick_tab(const PerK<float> *f, const PerK<double> *d);
template <>
__device__ __forceinline__ const PerK<float> *pick_tab<float>(const PerK<float> *f, const PerK<double> *) { return f; }
template <>
__device__ __forceinline__ const PerK<double> *pick_tab<double>(const PerK<float> *, const PerK<double> *d) { return d; }

// Solve one in-band point.  Returns false when the Cholesky factorisation
// fails (the reference then keeps the prior, src/crava.cpp:1109).
template <typename T, int NT>
__device__ __forceinline__ bool solve_point(const PerK<T> *__restrict__ tb, const T (&A)[kMaxTheta][3], const T (&S)[6],
                                            T lam, cx<T> (&mu)[3], cx<T> (&d)[NT], cx<T> (&post)[6]) {
  const T Sm[3][3] = {{S[0], S[3], S[4]}, {S[3], S[1], S[5]}, {S[4], S[5], S[2]}};
  T AS[NT][3];
#pragma unroll
  for (int l = 0; l < NT; ++l)
#pragma unroll
    for (int b = 0; b < 3; ++b) AS[l][b] = A[l][0] * Sm[0][b] + A[l][1] * Sm[1][b] + A[l][2] * Sm[2][b];

  cx<T> L[NT][NT];
#pragma unroll
  for (int l = 0; l < NT; ++l)
#pragma unroll
    for (int m = 0; m <= l; ++m) {
      const int q = l * (l + 1) / 2 + m;
      const T B = AS[l][0] * A[m][0] + AS[l][1] * A[m][1] + AS[l][2] * A[m][2];
      const cx<T> pv = tb->P[q], ev = tb->Ec[q];
      L[l][m] = mk(pv.x * B + lam * ev.x, pv.y * B + lam * ev.y);
    }
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    L[l][l].x += tb->Ed[l];
    L[l][l].y = (T)0;
  }

  // Cholesky, lower.  Failure rule of lib_matrCholCpx: M00 <= 0, or a
  // diagonal / pivot <= 1e-20 after normalising by M00.
  const T f = L[0][0].x;
  if (f <= (T)0) return false;
  const T tol = (T)1e-20 * f;
  T rinv[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    if (L[i][i].x <= tol) return false;
#pragma unroll
    for (int j = 0; j < i; ++j) {
      cx<T> s = L[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) s = sub(s, mulc(L[i][k], L[j][k]));
      L[i][j] = scl(s, rinv[j]);
    }
    T dd = L[i][i].x;
#pragma unroll
    for (int k = 0; k < i; ++k) dd -= norm2(L[i][k]);
    if (dd <= tol) return false;
    const T li = sqrt(dd);
    L[i][i].x = li;
    rinv[i] = (T)1 / li;
  }

  // Y = L^-1 (K S), z = L^-1 (d - K mu)
  cx<T> Y[NT][3], z[NT];
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    const cx<T> w = tb->kwd[l];
    cx<T> km = mk(A[l][0] * mu[0].x + A[l][1] * mu[1].x + A[l][2] * mu[2].x,
                  A[l][0] * mu[0].y + A[l][1] * mu[1].y + A[l][2] * mu[2].y);
    cx<T> r = sub(d[l], mul(w, km));
#pragma unroll
    for (int b = 0; b < 3; ++b) Y[l][b] = scl(w, AS[l][b]);
#pragma unroll
    for (int m = 0; m < l; ++m) {
#pragma unroll
      for (int b = 0; b < 3; ++b) Y[l][b] = sub(Y[l][b], mul(L[l][m], Y[m][b]));
      r = sub(r, mul(L[l][m], z[m]));
    }
#pragma unroll
    for (int b = 0; b < 3; ++b) Y[l][b] = scl(Y[l][b], rinv[l]);
    z[l] = scl(r, rinv[l]);
  }

  // Sigma_post = S - Y^H Y; mu += Y^H z
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    cx<T> acc = mk((T)0, (T)0);
#pragma unroll
    for (int l = 0; l < NT; ++l) acc = add(acc, cmul(Y[l][ia[q]], Y[l][ib[q]]));
    post[q] = mk(S[q] - acc.x, -acc.y);
  }
#pragma unroll
  for (int a = 0; a < 3; ++a) {
    cx<T> acc = mk((T)0, (T)0);
#pragma unroll
    for (int l = 0; l < NT; ++l) acc = add(acc, cmul(Y[l][a], z[l]));
    mu[a] = add(mu[a], acc);
  }
  // residual = d - K mu_post
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    cx<T> km = mk(A[l][0] * mu[0].x + A[l][1] * mu[1].x + A[l][2] * mu[2].x,
                  A[l][0] * mu[0].y + A[l][1] * mu[1].y + A[l][2] * mu[2].y);
    d[l] = sub(d[l], mul(tb->kwd[l], km));
  }
  return true;
}

template <typename D, typename S_>
__device__ __forceinline__ cx<D> cvt(cx<S_> a) { return mk((D)a.x, (D)a.y); }

// FP32 square-root information form of the same posterior (mixed form, the
// points at or above the noise/signal threshold).  With E = Le Le^H,
// S = Ls Ls^T, C = Le^-1 K Ls and y = Le^-1 (d - K mu), the posterior is the
// regularised least-squares problem min |C x - y|^2 + |x|^2:
//   [C; I] = Q R (modified Gram-Schmidt, applied to [y; 0] as well),
//   mu_post    = mu + Ls R^-1 (Q^H [y; 0])
//   Sigma_post = (Ls R^-1)(Ls R^-1)^H
//   residual   = Le (y - C x), the first ntheta entries of the Gram-Schmidt
//                remainder of [y; 0]
// which equals S - S K^H M^-1 K S etc. algebraically.  Neither M = K S K^H + E
// nor S - R is formed, so nothing cancels at low noise: the rounding errors
// grow with sqrt(cond M) instead of cond M.  Returns false when S or E is
// not positive definite in float (the wave then takes the FP64 form, which
// applies the reference's failure rule).
template <int NT>
__device__ __forceinline__ bool solve_sqrt(const PerK<float> *__restrict__ tb, const float (&A)[kMaxTheta][3],
                                           const float (&S)[6], float lam, cx<float> (&mu)[3], cx<float> (&d)[NT],
                                           cx<float> (&post)[6]) {
  // Ls = chol(S); a conditional variance below 1e-5 of the marginal counts as
  // singular
  if (!(S[0] > 0.0f)) return false;
  const float l00 = sqrtf(S[0]), i00 = 1.0f / l00;
  const float l10 = S[3] * i00, l20 = S[4] * i00;
  const float d11 = S[1] - l10 * l10;
  if (!(d11 > 1e-5f * S[1])) return false;
  const float l11 = sqrtf(d11), i11 = 1.0f / l11;
  const float l21 = (S[5] - l20 * l10) * i11;
  const float d22 = S[2] - l20 * l20 - l21 * l21;
  if (!(d22 > 1e-5f * S[2])) return false;
  const float l22 = sqrtf(d22);

  // Le = chol(E), E = lam Ec + diag(Ed), lower; rd = 1 / diag
  cx<float> L[NT][NT];
  float rd[NT];
#pragma unroll
  for (int i = 0; i < NT; ++i) {
#pragma unroll
    for (int j = 0; j < i; ++j) {
      const cx<float> e = tb->Ec[i * (i + 1) / 2 + j];
      cx<float> s = mk(lam * e.x, lam * e.y);
#pragma unroll
      for (int k = 0; k < j; ++k) s = sub(s, mulc(L[i][k], L[j][k]));
      L[i][j] = scl(s, rd[j]);
    }
    const float eii = lam * tb->Ec[i * (i + 1) / 2 + i].x + tb->Ed[i];
    float dd = eii;
#pragma unroll
    for (int k = 0; k < i; ++k) dd -= norm2(L[i][k]);
    if (!(dd > 1e-12f * eii)) return false;
    const float li = sqrtf(dd);
    L[i][i] = mk(li, 0.0f);
    rd[i] = 1.0f / li;
  }

  // C = Le^-1 (kWD_l (A_l Ls)), y = Le^-1 (d - K mu), forward substitution
  cx<float> C[NT][3], y[NT];
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    const cx<float> w = tb->kwd[l];
    const float a0 = A[l][0] * l00 + A[l][1] * l10 + A[l][2] * l20;
    const float a1 = A[l][1] * l11 + A[l][2] * l21;
    const float a2 = A[l][2] * l22;
    cx<float> c0 = scl(w, a0), c1 = scl(w, a1), c2 = scl(w, a2);
    const cx<float> am = mk(A[l][0] * mu[0].x + A[l][1] * mu[1].x + A[l][2] * mu[2].x,
                            A[l][0] * mu[0].y + A[l][1] * mu[1].y + A[l][2] * mu[2].y);
    cx<float> r = sub(d[l], mul(w, am));
#pragma unroll
    for (int m = 0; m < l; ++m) {
      c0 = sub(c0, mul(L[l][m], C[m][0]));
      c1 = sub(c1, mul(L[l][m], C[m][1]));
      c2 = sub(c2, mul(L[l][m], C[m][2]));
      r = sub(r, mul(L[l][m], y[m]));
    }
    C[l][0] = scl(c0, rd[l]);
    C[l][1] = scl(c1, rd[l]);
    C[l][2] = scl(c2, rd[l]);
    y[l] = scl(r, rd[l]);
  }

  // modified Gram-Schmidt on [C; I] with [y; 0] alongside; the identity block
  // is tracked explicitly: u* are the lower three entries of columns 1, 2 and
  // of the right-hand side as the columns are orthogonalised
  // column 0
  float n0 = 1.0f;
#pragma unroll
  for (int l = 0; l < NT; ++l) n0 += norm2(C[l][0]);
  const float r00 = sqrtf(n0), ir0 = 1.0f / r00;
  cx<float> r01 = mk(0.0f, 0.0f), r02 = mk(0.0f, 0.0f), t0 = mk(0.0f, 0.0f);
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    C[l][0] = scl(C[l][0], ir0);                // q0 (upper part)
    r01 = add(r01, cmul(C[l][0], C[l][1]));
    r02 = add(r02, cmul(C[l][0], C[l][2]));
    t0 = add(t0, cmul(C[l][0], y[l]));
  }
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    C[l][1] = sub(C[l][1], mul(r01, C[l][0]));
    C[l][2] = sub(C[l][2], mul(r02, C[l][0]));
    y[l] = sub(y[l], mul(t0, C[l][0]));
  }
  // lower block: q0 = (ir0, 0, 0); column 1 -> (-r01 ir0, 1, 0),
  // column 2 -> (-r02 ir0, 0, 1), rhs -> (-t0 ir0, 0, 0)
  const cx<float> u10 = scl(r01, -ir0), u20 = scl(r02, -ir0), uy0 = scl(t0, -ir0);
  // column 1
  float n1 = 1.0f + norm2(u10);
#pragma unroll
  for (int l = 0; l < NT; ++l) n1 += norm2(C[l][1]);
  const float r11 = sqrtf(n1), ir1 = 1.0f / r11;
  const cx<float> q10 = scl(u10, ir1);          // q1 lower block = (q10, ir1, 0)
  cx<float> r12 = cmul(q10, u20), t1 = cmul(q10, uy0);
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    C[l][1] = scl(C[l][1], ir1);
    r12 = add(r12, cmul(C[l][1], C[l][2]));
    t1 = add(t1, cmul(C[l][1], y[l]));
  }
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    C[l][2] = sub(C[l][2], mul(r12, C[l][1]));
    y[l] = sub(y[l], mul(t1, C[l][1]));
  }
  // column 2 -> (u20 - r12 q10, -r12 ir1, 1), rhs -> (uy0 - t1 q10, -t1 ir1, 0)
  const cx<float> v20 = sub(u20, mul(r12, q10)), v21 = scl(r12, -ir1);
  const cx<float> vy0 = sub(uy0, mul(t1, q10)), vy1 = scl(t1, -ir1);
  // column 2
  float n2 = 1.0f + norm2(v20) + norm2(v21);
#pragma unroll
  for (int l = 0; l < NT; ++l) n2 += norm2(C[l][2]);
  const float r22 = sqrtf(n2), ir2 = 1.0f / r22;
  cx<float> t2 = add(cmul(v20, vy0), cmul(v21, vy1));
#pragma unroll
  for (int l = 0; l < NT; ++l) t2 = add(t2, cmul(C[l][2], y[l]));
  t2 = scl(t2, ir2);
#pragma unroll
  for (int l = 0; l < NT; ++l) y[l] = sub(y[l], mul(t2, scl(C[l][2], ir2)));

  // x = R^-1 t; mu += Ls x
  const cx<float> x2 = scl(t2, ir2);
  const cx<float> x1 = scl(sub(t1, mul(r12, x2)), ir1);
  const cx<float> x0 = scl(sub(sub(t0, mul(r01, x1)), mul(r02, x2)), ir0);
  mu[0] = add(mu[0], scl(x0, l00));
  mu[1] = add(mu[1], add(scl(x0, l10), scl(x1, l11)));
  mu[2] = add(mu[2], add(add(scl(x0, l20), scl(x1, l21)), scl(x2, l22)));

  // Sigma_post = T T^H, T = Ls R^-1 (R^-1 upper: g)
  const cx<float> g01 = scl(r01, -ir0 * ir1), g12 = scl(r12, -ir1 * ir2);
  const cx<float> g02 = scl(sub(scl(mul(r01, r12), ir1), r02), ir0 * ir2);
  const cx<float> T00 = mk(l00 * ir0, 0.0f), T01 = scl(g01, l00), T02 = scl(g02, l00);
  const cx<float> T10 = mk(l10 * ir0, 0.0f), T11 = add(scl(g01, l10), mk(l11 * ir1, 0.0f));
  const cx<float> T12 = add(scl(g02, l10), scl(g12, l11));
  const cx<float> T20 = mk(l20 * ir0, 0.0f), T21 = add(scl(g01, l20), mk(l21 * ir1, 0.0f));
  const cx<float> T22 = add(add(scl(g02, l20), scl(g12, l21)), mk(l22 * ir2, 0.0f));
  // Sigma_ab = sum_c T_ac conj(T_bc)
  post[0] = mk(norm2(T00) + norm2(T01) + norm2(T02), 0.0f);
  post[1] = mk(norm2(T10) + norm2(T11) + norm2(T12), 0.0f);
  post[2] = mk(norm2(T20) + norm2(T21) + norm2(T22), 0.0f);
  post[3] = add(add(mulc(T00, T10), mulc(T01, T11)), mulc(T02, T12));
  post[4] = add(add(mulc(T00, T20), mulc(T01, T21)), mulc(T02, T22));
  post[5] = add(add(mulc(T10, T20), mulc(T11, T21)), mulc(T12, T22));

  // residual = Le (whitened remainder)
#pragma unroll
  for (int l = NT - 1; l >= 0; --l) {
    cx<float> r = scl(y[l], L[l][l].x);
#pragma unroll
    for (int m = 0; m < l; ++m) r = add(r, mul(L[l][m], y[m]));
    d[l] = r;
  }
  return true;
}

// Cleaned prior S of point p (src/seismicparametersholder.cpp:373-426), or
// |Re corr| Sigma0 for a separable prior.
template <typename G>
__device__ __forceinline__ void load_prior(const PostArgs<G> &a, long long p, G (&S)[6]) {
  const G *s0 = nullptr;
  if constexpr (sizeof(G) == 4) s0 = a.s0f; else s0 = a.s0d;
  if (a.corr) {
    const G c = fabs(a.corr[p].x);
#pragma unroll
    for (int q = 0; q < 6; ++q) S[q] = c * s0[q];
    return;
  }
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    G v = a.cov[q][p].x;
    if (a.s0nz[q]) v = v / s0[q];
    S[q] = fabs(v) * s0[q];
  }
}

template <typename G>
__device__ __forceinline__ void count(const PostArgs<G> &a, bool fail, bool wide) {
  if (a.counters) {
    const unsigned long long fails = __ballot(fail);
    const unsigned long long n64 = __ballot(wide);
    if ((threadIdx.x & 63) == 0) {
      if (fails) atomicAdd(&a.counters[0], (unsigned long long)__popcll(fails));
      if (n64) atomicAdd(&a.counters[1], (unsigned long long)__popcll(n64));
    }
  }
}

// FP32 (MODE 0) or FP64 (G = double) point in the reference's algorithm.
template <typename G, int NT>
__device__ __forceinline__ void post_point(const PostArgs<G> &a, int k, long long ip) {
  const bool live = ip < a.plane;
  const long long p = (long long)k * a.plane + (live ? ip : 0);
  const PerK<G> *tb = pick_tab<G>(a.tabf, a.tabd) + k;

  G S[6];
  if (live) load_prior(a, p, S);
  if (!tb->inv) {  // out of band: prior passes through, cleaned
    if (live) {
#pragma unroll
      for (int q = 0; q < 6; ++q) a.cov[q][p] = mk(S[q], (G)0);
    }
    return;
  }
  if (!live) return;

  G A[kMaxTheta][3];
#pragma unroll
  for (int l = 0; l < NT; ++l)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      if constexpr (sizeof(G) == 4) A[l][b] = a.Af[l][b]; else A[l][b] = a.Ad[l][b];
    }
  const G lam = fabs(a.ec[(long long)k * a.ec_kstride + ip].x);
  cx<G> mu[3], d[NT], post[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) mu[c] = a.mu[c][p];
#pragma unroll
  for (int l = 0; l < NT; ++l) d[l] = a.d[l][p];

  const bool ok = solve_point<G, NT>(tb, A, S, lam, mu, d, post);
  if (ok) {
#pragma unroll
    for (int c = 0; c < 3; ++c) a.mu[c][p] = mu[c];
#pragma unroll
    for (int q = 0; q < 6; ++q) a.cov[q][p] = post[q];
#pragma unroll
    for (int l = 0; l < NT; ++l) a.d[l][p] = d[l];
  } else {
#pragma unroll
    for (int q = 0; q < 6; ++q) a.cov[q][p] = mk(S[q], (G)0);
  }
  count(a, !ok, false);
}

// FP64 solve of one in-band point of float grids (the mixed form's queued
// waves): float inputs, FP64 reference-form solve, one rounding at the end.
template <int NT>
__device__ __forceinline__ void post_point_wide(const PostArgs<float> &a, int k, long long ip) {
  const bool live = ip < a.plane;
  const long long p = (long long)k * a.plane + (live ? ip : 0);
  bool ok = true;
  if (live) {
    float Sf[6];
    load_prior(a, p, Sf);
    double S[6], A[kMaxTheta][3];
#pragma unroll
    for (int q = 0; q < 6; ++q) S[q] = (double)Sf[q];
#pragma unroll
    for (int l = 0; l < NT; ++l)
#pragma unroll
      for (int b = 0; b < 3; ++b) A[l][b] = a.Ad[l][b];
    const double lam = (double)fabsf(a.ec[(long long)k * a.ec_kstride + ip].x);
    cx<double> mu[3], d[NT], post[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) mu[c] = cvt<double>(a.mu[c][p]);
#pragma unroll
    for (int l = 0; l < NT; ++l) d[l] = cvt<double>(a.d[l][p]);
    ok = solve_point<double, NT>(a.tabd + k, A, S, lam, mu, d, post);
    if (ok) {
#pragma unroll
      for (int c = 0; c < 3; ++c) a.mu[c][p] = cvt<float>(mu[c]);
#pragma unroll
      for (int q = 0; q < 6; ++q) a.cov[q][p] = cvt<float>(post[q]);
#pragma unroll
      for (int l = 0; l < NT; ++l) a.d[l][p] = cvt<float>(d[l]);
    } else {
#pragma unroll
      for (int q = 0; q < 6; ++q) a.cov[q][p] = mk(Sf[q], 0.0f);
    }
  }
  count(a, live && !ok, live && ok);
}

// Mixed form, FP32 kernel: point ip of k-plane k, in the 64-point wave that
// starts at point w0.  A wave in which any point lies below the noise/signal
// threshold tr(E)/tr(K S K^H) < thr, or whose FP32 factorisations fail, is
// queued whole for the FP64 kernel and writes nothing here.
template <int NT>
__device__ __forceinline__ void post_point_mixed(const PostArgs<float> &a, int k, long long ip, long long w0) {
  const bool live = ip < a.plane;
  const long long p = (long long)k * a.plane + (live ? ip : 0);
  const PerK<float> *tb = a.tabf + k;
  float S[6];
  if (live) load_prior(a, p, S);
  if (!tb->inv) {
    if (live) {
#pragma unroll
      for (int q = 0; q < 6; ++q) a.cov[q][p] = mk(S[q], 0.0f);
    }
    return;
  }
  float A[kMaxTheta][3];
#pragma unroll
  for (int l = 0; l < NT; ++l)
#pragma unroll
    for (int b = 0; b < 3; ++b) A[l][b] = a.Af[l][b];
  bool need64 = false;
  float lam = 0.0f;
  if (live) {
    lam = fabsf(a.ec[(long long)k * a.ec_kstride + ip].x);
    // tr(K S K^H) = sum_l |kWD_l|^2 (A S A^T)_ll
    float sig = 0.0f;
#pragma unroll
    for (int l = 0; l < NT; ++l) {
      const float b0 = A[l][0] * S[0] + A[l][1] * S[3] + A[l][2] * S[4];
      const float b1 = A[l][0] * S[3] + A[l][1] * S[1] + A[l][2] * S[5];
      const float b2 = A[l][0] * S[4] + A[l][1] * S[5] + A[l][2] * S[2];
      sig += tb->P[l * (l + 1) / 2 + l].x * (b0 * A[l][0] + b1 * A[l][1] + b2 * A[l][2]);
    }
    need64 = lam * tb->trEc + tb->trEd < a.thr * sig;
  }
  cx<float> mu[3], d[NT], post[6];
  bool ok = true;
  if (__ballot(need64) == 0ull && live) {
#pragma unroll
    for (int c = 0; c < 3; ++c) mu[c] = a.mu[c][p];
#pragma unroll
    for (int l = 0; l < NT; ++l) d[l] = a.d[l][p];
    ok = solve_sqrt<NT>(tb, A, S, lam, mu, d, post);
  }
  if (__ballot(need64 || !ok) != 0ull) {
    if ((threadIdx.x & 63) == 0) {
      const unsigned slot = atomicAdd(a.defer_count, 1u);
      a.defer_list[slot] = make_int2(k, (int)w0);
    }
    return;
  }
  if (live) {
#pragma unroll
    for (int c = 0; c < 3; ++c) a.mu[c][p] = mu[c];
#pragma unroll
    for (int q = 0; q < 6; ++q) a.cov[q][p] = post[q];
#pragma unroll
    for (int l = 0; l < NT; ++l) a.d[l][p] = d[l];
  }
}

template <typename G, int NT, int MODE>
__device__ __forceinline__ void post_any(const PostArgs<G> &a, int k, long long ip, long long w0) {
  if constexpr (MODE == 1) post_point_mixed<NT>(a, k, ip, w0);
  else post_point<G, NT>(a, k, ip);
}

template <typename G, int NT, int MODE>
__global__ void __launch_bounds__(256) k_posterior(PostArgs<G> a, int koff) {
  // consecutive workgroups go to different XCDs, which spreads the in-band
  // k-planes (a contiguous range of k) over all of them
  const long long ip = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const int k = a.kmap ? a.kmap[koff + (int)blockIdx.y] : (int)blockIdx.y;
  post_any<G, NT, MODE>(a, k, ip, ip - (threadIdx.x & 63));
}

// Out-of-band k-planes (kmap[koff + blockIdx.y]): the cleaned prior passes
// through, as post_point / post_point_mixed write it there.  A kernel of its
// own so these planes, a plain stream of writes, run at the occupancy of a
// few registers rather than at that of the in-band solve.
template <typename G>
__global__ void __launch_bounds__(256) k_posterior_prior(PostArgs<G> a, int koff) {
  const long long ip = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ip >= a.plane) return;
  const int k = a.kmap[koff + (int)blockIdx.y];
  const long long p = (long long)k * a.plane + ip;
  G S[6];
  load_prior(a, p, S);
#pragma unroll
  for (int q = 0; q < 6; ++q) a.cov[q][p] = mk(S[q], (G)0);
}

// Mixed form, FP64 kernel: every queued wave; waves stride over the queue.
template <int NT>
__global__ void __launch_bounds__(256) k_posterior_queued(PostArgs<float> a) {
  const unsigned n = *a.defer_count;
  const unsigned nw = (gridDim.x * blockDim.x) >> 6;
  for (unsigned w = (blockIdx.x * blockDim.x + threadIdx.x) >> 6; w < n; w += nw) {
    const int2 e = a.defer_list[w];
    post_point_wide<NT>(a, e.x, (long long)e.y + (threadIdx.x & 63));
  }
}

__global__ void k_zero_queue_counts(const PostArgs<float> *__restrict__ vols, int nvol) {
  const int v = blockIdx.x * blockDim.x + threadIdx.x;
  if (v < nvol) *vols[v].defer_count = 0u;
}

__global__ void k_zero_queue_count(unsigned *c) { *c = 0u; }

// The queued waves of every volume of a batch (each volume's own queue).
__global__ void __launch_bounds__(256) k_posterior_queued_batch(const PostArgs<float> *__restrict__ vols, int nvol) {
  const unsigned nw = (gridDim.x * blockDim.x) >> 6;
  const unsigned w0 = (blockIdx.x * blockDim.x + threadIdx.x) >> 6;
  for (int v = 0; v < nvol; ++v) {
    const PostArgs<float> &a = vols[v];
    const unsigned n = *a.defer_count;
    for (unsigned w = w0; w < n; w += nw) {
      const int2 e = a.defer_list[w];
      const long long ip = (long long)e.y + (threadIdx.x & 63);
      switch (a.nt) {
        case 1: post_point_wide<1>(a, e.x, ip); break;
        case 2: post_point_wide<2>(a, e.x, ip); break;
        case 3: post_point_wide<3>(a, e.x, ip); break;
        case 4: post_point_wide<4>(a, e.x, ip); break;
        case 5: post_point_wide<5>(a, e.x, ip); break;
        case 6: post_point_wide<6>(a, e.x, ip); break;
        case 7: post_point_wide<7>(a, e.x, ip); break;
        default: post_point_wide<8>(a, e.x, ip); break;
      }
    }
  }
}

// Many volumes in one persistent launch.  Tile t = {volume, k, first point,
// ntheta} covers blockDim.x consecutive points of one k-plane of one
// volume; workgroups stride over the tile list.  Everything a tile reads
// from `vols` and `tiles` is wave-uniform.  This form holds the registers of
// the largest ntheta for every tile; k_posterior_class below is the same
// loop over the tiles of one ntheta.  Tiles start at multiples of 64 points,
// so the waves are those of the single-volume kernel.
// Out-of-band tiles [t0, t1) of a batch: the cleaned prior passes through.
template <typename G>
__global__ void __launch_bounds__(256) k_posterior_prior_tiles(const PostArgs<G> *__restrict__ vols,
                                                               const int4 *__restrict__ tiles, long long t0,
                                                               long long t1) {
  for (long long t = t0 + blockIdx.x; t < t1; t += gridDim.x) {
    const int4 tl = tiles[t];
    const PostArgs<G> &a = vols[tl.x];
    const long long ip = (long long)tl.z + threadIdx.x;
    if (ip >= a.plane) continue;
    const long long p = (long long)tl.y * a.plane + ip;
    G S[6];
    load_prior(a, p, S);
#pragma unroll
    for (int q = 0; q < 6; ++q) a.cov[q][p] = mk(S[q], (G)0);
  }
}

template <typename G, int NT, int MODE>
__global__ void __launch_bounds__(256) k_posterior_class(const PostArgs<G> *__restrict__ vols,
                                                         const int4 *__restrict__ tiles, long long t0,
                                                         long long t1) {
  for (long long t = t0 + blockIdx.x; t < t1; t += gridDim.x) {
    const int4 tl = tiles[t];
    const long long ip = (long long)tl.z + threadIdx.x;
    post_any<G, NT, MODE>(vols[tl.x], tl.y, ip, ip - (threadIdx.x & 63));
  }
}

template <typename G, int MODE>
__global__ void __launch_bounds__(256) k_posterior_batch(const PostArgs<G> *__restrict__ vols,
                                                         const int4 *__restrict__ tiles, long long ntiles) {
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int4 tl = tiles[t];
    const PostArgs<G> &a = vols[tl.x];
    const long long ip = (long long)tl.z + threadIdx.x;
    const long long w0 = ip - (threadIdx.x & 63);
    switch (tl.w) {
      case 1: post_any<G, 1, MODE>(a, tl.y, ip, w0); break;
      case 2: post_any<G, 2, MODE>(a, tl.y, ip, w0); break;
      case 3: post_any<G, 3, MODE>(a, tl.y, ip, w0); break;
      case 4: post_any<G, 4, MODE>(a, tl.y, ip, w0); break;
      case 5: post_any<G, 5, MODE>(a, tl.y, ip, w0); break;
      case 6: post_any<G, 6, MODE>(a, tl.y, ip, w0); break;
      case 7: post_any<G, 7, MODE>(a, tl.y, ip, w0); break;
      default: post_any<G, 8, MODE>(a, tl.y, ip, w0); break;
    }
  }
}

// ---------------------------------------------------------------------------
void build_perk(const PostProblem &pp, const unsigned char *invert, const float *kWD, const float *e1,
                const float *e2, const float *e3, PerK<float> *tf, PerK<double> *td) {
  const int n = pp.ntheta;
  for (int k = 0; k < pp.nzp; ++k) {
    PerK<double> t{};
    t.inv = invert[k] ? 1 : 0;
    cx<double> w[kMaxTheta], m1[kMaxTheta], m2[kMaxTheta];
    double m3[kMaxTheta];
    for (int l = 0; l < n; ++l) {
      const long long o = ((long long)k * n + l) * 2;
      w[l] = mk((double)kWD[o], (double)kWD[o + 1]);
      m1[l] = mk((double)e1[o], (double)e1[o + 1]);
      m2[l] = mk((double)e2[o], (double)e2[o + 1]);
      m3[l] = (double)e3[o];
      t.kwd[l] = w[l];
    }
    const double c0 = 0.5 * (1.0 - (double)pp.wnc);
    t.trEc = 0.0;
    t.trEd = 0.0;
    for (int l = 0; l < n; ++l) {
      for (int m = 0; m <= l; ++m) {
        const int q = l * (l + 1) / 2 + m;
        t.P[q] = mulc(w[l], w[m]);
        const cx<double> g = add(mulc(m1[l], m1[m]), mulc(m2[l], m2[m]));
        const double c = c0 * pp.err_theta_cov[l * n + m];
        t.Ec[q] = mk(c * g.x, l == m ? 0.0 : c * g.y);
      }
      t.P[l * (l + 1) / 2 + l].y = 0.0;
      t.Ed[l] = (double)pp.wnc * pp.err_theta_cov[l * n + l] * m3[l] * m3[l];
      t.trEc += t.Ec[l * (l + 1) / 2 + l].x;
      t.trEd += t.Ed[l];
    }
    td[k] = t;
    PerK<float> f{};
    f.inv = t.inv;
    for (int l = 0; l < kMaxTheta; ++l) {
      f.kwd[l] = mk((float)t.kwd[l].x, (float)t.kwd[l].y);
      f.Ed[l] = (float)t.Ed[l];
    }
    for (int q = 0; q < kTri; ++q) {
      f.P[q] = mk((float)t.P[q].x, (float)t.P[q].y);
      f.Ec[q] = mk((float)t.Ec[q].x, (float)t.Ec[q].y);
    }
    // the float table keeps the reference's float kWD exactly
    for (int l = 0; l < n; ++l) {
      const long long o = ((long long)k * n + l) * 2;
      f.kwd[l] = mk(kWD[o], kWD[o + 1]);
    }
    f.trEc = (float)t.trEc;
    f.trEd = (float)t.trEd;
    tf[k] = f;
  }
}

std::vector<int> posterior_kmap(int nzp, const unsigned char *invert, int *nk_in) {
  std::vector<int> m;
  m.reserve((size_t)nzp);
  for (int k = 0; k < nzp; ++k)
    if (invert[k]) m.push_back(k);
  *nk_in = (int)m.size();
  for (int k = 0; k < nzp; ++k)
    if (!invert[k]) m.push_back(k);
  return m;
}

long long defer_capacity(const PostProblem &pp, const unsigned char *invert) {
  const long long plane = (long long)(pp.nxp / 2 + 1) * pp.nyp;
  long long nk = 0;
  for (int k = 0; k < pp.nzp; ++k) nk += invert[k] ? 1 : 0;
  return nk * ((plane + 63) / 64);
}

// Workgroups of the FP64 kernel over the queue (4 waves each, striding).
static unsigned queue_blocks(long long cap) {
  return (unsigned)std::max<long long>(1, std::min<long long>(1024, (cap + 3) / 4));
}

// With a k-plane map: the solve over the in-band planes, the prior
// pass-through over the others; without: the solve kernel over every plane.
template <typename G, int NT, int MODE>
static void launch_nt(const PostArgs<G> &a, int nzp, int nk_in, long long cap, hipStream_t s) {
  const unsigned gx = (unsigned)((a.plane + 255) / 256);
  if constexpr (MODE == 1) hipLaunchKernelGGL(k_zero_queue_count, dim3(1), dim3(1), 0, s, a.defer_count);
  const int nsolve = a.kmap ? nk_in : nzp;
  if (nsolve > 0)
    hipLaunchKernelGGL((k_posterior<G, NT, MODE>), dim3(gx, (unsigned)nsolve), dim3(256), 0, s, a, 0);
  if (a.kmap && nzp > nk_in)
    hipLaunchKernelGGL((k_posterior_prior<G>), dim3(gx, (unsigned)(nzp - nk_in)), dim3(256), 0, s, a, nk_in);
  if constexpr (MODE == 1)
    hipLaunchKernelGGL((k_posterior_queued<NT>), dim3(queue_blocks(cap)), dim3(256), 0, s, a);
}

template <typename G, int MODE>
static int dispatch(const PostArgs<G> &a, int nt, int nzp, int nk_in, long long cap, hipStream_t s) {
  switch (nt) {
    case 1: launch_nt<G, 1, MODE>(a, nzp, nk_in, cap, s); break;
    case 2: launch_nt<G, 2, MODE>(a, nzp, nk_in, cap, s); break;
    case 3: launch_nt<G, 3, MODE>(a, nzp, nk_in, cap, s); break;
    case 4: launch_nt<G, 4, MODE>(a, nzp, nk_in, cap, s); break;
    case 5: launch_nt<G, 5, MODE>(a, nzp, nk_in, cap, s); break;
    case 6: launch_nt<G, 6, MODE>(a, nzp, nk_in, cap, s); break;
    case 7: launch_nt<G, 7, MODE>(a, nzp, nk_in, cap, s); break;
    case 8: launch_nt<G, 8, MODE>(a, nzp, nk_in, cap, s); break;
    default: set_error("ntheta must be 1..8"); return CRAVA_HIP_E_INVALID;
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

template <typename G>
static void fill_args(PostArgs<G> &a, const PostDevice &pd, void *const *mu, void *const *data, void *const *cov,
                      const void *errcorr, unsigned long long *counters, const void *corr = nullptr) {
  const PostProblem &pp = pd.pp;
  for (int c = 0; c < 3; ++c) a.mu[c] = reinterpret_cast<cx<G> *>(mu[c]);
  for (int l = 0; l < kMaxTheta; ++l) a.d[l] = l < pp.ntheta ? reinterpret_cast<cx<G> *>(data[l]) : nullptr;
  for (int q = 0; q < 6; ++q) a.cov[q] = reinterpret_cast<cx<G> *>(cov[q]);
  a.ec = reinterpret_cast<const cx<G> *>(errcorr);
  a.corr = reinterpret_cast<const cx<G> *>(corr);
  a.tabf = pd.tabf;
  a.tabd = pd.tabd;
  a.counters = counters;
  a.defer_list = nullptr;
  a.defer_count = nullptr;
  a.plane = (long long)(pp.nxp / 2 + 1) * pp.nyp;
  a.ec_kstride = pd.ec_plane ? 0 : a.plane;
  a.nt = pp.ntheta;
  for (int l = 0; l < kMaxTheta; ++l)
    for (int b = 0; b < 3; ++b) {
      a.Af[l][b] = pp.A[l * 3 + b];
      a.Ad[l][b] = (double)pp.A[l * 3 + b];
    }
  static const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
  for (int q = 0; q < 6; ++q) {
    const double s = pp.sigma0[ia[q] * 3 + ib[q]];
    a.s0nz[q] = s != 0.0;
    a.s0f[q] = (float)s;
    a.s0d[q] = s;
  }
  a.thr = (float)pp.mixed_threshold;
  a.kmap = pd.kmap;
}

int launch_posterior(const PostDevice &pd, int precision, void *const *mu, void *const *data, void *const *cov,
                     const void *errcorr, unsigned long long *counters, const DeferQueue *defer, hipStream_t s,
                     const void *corr) {
  const int nt = pd.pp.ntheta;
  if (precision == CRAVA_HIP_F64) {
    PostArgs<double> a{};
    fill_args(a, pd, mu, data, cov, errcorr, counters, corr);
    return dispatch<double, 0>(a, nt, pd.pp.nzp, pd.nk_in, 0, s);
  }
  PostArgs<float> a{};
  fill_args(a, pd, mu, data, cov, errcorr, counters, corr);
  if (precision == CRAVA_HIP_MIXED) {
    if (!defer || !defer->list || !defer->count || defer->cap < pd.defer_need) {
      set_error("the mixed posterior needs a queue of " + std::to_string(pd.defer_need) + " entries");
      return CRAVA_HIP_E_INVALID;
    }
    a.defer_list = defer->list;
    a.defer_count = defer->count;
    return dispatch<float, 1>(a, nt, pd.pp.nzp, pd.nk_in, defer->cap, s);
  }
  return dispatch<float, 0>(a, nt, pd.pp.nzp, pd.nk_in, 0, s);
}

// ---------------------------------------------------------------------------
PostBatch::~PostBatch() {
  if (args) (void)hipFree(args);
  if (tiles) (void)hipFree(tiles);
  if (queues) (void)hipFree(queues);
}

int PostBatch::init(int prec, const PostVolume *vols, int n) {
  if (n < 1 || !vols) {
    set_error("a posterior batch needs at least one volume");
    return CRAVA_HIP_E_INVALID;
  }
  precision = prec;
  nvol = n;
  nt.resize(n);
  nzp.resize(n);
  plane.resize(n);
  const size_t asz = prec == CRAVA_HIP_F64 ? sizeof(PostArgs<double>) : sizeof(PostArgs<float>);
  std::vector<unsigned char> host(asz * (size_t)n);
  // mixed form: one queue per volume, the counts first, then the lists
  vq.assign(n, DeferQueue{});
  if (prec == CRAVA_HIP_MIXED) {
    long long total = 0;
    for (int v = 0; v < n; ++v) total += vols[v].pd->defer_need;
    const size_t head = ((sizeof(unsigned) * (size_t)n + 255) / 256) * 256;
    CRAVA_HIP_CHECK(hipMalloc(&queues, head + sizeof(int2) * (size_t)std::max<long long>(total, 1)));
    unsigned *counts = static_cast<unsigned *>(queues);
    int2 *lists = reinterpret_cast<int2 *>(static_cast<unsigned char *>(queues) + head);
    long long off = 0;
    for (int v = 0; v < n; ++v) {
      vq[v].count = counts + v;
      vq[v].list = lists + off;
      vq[v].cap = vols[v].pd->defer_need;
      off += vq[v].cap;
    }
  }
  inv.assign(n, {});
  for (int v = 0; v < n; ++v) {
    const PostVolume &pv = vols[v];
    const PostProblem &pp = pv.pd->pp;
    inv[v] = pv.pd->invert;
    if ((int)inv[v].size() != pp.nzp) inv[v].assign(pp.nzp, 1);  // no gate kept: every plane through the solve
    nt[v] = pp.ntheta;
    nzp[v] = pp.nzp;
    plane[v] = (long long)(pp.nxp / 2 + 1) * pp.nyp;
    if (prec == CRAVA_HIP_F64) {
      PostArgs<double> a{};
      fill_args(a, *pv.pd, pv.mu, pv.data, pv.cov, pv.errcorr, pv.counters, pv.corr);
      std::memcpy(host.data() + asz * v, &a, sizeof(a));
    } else {
      PostArgs<float> a{};
      fill_args(a, *pv.pd, pv.mu, pv.data, pv.cov, pv.errcorr, pv.counters, pv.corr);
      a.defer_list = vq[v].list;
      a.defer_count = vq[v].count;
      std::memcpy(host.data() + asz * v, &a, sizeof(a));
    }
  }
  CRAVA_HIP_CHECK(hipMalloc(&args, host.size()));
  CRAVA_HIP_CHECK(hipMemcpy(args, host.data(), host.size(), hipMemcpyHostToDevice));
  return set_launch(256, 0);
}

int PostBatch::set_launch(int points_per_block, int nblocks) {
  if (points_per_block != 64 && points_per_block != 128 && points_per_block != 256) {
    set_error("points per block must be 64, 128 or 256");
    return CRAVA_HIP_E_INVALID;
  }
  if (nblocks < 0) {
    set_error("negative block count");
    return CRAVA_HIP_E_INVALID;
  }
  if (points_per_block != ppb || !tiles) {
    // in-band tiles ordered by ntheta, so each ntheta class is one
    // contiguous segment; then every out-of-band tile
    std::vector<int4> h;
    for (int c = 1; c <= kMaxTheta; ++c) {
      seg[c - 1] = (long long)h.size();
      for (int v = 0; v < nvol; ++v) {
        if (nt[v] != c) continue;
        for (int k = 0; k < nzp[v]; ++k)
          if (inv[v][k])
            for (long long p0 = 0; p0 < plane[v]; p0 += points_per_block) h.push_back(make_int4(v, k, (int)p0, c));
      }
    }
    seg[kMaxTheta] = (long long)h.size();
    for (int v = 0; v < nvol; ++v)
      for (int k = 0; k < nzp[v]; ++k)
        if (!inv[v][k])
          for (long long p0 = 0; p0 < plane[v]; p0 += points_per_block) h.push_back(make_int4(v, k, (int)p0, nt[v]));
    if (tiles) (void)hipFree(tiles);
    tiles = nullptr;
    CRAVA_HIP_CHECK(hipMalloc(&tiles, sizeof(int4) * h.size()));
    CRAVA_HIP_CHECK(hipMemcpy(tiles, h.data(), sizeof(int4) * h.size(), hipMemcpyHostToDevice));
    ntiles = (long long)h.size();
    ppb = points_per_block;
  }
  blocks = nblocks;
  return CRAVA_HIP_OK;
}

int PostBatch::num_classes() const {
  int n = 0;
  for (int c = 0; c < kMaxTheta; ++c) n += seg[c + 1] > seg[c];
  return n;
}

template <typename G, int NT, int MODE>
static void launch_class(const PostBatch &b, long long t0, long long t1, hipStream_t s) {
  const long long n = t1 - t0;
  const long long g = b.blocks > 0 ? std::min<long long>(b.blocks, n) : n;
  hipLaunchKernelGGL((k_posterior_class<G, NT, MODE>), dim3((unsigned)g), dim3((unsigned)b.ppb), 0, s,
                     reinterpret_cast<const PostArgs<G> *>(b.args), b.tiles, t0, t1);
}

template <typename G, int MODE>
static int launch_classes(const PostBatch &b, const hipStream_t *streams, int nstreams) {
  int i = 0;
  for (int c = 1; c <= kMaxTheta; ++c) {
    const long long t0 = b.seg[c - 1], t1 = b.seg[c];
    if (t1 <= t0) continue;
    hipStream_t s = streams[i++ % nstreams];
    switch (c) {
      case 1: launch_class<G, 1, MODE>(b, t0, t1, s); break;
      case 2: launch_class<G, 2, MODE>(b, t0, t1, s); break;
      case 3: launch_class<G, 3, MODE>(b, t0, t1, s); break;
      case 4: launch_class<G, 4, MODE>(b, t0, t1, s); break;
      case 5: launch_class<G, 5, MODE>(b, t0, t1, s); break;
      case 6: launch_class<G, 6, MODE>(b, t0, t1, s); break;
      case 7: launch_class<G, 7, MODE>(b, t0, t1, s); break;
      default: launch_class<G, 8, MODE>(b, t0, t1, s); break;
    }
    CRAVA_HIP_CHECK(hipGetLastError());
  }
  return CRAVA_HIP_OK;
}

int PostBatch::launch(const hipStream_t *streams, int nstreams) const {
  if (ntiles < 1) return CRAVA_HIP_OK;
  // the out-of-band tiles, on the last stream
  const long long p0 = seg[kMaxTheta], np = ntiles - p0;
  if (np > 0) {
    const long long g = blocks > 0 ? std::min<long long>(blocks, np) : np;
    hipStream_t s = streams[nstreams - 1];
    if (precision == CRAVA_HIP_F64)
      hipLaunchKernelGGL((k_posterior_prior_tiles<double>), dim3((unsigned)g), dim3((unsigned)ppb), 0, s,
                         reinterpret_cast<const PostArgs<double> *>(args), tiles, p0, ntiles);
    else
      hipLaunchKernelGGL((k_posterior_prior_tiles<float>), dim3((unsigned)g), dim3((unsigned)ppb), 0, s,
                         reinterpret_cast<const PostArgs<float> *>(args), tiles, p0, ntiles);
    CRAVA_HIP_CHECK(hipGetLastError());
  }
  if (form == kFormPerClass) {
    if (precision == CRAVA_HIP_F64) return launch_classes<double, 0>(*this, streams, nstreams);
    if (precision == CRAVA_HIP_MIXED) return launch_classes<float, 1>(*this, streams, nstreams);
    return launch_classes<float, 0>(*this, streams, nstreams);
  }
  const long long ns = seg[kMaxTheta];
  if (ns < 1) return CRAVA_HIP_OK;
  const long long g = blocks > 0 ? std::min<long long>(blocks, ns) : ns;
  const dim3 grid((unsigned)g), block((unsigned)ppb);
  hipStream_t s = streams[0];
  if (precision == CRAVA_HIP_F64)
    hipLaunchKernelGGL((k_posterior_batch<double, 0>), grid, block, 0, s,
                       reinterpret_cast<const PostArgs<double> *>(args), tiles, ns);
  else if (precision == CRAVA_HIP_MIXED)
    hipLaunchKernelGGL((k_posterior_batch<float, 1>), grid, block, 0, s,
                       reinterpret_cast<const PostArgs<float> *>(args), tiles, ns);
  else
    hipLaunchKernelGGL((k_posterior_batch<float, 0>), grid, block, 0, s,
                       reinterpret_cast<const PostArgs<float> *>(args), tiles, ns);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int PostBatch::reset_queues(hipStream_t s) const {
  if (precision != CRAVA_HIP_MIXED) return CRAVA_HIP_OK;
  hipLaunchKernelGGL(k_zero_queue_counts, dim3((unsigned)((nvol + 255) / 256)), dim3(256), 0, s,
                     reinterpret_cast<const PostArgs<float> *>(args), nvol);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int PostBatch::launch_queued(hipStream_t s) const {
  if (precision != CRAVA_HIP_MIXED) return CRAVA_HIP_OK;
  long long cap = 0;
  for (const DeferQueue &q : vq) cap = std::max(cap, q.cap);
  hipLaunchKernelGGL(k_posterior_queued_batch, dim3(queue_blocks(cap)), dim3(256), 0, s,
                     reinterpret_cast<const PostArgs<float> *>(args), nvol);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

}  // namespace crava
