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
  const PerK<float> *tabf;
  const PerK<double> *tabd;
  unsigned long long *counters;  // [0] failed Cholesky, [1] points solved in FP64
  long long plane;
  float Af[kMaxTheta][3];
  double Ad[kMaxTheta][3];
  float s0f[6];
  double s0d[6];
  int s0nz[6];
  float thr;
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

// MODE 0: arithmetic in G.  MODE 1 (G = float): FP32, except points whose
// noise/signal ratio tr(E)/tr(K S K^H) is below thr, and the other points of
// their 64-lane wave, which are solved end-to-end in FP64 from the FP32
// inputs and rounded once at the end.
template <typename G, int NT, int MODE>
__device__ __forceinline__ void post_point(const PostArgs<G> &a, int k, long long ip) {
  const bool live = ip < a.plane;
  const long long p = (long long)k * a.plane + (live ? ip : 0);
  const PerK<G> *tb = pick_tab<G>(a.tabf, a.tabd) + k;
  const G *s0 = nullptr;
  if constexpr (sizeof(G) == 4) s0 = a.s0f; else s0 = a.s0d;

  G S[6];
  if (live) {
#pragma unroll
    for (int q = 0; q < 6; ++q) {
      G v = a.cov[q][p].x;
      if (a.s0nz[q]) v = v / s0[q];
      S[q] = fabs(v) * s0[q];
    }
  }
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
  const G lam = fabs(a.ec[p].x);
  cx<G> mu[3], d[NT], post[6];
#pragma unroll
  for (int c = 0; c < 3; ++c) mu[c] = a.mu[c][p];
#pragma unroll
  for (int l = 0; l < NT; ++l) d[l] = a.d[l][p];

  bool use64 = false;
  if constexpr (MODE == 1) {
    // tr(K S K^H) = sum_l |kWD_l|^2 (A S A^T)_ll
    G sig = (G)0;
#pragma unroll
    for (int l = 0; l < NT; ++l) {
      const G b0 = A[l][0] * S[0] + A[l][1] * S[3] + A[l][2] * S[4];
      const G b1 = A[l][0] * S[3] + A[l][1] * S[1] + A[l][2] * S[5];
      const G b2 = A[l][0] * S[4] + A[l][1] * S[5] + A[l][2] * S[2];
      sig += tb->P[l * (l + 1) / 2 + l].x * (b0 * A[l][0] + b1 * A[l][1] + b2 * A[l][2]);
    }
    const G noise = lam * tb->trEc + tb->trEd;
    // one path per wave: when any of its points needs FP64 the whole wave
    // solves in FP64, instead of running both paths one after the other
    use64 = __ballot(noise < (G)a.thr * sig) != 0ull;
  }

  bool ok;
  if (MODE == 1 && use64) {
    const PerK<double> *td = a.tabd + k;
    double Sd[6], Ad2[kMaxTheta][3];
#pragma unroll
    for (int q = 0; q < 6; ++q) Sd[q] = (double)S[q];
#pragma unroll
    for (int l = 0; l < NT; ++l)
#pragma unroll
      for (int b = 0; b < 3; ++b) Ad2[l][b] = a.Ad[l][b];
    cx<double> mud[3], dd[NT], postd[6];
#pragma unroll
    for (int c = 0; c < 3; ++c) mud[c] = cvt<double>(mu[c]);
#pragma unroll
    for (int l = 0; l < NT; ++l) dd[l] = cvt<double>(d[l]);
    ok = solve_point<double, NT>(td, Ad2, Sd, (double)lam, mud, dd, postd);
    if (ok) {
#pragma unroll
      for (int c = 0; c < 3; ++c) mu[c] = cvt<G>(mud[c]);
#pragma unroll
      for (int l = 0; l < NT; ++l) d[l] = cvt<G>(dd[l]);
#pragma unroll
      for (int q = 0; q < 6; ++q) post[q] = cvt<G>(postd[q]);
    }
  } else {
    ok = solve_point<G, NT>(tb, A, S, lam, mu, d, post);
  }

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
  if (a.counters) {
    const unsigned long long fails = __ballot(!ok);
    const unsigned long long wide = __ballot(MODE == 1 && use64 && ok);
    if ((threadIdx.x & 63) == 0) {
      if (fails) atomicAdd(&a.counters[0], (unsigned long long)__popcll(fails));
      if (wide) atomicAdd(&a.counters[1], (unsigned long long)__popcll(wide));
    }
  }
}

template <typename G, int NT, int MODE>
__global__ void __launch_bounds__(256) k_posterior(PostArgs<G> a) {
  // consecutive workgroups go to different XCDs, which spreads the in-band
  // k-planes (a contiguous range of k) over all of them
  post_point<G, NT, MODE>(a, blockIdx.y, (long long)blockIdx.x * blockDim.x + threadIdx.x);
}

// Many volumes in one persistent launch.  Tile t = {volume, k, first point,
// ntheta} covers blockDim.x consecutive points of one k-plane of one
// volume; workgroups stride over the tile list.  Everything a tile reads
// from `vols` and `tiles` is wave-uniform.  This form holds the registers of
// the largest ntheta for every tile; k_posterior_class below is the same
// loop over the tiles of one ntheta.
template <typename G, int NT, int MODE>
__global__ void __launch_bounds__(256) k_posterior_class(const PostArgs<G> *__restrict__ vols,
                                                         const int4 *__restrict__ tiles, long long t0,
                                                         long long t1) {
  for (long long t = t0 + blockIdx.x; t < t1; t += gridDim.x) {
    const int4 tl = tiles[t];
    post_point<G, NT, MODE>(vols[tl.x], tl.y, (long long)tl.z + threadIdx.x);
  }
}

template <typename G, int MODE>
__global__ void __launch_bounds__(256) k_posterior_batch(const PostArgs<G> *__restrict__ vols,
                                                         const int4 *__restrict__ tiles, long long ntiles) {
  for (long long t = blockIdx.x; t < ntiles; t += gridDim.x) {
    const int4 tl = tiles[t];
    const PostArgs<G> &a = vols[tl.x];
    const long long ip = (long long)tl.z + threadIdx.x;
    switch (tl.w) {
      case 1: post_point<G, 1, MODE>(a, tl.y, ip); break;
      case 2: post_point<G, 2, MODE>(a, tl.y, ip); break;
      case 3: post_point<G, 3, MODE>(a, tl.y, ip); break;
      case 4: post_point<G, 4, MODE>(a, tl.y, ip); break;
      case 5: post_point<G, 5, MODE>(a, tl.y, ip); break;
      case 6: post_point<G, 6, MODE>(a, tl.y, ip); break;
      case 7: post_point<G, 7, MODE>(a, tl.y, ip); break;
      default: post_point<G, 8, MODE>(a, tl.y, ip); break;
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

template <typename G, int NT, int MODE>
static void launch_nt(const PostArgs<G> &a, int nzp, hipStream_t s) {
  const unsigned gx = (unsigned)((a.plane + 255) / 256);
  hipLaunchKernelGGL((k_posterior<G, NT, MODE>), dim3(gx, (unsigned)nzp), dim3(256), 0, s, a);
}

template <typename G, int MODE>
static int dispatch(const PostArgs<G> &a, int nt, int nzp, hipStream_t s) {
  switch (nt) {
    case 1: launch_nt<G, 1, MODE>(a, nzp, s); break;
    case 2: launch_nt<G, 2, MODE>(a, nzp, s); break;
    case 3: launch_nt<G, 3, MODE>(a, nzp, s); break;
    case 4: launch_nt<G, 4, MODE>(a, nzp, s); break;
    case 5: launch_nt<G, 5, MODE>(a, nzp, s); break;
    case 6: launch_nt<G, 6, MODE>(a, nzp, s); break;
    case 7: launch_nt<G, 7, MODE>(a, nzp, s); break;
    case 8: launch_nt<G, 8, MODE>(a, nzp, s); break;
    default: set_error("ntheta must be 1..8"); return CRAVA_HIP_E_INVALID;
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

template <typename G>
static void fill_args(PostArgs<G> &a, const PostDevice &pd, void *const *mu, void *const *data, void *const *cov,
                      const void *errcorr, unsigned long long *counters) {
  const PostProblem &pp = pd.pp;
  for (int c = 0; c < 3; ++c) a.mu[c] = reinterpret_cast<cx<G> *>(mu[c]);
  for (int l = 0; l < kMaxTheta; ++l) a.d[l] = l < pp.ntheta ? reinterpret_cast<cx<G> *>(data[l]) : nullptr;
  for (int q = 0; q < 6; ++q) a.cov[q] = reinterpret_cast<cx<G> *>(cov[q]);
  a.ec = reinterpret_cast<const cx<G> *>(errcorr);
  a.tabf = pd.tabf;
  a.tabd = pd.tabd;
  a.counters = counters;
  a.plane = (long long)(pp.nxp / 2 + 1) * pp.nyp;
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
}

int launch_posterior(const PostDevice &pd, int precision, void *const *mu, void *const *data, void *const *cov,
                     const void *errcorr, unsigned long long *counters, hipStream_t s) {
  const int nt = pd.pp.ntheta;
  if (precision == CRAVA_HIP_F64) {
    PostArgs<double> a{};
    fill_args(a, pd, mu, data, cov, errcorr, counters);
    return dispatch<double, 0>(a, nt, pd.pp.nzp, s);
  }
  PostArgs<float> a{};
  fill_args(a, pd, mu, data, cov, errcorr, counters);
  if (precision == CRAVA_HIP_MIXED) return dispatch<float, 1>(a, nt, pd.pp.nzp, s);
  return dispatch<float, 0>(a, nt, pd.pp.nzp, s);
}

// ---------------------------------------------------------------------------
PostBatch::~PostBatch() {
  if (args) (void)hipFree(args);
  if (tiles) (void)hipFree(tiles);
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
  for (int v = 0; v < n; ++v) {
    const PostVolume &pv = vols[v];
    const PostProblem &pp = pv.pd->pp;
    nt[v] = pp.ntheta;
    nzp[v] = pp.nzp;
    plane[v] = (long long)(pp.nxp / 2 + 1) * pp.nyp;
    if (prec == CRAVA_HIP_F64) {
      PostArgs<double> a{};
      fill_args(a, *pv.pd, pv.mu, pv.data, pv.cov, pv.errcorr, pv.counters);
      std::memcpy(host.data() + asz * v, &a, sizeof(a));
    } else {
      PostArgs<float> a{};
      fill_args(a, *pv.pd, pv.mu, pv.data, pv.cov, pv.errcorr, pv.counters);
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
    // tiles ordered by ntheta, so each ntheta class is one contiguous segment
    std::vector<int4> h;
    for (int c = 1; c <= kMaxTheta; ++c) {
      seg[c - 1] = (long long)h.size();
      for (int v = 0; v < nvol; ++v) {
        if (nt[v] != c) continue;
        for (int k = 0; k < nzp[v]; ++k)
          for (long long p0 = 0; p0 < plane[v]; p0 += points_per_block) h.push_back(make_int4(v, k, (int)p0, c));
      }
    }
    seg[kMaxTheta] = (long long)h.size();
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
  if (form == kFormPerClass) {
    if (precision == CRAVA_HIP_F64) return launch_classes<double, 0>(*this, streams, nstreams);
    if (precision == CRAVA_HIP_MIXED) return launch_classes<float, 1>(*this, streams, nstreams);
    return launch_classes<float, 0>(*this, streams, nstreams);
  }
  const long long g = blocks > 0 ? std::min<long long>(blocks, ntiles) : ntiles;
  const dim3 grid((unsigned)g), block((unsigned)ppb);
  hipStream_t s = streams[0];
  if (precision == CRAVA_HIP_F64)
    hipLaunchKernelGGL((k_posterior_batch<double, 0>), grid, block, 0, s,
                       reinterpret_cast<const PostArgs<double> *>(args), tiles, ntiles);
  else if (precision == CRAVA_HIP_MIXED)
    hipLaunchKernelGGL((k_posterior_batch<float, 1>), grid, block, 0, s,
                       reinterpret_cast<const PostArgs<float> *>(args), tiles, ntiles);
  else
    hipLaunchKernelGGL((k_posterior_batch<float, 0>), grid, block, 0, s,
                       reinterpret_cast<const PostArgs<float> *>(args), tiles, ntiles);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

}  // namespace crava
