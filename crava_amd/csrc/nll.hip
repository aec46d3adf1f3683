// Negative log marginal likelihood of the seismic data and its adjoint
// gradient (no reference counterpart: CRAVA evaluates the posterior only).
//
// Model per in-band half-complex Fourier point p (the one the posterior loop
// of src/crava.cpp:1087-1130 assumes):
//   d(p) ~ CN(K mu(p), M(p)),  M = K S K^H + E,
//   K = diag(s) K0,  S = |c(p)| Sigma0,  Sigma0 = L0 L0^T,
//   E = lambda(p) Ec(Sigma_e) + diag(wnc Sigma_e,ll),
//   Sigma_e,lm = sqrt(v_l v_m) rho_lm.
// NLL = 1/2 sum_p w_p [ r^H M^-1 r + log det M ],  r = d - K mu,
// w_p = 1 on the self-conjugate x planes, 2 elsewhere (the half spectrum
// stands for the full one of a real field).
//
// With L = chol(M) recomputed per point, z = L^-1 r, alpha = L^-H z and
// G = M^-1 - alpha alpha^H:
//   d f / d log s_l   = 2 Re[(K S K^H G)_ll - conj(alpha_l) (K mu)_l]
//   d f / d Sigma_e,lm = lambda c0 G_ml g_lm + delta_lm wnc e3_l^2 G_ll
//   d f / d Sigma0_ab  = |c| Re(K^H G K)_ba
// Each workgroup writes its partial sums; the host reduces them in FP64.
#include "../../include/crava_hip.h"
#include "common.hpp"
#include "nll.hpp"

namespace crava {

template <int NT>
struct NllAcc {
  static constexpr int kF = 0;
  static constexpr int kS = 1;                          // NT log-scale terms
  static constexpr int kN = 1 + NT;                     // NT(NT+1)/2 noise terms (tri, l >= m)
  static constexpr int kP = 1 + NT + NT * (NT + 1) / 2; // 6 prior terms (aa, bb, rr, ab, ar, br)
  static constexpr int kCount = kP + 6;
};

template <typename G>
struct NllArgs {
  const cx<G> *mu[3];
  const cx<G> *d[kMaxTheta];
  const cx<G> *cov0;
  const cx<G> *ec;
  long long ec_kstride;   // ec of point (k, ip) is ec[k ec_kstride + ip]
  const NllTab *tab;      // per k
  const int *klist;       // in-band planes
  long long plane;
  long long total;        // n_inband * plane
  int cnxp;
  int nxp_even;
  double A[kMaxTheta][3];
  double sig0[6];         // Sigma0 (aa, bb, rr, ab, ar, br)
  double base00;          // Sigma0_aa the covariance grids were built with
  double *partial;        // [gridDim.x][NllAcc::kCount]
};

template <int NT>
__device__ __forceinline__ void point_terms(const NllTab &t, const double (&A)[kMaxTheta][3], const double (&s0)[6],
                                            double cabs, double lam, const cx<double> (&mu)[3],
                                            const cx<double> (&d)[NT], double w, double *acc) {
  using C = cx<double>;
  const double Sm[3][3] = {{s0[0], s0[3], s0[4]}, {s0[3], s0[1], s0[5]}, {s0[4], s0[5], s0[2]}};
  // K = diag(kwd) A, signal M_s = K S K^H with S = cabs Sigma0
  double AS[NT][3];
#pragma unroll
  for (int l = 0; l < NT; ++l)
#pragma unroll
    for (int b = 0; b < 3; ++b)
      AS[l][b] = cabs * (A[l][0] * Sm[0][b] + A[l][1] * Sm[1][b] + A[l][2] * Sm[2][b]);
  C Ms[NT][NT], M[NT][NT];
#pragma unroll
  for (int l = 0; l < NT; ++l)
#pragma unroll
    for (int m = 0; m <= l; ++m) {
      const int q = l * (l + 1) / 2 + m;
      const double B = AS[l][0] * A[m][0] + AS[l][1] * A[m][1] + AS[l][2] * A[m][2];
      const C pv = mulc(t.kwd[l], t.kwd[m]);
      Ms[l][m] = scl(pv, B);
      C e = scl(t.Ec[q], lam);
      if (l == m) e = mk(e.x + t.Ed[l], 0.0);
      M[l][m] = add(Ms[l][m], e);
      if (l == m) M[l][m].y = 0.0;
    }
  // Cholesky (lower) of M
  C L[NT][NT];
  double rinv[NT], logdet = 0.0;
#pragma unroll
  for (int i = 0; i < NT; ++i) {
#pragma unroll
    for (int j = 0; j < i; ++j) {
      C sacc = M[i][j];
#pragma unroll
      for (int k = 0; k < j; ++k) sacc = sub(sacc, mulc(L[i][k], L[j][k]));
      L[i][j] = scl(sacc, rinv[j]);
    }
    double dd = M[i][i].x;
#pragma unroll
    for (int k = 0; k < i; ++k) dd -= norm2(L[i][k]);
    dd = dd > 0.0 ? dd : 1e-300;
    const double li = sqrt(dd);
    L[i][i] = mk(li, 0.0);
    rinv[i] = 1.0 / li;
    logdet += 2.0 * log(li);
  }
  // Linv = L^-1 (lower)
  C Li[NT][NT];
#pragma unroll
  for (int j = 0; j < NT; ++j) {
#pragma unroll
    for (int i = 0; i < NT; ++i) Li[i][j] = mk(0.0, 0.0);
    Li[j][j] = mk(rinv[j], 0.0);
#pragma unroll
    for (int i = j + 1; i < NT; ++i) {
      C sacc = mk(0.0, 0.0);
#pragma unroll
      for (int k = j; k < i; ++k) sacc = add(sacc, mul(L[i][k], Li[k][j]));
      Li[i][j] = scl(sacc, -rinv[i]);
    }
  }
  // r = d - K mu, z = Linv r, alpha = Linv^H z
  C Kmu[NT], z[NT], alpha[NT];
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    const C am = mk(A[l][0] * mu[0].x + A[l][1] * mu[1].x + A[l][2] * mu[2].x,
                    A[l][0] * mu[0].y + A[l][1] * mu[1].y + A[l][2] * mu[2].y);
    Kmu[l] = mul(t.kwd[l], am);
  }
  double quad = 0.0;
#pragma unroll
  for (int i = 0; i < NT; ++i) {
    C sacc = mk(0.0, 0.0);
#pragma unroll
    for (int k = 0; k <= i; ++k) sacc = add(sacc, mul(Li[i][k], sub(d[k], Kmu[k])));
    z[i] = sacc;
    quad += norm2(sacc);
  }
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    C sacc = mk(0.0, 0.0);
#pragma unroll
    for (int i = l; i < NT; ++i) sacc = add(sacc, cmul(Li[i][l], z[i]));
    alpha[l] = sacc;
  }
  acc[NllAcc<NT>::kF] += w * (quad + logdet);
  // G = M^-1 - alpha alpha^H, full Hermitian (Minv_lm = sum_i conj(Li_il) Li_im)
  C Gm[NT][NT];
#pragma unroll
  for (int l = 0; l < NT; ++l)
#pragma unroll
    for (int m = 0; m < NT; ++m) {
      C sacc = mk(0.0, 0.0);
      const int i0 = l > m ? l : m;
#pragma unroll
      for (int i = 0; i < NT; ++i)
        if (i >= i0) sacc = add(sacc, cmul(Li[i][l], Li[i][m]));
      Gm[l][m] = sub(sacc, mulc(alpha[l], alpha[m]));
    }
  // wavelet log-scales: 2 Re[(Ms G)_ll - conj(alpha_l) Kmu_l]  (dr = -dK mu)
#pragma unroll
  for (int l = 0; l < NT; ++l) {
    double tr = 0.0;
#pragma unroll
    for (int m = 0; m < NT; ++m) {
      const C msl = m <= l ? Ms[l][m] : mk(Ms[m][l].x, -Ms[m][l].y);
      const C p = mul(msl, Gm[m][l]);
      tr += p.x;
    }
    acc[NllAcc<NT>::kS + l] += w * 2.0 * (tr - cmul(alpha[l], Kmu[l]).x);
  }
  // noise: T_lm = lam c0 G_ml g_lm + delta wd_l G_ll, stored as T_ll and Re(T_lm + T_ml), l > m
#pragma unroll
  for (int l = 0; l < NT; ++l)
#pragma unroll
    for (int m = 0; m <= l; ++m) {
      const int q = l * (l + 1) / 2 + m;
      const C g = t.g[q];                          // g_lm, l >= m
      double v;
      if (l == m) {
        v = lam * t.c0 * Gm[l][l].x * g.x + t.wd[l] * Gm[l][l].x;
      } else {
        // T_lm + T_ml = lam c0 (G_ml g_lm + G_lm conj(g_lm))
        const C a = mul(Gm[m][l], g);
        const C b = mulc(Gm[l][m], g);
        v = lam * t.c0 * (a.x + b.x);
      }
      acc[NllAcc<NT>::kN + q] += w * v;
    }
  // prior: |c| Re(K^H G K), K = diag(kwd) A
  C GK[NT][3];
#pragma unroll
  for (int l = 0; l < NT; ++l)
#pragma unroll
    for (int b = 0; b < 3; ++b) {
      C sacc = mk(0.0, 0.0);
#pragma unroll
      for (int m = 0; m < NT; ++m) sacc = add(sacc, scl(mul(Gm[l][m], t.kwd[m]), A[m][b]));
      GK[l][b] = sacc;
    }
  const int ia[6] = {0, 1, 2, 0, 0, 1}, ib[6] = {0, 1, 2, 1, 2, 2};
#pragma unroll
  for (int q = 0; q < 6; ++q) {
    double sacc = 0.0;
#pragma unroll
    for (int l = 0; l < NT; ++l) sacc += A[l][ia[q]] * cmul(t.kwd[l], GK[l][ib[q]]).x;
    acc[NllAcc<NT>::kP + q] += w * cabs * sacc;
  }
}

template <typename G, int NT>
__global__ void __launch_bounds__(256) k_nll(NllArgs<G> a) {
  constexpr int NA = NllAcc<NT>::kCount;
  double acc[NA];
#pragma unroll
  for (int i = 0; i < NA; ++i) acc[i] = 0.0;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < a.total;
       idx += (long long)gridDim.x * blockDim.x) {
    const long long kk = idx / a.plane;
    const long long ip = idx - kk * a.plane;
    const int k = a.klist[kk];
    const long long p = (long long)k * a.plane + ip;
    const int i = (int)(ip % a.cnxp);
    const double w = (i == 0 || (a.nxp_even && i == a.cnxp - 1)) ? 1.0 : 2.0;
    const NllTab &t = a.tab[k];
    cx<double> mu[3], d[NT];
#pragma unroll
    for (int c = 0; c < 3; ++c) mu[c] = mk((double)a.mu[c][p].x, (double)a.mu[c][p].y);
#pragma unroll
    for (int l = 0; l < NT; ++l) d[l] = mk((double)a.d[l][p].x, (double)a.d[l][p].y);
    const double cabs = fabs((double)a.cov0[p].x / a.base00);
    const double lam = fabs((double)a.ec[(long long)k * a.ec_kstride + ip].x);
    point_terms<NT>(t, a.A, a.sig0, cabs, lam, mu, d, w, acc);
  }
  // block reduction through LDS, one value at a time
  __shared__ double red[256];
  for (int v = 0; v < NA; ++v) {
    red[threadIdx.x] = acc[v];
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
      if ((int)threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
      __syncthreads();
    }
    if (threadIdx.x == 0) a.partial[(long long)blockIdx.x * NA + v] = red[0];
    __syncthreads();
  }
}

int nll_accumulator_count(int nt) { return 1 + nt + nt * (nt + 1) / 2 + 6; }

template <typename G, int NT>
static void launch_nt(const NllArgs<G> &a, int blocks, hipStream_t s) {
  hipLaunchKernelGGL((k_nll<G, NT>), dim3(blocks), dim3(256), 0, s, a);
}

template <typename G>
static int launch_g(const NllCall &c, hipStream_t s) {
  NllArgs<G> a{};
  for (int q = 0; q < 3; ++q) a.mu[q] = reinterpret_cast<const cx<G> *>(c.mu[q]);
  for (int l = 0; l < kMaxTheta; ++l) a.d[l] = l < c.nt ? reinterpret_cast<const cx<G> *>(c.data[l]) : nullptr;
  a.cov0 = reinterpret_cast<const cx<G> *>(c.cov0);
  a.ec = reinterpret_cast<const cx<G> *>(c.errcorr);
  a.tab = c.tab;
  a.klist = c.klist;
  a.cnxp = c.nxp / 2 + 1;
  a.nxp_even = (c.nxp % 2) == 0;
  a.plane = (long long)a.cnxp * c.nyp;
  a.total = (long long)c.n_inband * a.plane;
  a.ec_kstride = c.ec_plane ? 0 : a.plane;
  for (int l = 0; l < kMaxTheta; ++l)
    for (int b = 0; b < 3; ++b) a.A[l][b] = c.A[l * 3 + b];
  for (int q = 0; q < 6; ++q) a.sig0[q] = c.sig0[q];
  a.base00 = c.base00;
  a.partial = c.partial;
  switch (c.nt) {
    case 1: launch_nt<G, 1>(a, c.blocks, s); break;
    case 2: launch_nt<G, 2>(a, c.blocks, s); break;
    case 3: launch_nt<G, 3>(a, c.blocks, s); break;
    case 4: launch_nt<G, 4>(a, c.blocks, s); break;
    case 5: launch_nt<G, 5>(a, c.blocks, s); break;
    case 6: launch_nt<G, 6>(a, c.blocks, s); break;
    case 7: launch_nt<G, 7>(a, c.blocks, s); break;
    default: launch_nt<G, 8>(a, c.blocks, s); break;
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_nll(const NllCall &c, hipStream_t s) {
  if (c.nt < 1 || c.nt > kMaxTheta) {
    set_error("ntheta must be 1..8");
    return CRAVA_HIP_E_INVALID;
  }
  return c.dbl ? launch_g<double>(c, s) : launch_g<float>(c, s);
}

}  // namespace crava
