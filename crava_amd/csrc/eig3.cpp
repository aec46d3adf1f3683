// Right eigenvectors of a dense real 3x3 matrix with the order and signs
// that LAPACK's DGEEV (netlib reference implementation, JOBVL='N',
// JOBVR='V') gives them.  NRLib::ComputeEigenVectors
// (libs/nrlib/flens/nrlib_flens.cpp:114-121) takes its eigenvectors from
// DGEEV, and Crava::computeG (src/crava.cpp:2182) uses their signs.
//
// The steps are LAPACK's published algorithm for this case:
//   DGEBAL('B')  no permutation when every off-diagonal entry is nonzero;
//                its radix-2 scaling sweep restated, and matrices it would
//                scale (as it would not a symmetric one) not handled here;
//   DGEHD2       one Householder reflector (DLARFG) on column 1;
//   DORG2R       the 2x2 block of Q = diag(1, I - tau v v^T);
//   DLAHQR       Francis double-shift QR sweeps on the Hessenberg matrix with
//                the Schur vectors accumulated (WANTT, WANTZ), DLANV2 on a
//                deflated 2x2 block;
//   DTREVC3      back substitution in the quasi-triangular Schur form and
//                back transformation, each vector scaled by its largest
//                entry, then DGEEV's scaling to unit 2-norm.
// Only real eigenvalues are handled (a symmetric matrix has nothing else).
#include <algorithm>
#include <cfloat>
#include <cmath>

#include "eig3.hpp"

namespace crava {
namespace {

constexpr double kSafmin = 2.2250738585072014e-308;  // DLAMCH('S')
constexpr double kUlp = 2.2204460492503131e-16;      // DLAMCH('P'), 2^-52

inline double fsign(double a, double b) { return std::signbit(b) ? -std::fabs(a) : std::fabs(a); }

// DLAPY2
double lapy2(double x, double y) {
  const double xa = std::fabs(x), ya = std::fabs(y);
  const double w = xa > ya ? xa : ya, z = xa > ya ? ya : xa;
  if (z == 0.0) return w;
  return w * std::sqrt(1.0 + (z / w) * (z / w));
}

// DLARFG on (alpha, x[0..n-2]); returns tau.
double larfg(int n, double &alpha, double *x) {
  if (n <= 1) return 0.0;
  double ss = 0.0;
  for (int i = 0; i < n - 1; ++i) ss += x[i] * x[i];
  const double xnorm = std::sqrt(ss);
  if (xnorm == 0.0) return 0.0;
  const double beta = -fsign(lapy2(alpha, xnorm), alpha);
  const double tau = (beta - alpha) / beta;
  const double sc = 1.0 / (alpha - beta);
  for (int i = 0; i < n - 1; ++i) x[i] *= sc;
  alpha = beta;
  return tau;
}

// DLANV2: standardise the 2x2 block [a b; c d]; returns (cs, sn).
void lanv2(double &a, double &b, double &c, double &d, double &cs, double &sn) {
  if (c == 0.0) {
    cs = 1.0;
    sn = 0.0;
  } else if (b == 0.0) {
    cs = 0.0;
    sn = 1.0;
    const double t = d;
    d = a;
    a = t;
    b = -c;
    c = 0.0;
  } else if ((a - d) == 0.0 && std::signbit(b) != std::signbit(c)) {
    cs = 1.0;
    sn = 0.0;
  } else {
    const double temp = a - d;
    double p = 0.5 * temp;
    const double bcmax = std::fmax(std::fabs(b), std::fabs(c));
    const double bcmis = std::fmin(std::fabs(b), std::fabs(c)) * fsign(1.0, b) * fsign(1.0, c);
    const double scale = std::fmax(std::fabs(p), bcmax);
    double z = (p / scale) * p + (bcmax / scale) * bcmis;
    if (z >= 4.0 * kUlp) {
      z = p + fsign(std::sqrt(scale) * std::sqrt(z), p);
      a = d + z;
      d = d - (bcmax / z) * bcmis;
      const double tau = lapy2(c, z);
      cs = z / tau;
      sn = c / tau;
      b = b - c;
      c = 0.0;
    } else {
      const double sigma = b + c;
      const double tau = lapy2(sigma, temp);
      cs = std::sqrt(0.5 * (1.0 + std::fabs(sigma) / tau));
      sn = -(p / (tau * cs)) * fsign(1.0, sigma);
      const double aa = a * cs + b * sn, bb = -a * sn + b * cs;
      const double cc = c * cs + d * sn, dd = -c * sn + d * cs;
      a = aa * cs + cc * sn;
      b = bb * cs + dd * sn;
      c = -aa * sn + cc * cs;
      d = -bb * sn + dd * cs;
      const double tmp = 0.5 * (a + d);
      a = tmp;
      d = tmp;
      if (c != 0.0) {
        if (b != 0.0) {
          if (std::signbit(b) == std::signbit(c)) {
            const double sab = std::sqrt(std::fabs(b)), sac = std::sqrt(std::fabs(c));
            p = fsign(sab * sac, c);
            const double tau2 = 1.0 / std::sqrt(std::fabs(b + c));
            a = tmp + p;
            d = tmp - p;
            b = b - c;
            c = 0.0;
            const double cs1 = sab * tau2, sn1 = sac * tau2;
            const double t2 = cs * cs1 - sn * sn1;
            sn = cs * sn1 + sn * cs1;
            cs = t2;
          }
        } else {
          b = -c;
          c = 0.0;
          const double t2 = cs;
          cs = -sn;
          sn = t2;
        }
      }
    }
  }
}

}  // namespace

bool dgeev3(const double m[3][3], double wr[3], double vr[3][3]) {
  // 1-based copies, as in the Fortran
  double H[4][4] = {}, Z[4][4] = {};
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) {
      H[r + 1][c + 1] = m[r][c];
      if (r != c && m[r][c] == 0.0) return false;  // DGEBAL could permute
      if (!std::isfinite(m[r][c])) return false;
    }
  // DGEBAL's scaling sweep (ILO = 1, IHI = 3: no zero off-diagonal entry, so
  // no permutation).  The helper proceeds only when the first sweep scales
  // no index: then DGEBAL's NOCONV stays false and the matrix is unbalanced.
  {
    const double radix = 2.0, factor = 0.95;
    const double sfmin1 = DBL_MIN / (DBL_EPSILON * 0.5 * radix), sfmin2 = sfmin1 * radix, sfmax2 = 1.0 / sfmin2;
    for (int i = 1; i <= 3; ++i) {
      double c = 0.0, r = 0.0, ca = 0.0, ra = 0.0;
      for (int j = 1; j <= 3; ++j) {       // DNRM2 of column i and row i, IDAMAX of each
        c += H[j][i] * H[j][i];
        r += H[i][j] * H[i][j];
        ca = std::max(ca, std::fabs(H[j][i]));
        ra = std::max(ra, std::fabs(H[i][j]));
      }
      c = std::sqrt(c);
      r = std::sqrt(r);
      if (c == 0.0 || r == 0.0) continue;
      double g = r / radix, f = 1.0;
      const double s = c + r;
      while (c < g && std::max(f, std::max(c, ca)) < sfmax2 && std::min(r, std::min(g, ra)) > sfmin2) {
        f *= radix;
        c *= radix;
        ca *= radix;
        r /= radix;
        g /= radix;
        ra /= radix;
      }
      g = c / radix;
      while (g >= r && std::max(r, ra) < sfmax2 && std::min(std::min(f, c), std::min(g, ca)) > sfmin2) {
        f /= radix;
        c /= radix;
        g /= radix;
        ca /= radix;
        r *= radix;
        ra *= radix;
      }
      if (c + r >= factor * s) continue;   // not worth scaling
      if (f != 1.0) return false;          // DGEBAL would scale index i
    }
  }
  // DGEHD2, i = 1 (i = 2 gives tau = 0)
  double alpha = H[2][1], x = H[3][1];
  const double tau = larfg(2, alpha, &x);
  const double v2 = x;
  // H := H (I - tau v v^T) on columns 2..3, rows 1..3 (DLARF 'Right')
  for (int r = 1; r <= 3; ++r) {
    const double w = H[r][2] + H[r][3] * v2;
    H[r][2] -= tau * w;
    H[r][3] -= tau * w * v2;
  }
  // H := (I - tau v v^T) H on rows 2..3, columns 2..3 (DLARF 'Left')
  for (int c = 2; c <= 3; ++c) {
    const double w = H[2][c] + v2 * H[3][c];
    H[2][c] -= tau * w;
    H[3][c] -= tau * v2 * w;
  }
  H[2][1] = alpha;
  H[3][1] = 0.0;  // DLAHQR clears H(ihi, ihi-2)
  // DORGHR / DORG2R
  Z[1][1] = 1.0;
  Z[2][2] = 1.0 - tau;
  Z[3][2] = -tau * v2;
  Z[2][3] = -tau * v2;
  Z[3][3] = 1.0 - tau * v2 * v2;

  // DLAHQR, WANTT = WANTZ = true, ILO = 1, IHI = N = 3
  const int n = 3, ilo = 1, ihi = 3, i1 = 1, i2 = 3;
  const double smlnum = kSafmin * (3.0 / kUlp);
  const int itmax = 30 * 10;
  const double dat1 = 0.75, dat2 = -0.4375;
  const int kexsh = 10;
  int kdefl = 0;
  int I = ihi;
  double WR[4] = {};
  while (I >= ilo) {
    int L = ilo;
    bool converged = false;
    for (int its = 0; its <= itmax; ++its) {
      int K;
      for (K = I; K >= L + 1; --K) {
        if (std::fabs(H[K][K - 1]) <= smlnum) break;
        double tst = std::fabs(H[K - 1][K - 1]) + std::fabs(H[K][K]);
        if (tst == 0.0) {
          if (K - 2 >= ilo) tst += std::fabs(H[K - 1][K - 2]);
          if (K + 1 <= ihi) tst += std::fabs(H[K + 1][K]);
        }
        if (std::fabs(H[K][K - 1]) <= kUlp * tst) {
          const double ab = std::fmax(std::fabs(H[K][K - 1]), std::fabs(H[K - 1][K]));
          const double ba = std::fmin(std::fabs(H[K][K - 1]), std::fabs(H[K - 1][K]));
          const double aa = std::fmax(std::fabs(H[K][K]), std::fabs(H[K - 1][K - 1] - H[K][K]));
          const double bb = std::fmin(std::fabs(H[K][K]), std::fabs(H[K - 1][K - 1] - H[K][K]));
          const double s = aa + ab;
          if (ba * (ab / s) <= std::fmax(smlnum, kUlp * (bb * (aa / s)))) break;
        }
      }
      L = K;
      if (L > ilo) H[L][L - 1] = 0.0;
      if (L >= I - 1) {
        converged = true;
        break;
      }
      ++kdefl;
      double h11, h12, h21, h22;
      if (kdefl % (2 * kexsh) == 0) {
        const double s = std::fabs(H[I][I - 1]) + std::fabs(H[I - 1][I - 2]);
        h11 = dat1 * s + H[I][I];
        h12 = dat2 * s;
        h21 = s;
        h22 = h11;
      } else if (kdefl % kexsh == 0) {
        const double s = std::fabs(H[L + 1][L]) + std::fabs(H[L + 2][L + 1]);
        h11 = dat1 * s + H[L][L];
        h12 = dat2 * s;
        h21 = s;
        h22 = h11;
      } else {
        h11 = H[I - 1][I - 1];
        h21 = H[I][I - 1];
        h12 = H[I - 1][I];
        h22 = H[I][I];
      }
      double rt1r, rt1i, rt2r, rt2i;
      {
        const double s = std::fabs(h11) + std::fabs(h12) + std::fabs(h21) + std::fabs(h22);
        if (s == 0.0) {
          rt1r = rt1i = rt2r = rt2i = 0.0;
        } else {
          h11 /= s;
          h21 /= s;
          h12 /= s;
          h22 /= s;
          const double tr = (h11 + h22) / 2.0;
          const double det = (h11 - tr) * (h22 - tr) - h12 * h21;
          const double rtdisc = std::sqrt(std::fabs(det));
          if (det >= 0.0) {
            rt1r = tr * s;
            rt2r = rt1r;
            rt1i = rtdisc * s;
            rt2i = -rt1i;
          } else {
            rt1r = tr + rtdisc;
            rt2r = tr - rtdisc;
            if (std::fabs(rt1r - h22) <= std::fabs(rt2r - h22)) {
              rt1r *= s;
              rt2r = rt1r;
            } else {
              rt2r *= s;
              rt1r = rt2r;
            }
            rt1i = rt2i = 0.0;
          }
        }
      }
      double v[3];
      int M;
      for (M = I - 2; M >= L; --M) {
        double h21s = H[M + 1][M];
        double s = std::fabs(H[M][M] - rt2r) + std::fabs(rt2i) + std::fabs(h21s);
        h21s = H[M + 1][M] / s;
        v[0] = h21s * H[M][M + 1] + (H[M][M] - rt1r) * ((H[M][M] - rt2r) / s) - rt1i * (rt2i / s);
        v[1] = h21s * (H[M][M] + H[M + 1][M + 1] - rt1r - rt2r);
        v[2] = h21s * H[M + 2][M + 1];
        s = std::fabs(v[0]) + std::fabs(v[1]) + std::fabs(v[2]);
        v[0] /= s;
        v[1] /= s;
        v[2] /= s;
        if (M == L) break;
        const double h00 = std::fabs(H[M][M - 1]) * (std::fabs(v[1]) + std::fabs(v[2]));
        const double h01 = std::fabs(v[0]) * (std::fabs(H[M - 1][M - 1]) + std::fabs(H[M][M]) + std::fabs(H[M + 1][M + 1]));
        if (h00 <= kUlp * h01) break;
      }
      for (int Kk = M; Kk <= I - 1; ++Kk) {
        const int nr = (3 < I - Kk + 1) ? 3 : I - Kk + 1;
        if (Kk > M)
          for (int q = 0; q < nr; ++q) v[q] = H[Kk + q][Kk - 1];
        double a0 = v[0];
        const double t1 = larfg(nr, a0, v + 1);
        v[0] = a0;
        if (Kk > M) {
          H[Kk][Kk - 1] = v[0];
          H[Kk + 1][Kk - 1] = 0.0;
          if (Kk < I - 1) H[Kk + 2][Kk - 1] = 0.0;
        } else if (M > L) {
          H[Kk][Kk - 1] *= (1.0 - t1);
        }
        const double vv2 = v[1], t2 = t1 * vv2;
        if (nr == 3) {
          const double vv3 = v[2], t3 = t1 * vv3;
          for (int j = Kk; j <= i2; ++j) {
            const double sum = H[Kk][j] + vv2 * H[Kk + 1][j] + vv3 * H[Kk + 2][j];
            H[Kk][j] -= sum * t1;
            H[Kk + 1][j] -= sum * t2;
            H[Kk + 2][j] -= sum * t3;
          }
          const int jmax = (Kk + 3 < I) ? Kk + 3 : I;
          for (int j = i1; j <= jmax; ++j) {
            const double sum = H[j][Kk] + vv2 * H[j][Kk + 1] + vv3 * H[j][Kk + 2];
            H[j][Kk] -= sum * t1;
            H[j][Kk + 1] -= sum * t2;
            H[j][Kk + 2] -= sum * t3;
          }
          for (int j = 1; j <= n; ++j) {
            const double sum = Z[j][Kk] + vv2 * Z[j][Kk + 1] + vv3 * Z[j][Kk + 2];
            Z[j][Kk] -= sum * t1;
            Z[j][Kk + 1] -= sum * t2;
            Z[j][Kk + 2] -= sum * t3;
          }
        } else if (nr == 2) {
          for (int j = Kk; j <= i2; ++j) {
            const double sum = H[Kk][j] + vv2 * H[Kk + 1][j];
            H[Kk][j] -= sum * t1;
            H[Kk + 1][j] -= sum * t2;
          }
          for (int j = i1; j <= I; ++j) {
            const double sum = H[j][Kk] + vv2 * H[j][Kk + 1];
            H[j][Kk] -= sum * t1;
            H[j][Kk + 1] -= sum * t2;
          }
          for (int j = 1; j <= n; ++j) {
            const double sum = Z[j][Kk] + vv2 * Z[j][Kk + 1];
            Z[j][Kk] -= sum * t1;
            Z[j][Kk + 1] -= sum * t2;
          }
        }
      }
    }
    if (!converged) return false;
    if (L == I) {
      WR[I] = H[I][I];
    } else if (L == I - 1) {
      double cs, sn;
      lanv2(H[I - 1][I - 1], H[I - 1][I], H[I][I - 1], H[I][I], cs, sn);
      WR[I - 1] = H[I - 1][I - 1];
      WR[I] = H[I][I];
      if (H[I][I - 1] != 0.0) return false;  // complex pair: not handled here
      // DROT on rows I-1, I (columns I+1..I2), columns I-1, I (rows I1..I-2), and Z
      for (int j = I + 1; j <= i2; ++j) {
        const double xa = H[I - 1][j], ya = H[I][j];
        H[I - 1][j] = cs * xa + sn * ya;
        H[I][j] = cs * ya - sn * xa;
      }
      for (int j = i1; j <= I - 2; ++j) {
        const double xa = H[j][I - 1], ya = H[j][I];
        H[j][I - 1] = cs * xa + sn * ya;
        H[j][I] = cs * ya - sn * xa;
      }
      for (int j = 1; j <= n; ++j) {
        const double xa = Z[j][I - 1], ya = Z[j][I];
        Z[j][I - 1] = cs * xa + sn * ya;
        Z[j][I] = cs * ya - sn * xa;
      }
    }
    kdefl = 0;
    I = L - 1;
  }

  // DTREVC3 (right vectors, back transformed), then DGEEV's 2-norm scaling
  double VR[4][4];
  for (int r = 1; r <= 3; ++r)
    for (int c = 1; c <= 3; ++c) VR[r][c] = Z[r][c];
  const double unfl = kSafmin;
  const double smlt = unfl * (n / kUlp);
  for (int ki = n; ki >= 1; --ki) {
    const double wr0 = H[ki][ki];
    const double smin = std::fmax(kUlp * std::fabs(wr0), smlt);
    double w[4] = {};
    w[ki] = 1.0;
    for (int k = 1; k <= ki - 1; ++k) w[k] = -H[k][ki];
    for (int j = ki - 1; j >= 1; --j) {
      double den = H[j][j] - wr0;
      if (std::fabs(den) < smin) den = smin;
      const double xj = w[j] / den;
      w[j] = xj;
      for (int k = 1; k <= j - 1; ++k) w[k] -= xj * H[k][j];
    }
    double col[4];
    for (int r = 1; r <= n; ++r) {
      double s = w[ki] * VR[r][ki];
      for (int k = 1; k <= ki - 1; ++k) s += VR[r][k] * w[k];
      col[r] = s;
    }
    int ii = 1;
    for (int r = 2; r <= n; ++r)
      if (std::fabs(col[r]) > std::fabs(col[ii])) ii = r;
    const double remax = 1.0 / std::fabs(col[ii]);
    for (int r = 1; r <= n; ++r) VR[r][ki] = col[r] * remax;
  }
  for (int c = 1; c <= 3; ++c) {
    double ss = 0.0;
    for (int r = 1; r <= 3; ++r) ss += VR[r][c] * VR[r][c];
    const double scl = 1.0 / std::sqrt(ss);
    for (int r = 1; r <= 3; ++r) vr[r - 1][c - 1] = VR[r][c] * scl;
    wr[c - 1] = WR[c];
  }
  return true;
}

}  // namespace crava
