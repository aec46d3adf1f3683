// Host side of the facies density tables (see kde.hpp).
#include "kde.hpp"

#include <algorithm>
#include <cmath>

namespace crava {

bool invert_spd3(const double s[3][3], double inv[3][3]) {
  // lib_matrCholR (libs/lib/lib_matr.c:21-76): L in the lower triangle
  double l[3][3];
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) l[r][c] = s[r][c];
  const double tol = 1e-20;   // lib_matr.c:20
  for (int i = 0; i < 3; ++i) {
    if (l[i][i] <= tol) return false;
    for (int j = 0; j < i; ++j) {
      double acc = 0.0;
      for (int k = 0; k < j; ++k) acc += l[i][k] * l[j][k];
      l[i][j] = (l[i][j] - acc) / l[j][j];
    }
    double acc = 0.0;
    for (int k = 0; k < i; ++k) acc += l[i][k] * l[i][k];
    acc = l[i][i] - acc;
    if (acc <= tol) return false;
    l[i][i] = std::sqrt(acc);
  }
  // lib_matrAXeqBMatR (:280-306) on the identity, column by column, with its
  // (x L_ii) / L_ii^2 divisions
  for (int r = 0; r < 3; ++r)
    for (int c = 0; c < 3; ++c) inv[r][c] = r == c ? 1.0 : 0.0;
  for (int c = 0; c < 3; ++c) {
    for (int i = 0; i < 3; ++i) {
      double x = inv[i][c];
      for (int j = 0; j < i; ++j) x -= inv[j][c] * l[i][j];
      const double h = l[i][i] * l[i][i];
      inv[i][c] = (x * l[i][i]) / h;
    }
    for (int i = 2; i >= 0; --i) {
      double x = inv[i][c];
      for (int j = 2; j > i; --j) x = x - inv[j][c] * l[j][i];
      const double h = l[i][i] * l[i][i];
      inv[i][c] = (x * l[i][i]) / h;
    }
  }
  return true;
}

void kde_histogram(long long npts, const double *d1, const double *d2, const double *d3, int n1, int n2, int n3,
                   const double lim[6], float *grid) {
  const int rn1 = 2 * (n1 / 2 + 1);
  const long long rsize = (long long)rn1 * n2 * n3;
  std::fill(grid, grid + rsize, 0.0f);
  const double dx = (lim[1] - lim[0]) / n1, dy = (lim[3] - lim[2]) / n2, dz = (lim[5] - lim[4]) / n3;
  for (long long p = 0; p < npts; ++p) {
    const double fi = std::floor((d1[p] - lim[0]) / dx), fj = std::floor((d2[p] - lim[2]) / dy),
                 fk = std::floor((d3[p] - lim[4]) / dz);
    // FFTGrid::setRealValue ignores a point outside the grid
    if (!(fi >= 0.0 && fi < n1 && fj >= 0.0 && fj < n2 && fk >= 0.0 && fk < n3)) continue;
    const long long o = (long long)fi + (long long)rn1 * ((long long)fj + (long long)n2 * (long long)fk);
    grid[o] = grid[o] + 1.0f;
  }
  // multiplyByScalar(float(1.0f / dim)) over the whole real grid
  const float sc = float(1.0f / (float)npts);
  for (long long o = 0; o < rsize; ++o) grid[o] *= sc;
}

bool kde_smoother(const double sigma[3][3], int n1, int n2, int n3, const double lim[6], float *grid) {
  double si[3][3];
  if (!invert_spd3(sigma, si)) return false;
  const double dx = (lim[1] - lim[0]) / n1, dy = (lim[3] - lim[2]) / n2, dz = (lim[5] - lim[4]) / n3;
  const int rn1 = 2 * (n1 / 2 + 1);
  std::fill(grid, grid + (long long)rn1 * n2 * n3, 0.0f);
  // lags in cells: j for j <= n/2, j - n after (the wrap of :367-388)
  float sum = 0.0f;
  for (int l = 0; l < n3; ++l) {
    const int ll = l <= n3 / 2 ? l : l - n3;
    for (int k = 0; k < n2; ++k) {
      const int kk = k <= n2 / 2 ? k : k - n2;
      for (int j = 0; j < n1; ++j) {
        const int jj = j <= n1 / 2 ? j : j - n1;
        const double q = jj * dx * jj * dx * si[0][0] + kk * dy * kk * dy * si[1][1] + ll * dz * ll * dz * si[2][2] +
                         2 * jj * dx * kk * dy * si[1][0] + 2 * jj * dx * ll * dz * si[2][0] +
                         2 * kk * dy * ll * dz * si[2][1];
        const float v = float(std::exp(-0.5f * q));
        grid[j + (long long)rn1 * (k + (long long)n2 * l)] = v;
        sum += v;
      }
    }
  }
  for (int l = 0; l < n3; ++l)
    for (int k = 0; k < n2; ++k)
      for (int j = 0; j < n1; ++j) grid[j + (long long)rn1 * (k + (long long)n2 * l)] /= sum;
  return true;
}

void facies_sample_variances(int n, const float *a, const float *b, const float *r, const int *facies, float *var) {
  int cnt = 0;
  float ma = 0.0f, mb = 0.0f, mr = 0.0f;
  for (int i = 0; i < n; ++i)
    if (facies[i] != kIMissing) {
      ma += a[i];
      mb += b[i];
      mr += r[i];
      ++cnt;
    }
  ma /= cnt;
  mb /= cnt;
  mr /= cnt;
  float va = 0.0f, vb = 0.0f, vr = 0.0f;
  for (int i = 0; i < n; ++i)
    if (facies[i] != kIMissing) {
      // pow(float, 2) is a double square; the float accumulator rounds each sum
      va = (float)((double)va + std::pow((double)(a[i] - ma), 2));
      vb = (float)((double)vb + std::pow((double)(b[i] - mb), 2));
      vr = (float)((double)vr + std::pow((double)(r[i] - mr), 2));
    }
  var[0] = va / (cnt - 1);
  var[1] = vb / (cnt - 1);
  var[2] = vr / (cnt - 1);
}

void facies_density_setup(int n, const float *a, const float *b, const float *r, const int *facies,
                          const double *sigma_e, bool no_vs, int n1, int n2, int n3, double sigma[3][3],
                          double lim[6]) {
  int nobs = 0;
  for (int i = 0; i < n; ++i)
    if (facies[i] != kIMissing) ++nobs;
  float var[3];
  facies_sample_variances(n, a, b, r, facies, var);
  if (no_vs) var[1] = 5 * var[0];
  const float hopt = (float)(std::pow(4.0 / 7, 1.0 / 7) * std::pow((double)nobs, -1.0 / 7));
  const float kstd[3] = {hopt * std::sqrt(var[0]), hopt * std::sqrt(var[1]), hopt * std::sqrt(var[2])};
  double se[3][3] = {{0, 0, 0}, {0, 0, 0}, {0, 0, 0}};
  if (sigma_e) {
    if (!no_vs) {
      for (int j = 0; j < 3; ++j)
        for (int k = 0; k <= j; ++k) se[k][j] = se[j][k] = sigma_e[k * 3 + j];
    } else {
      // the 2x2 (alpha, rho) matrix in the corners of sigma_e's upper triangle
      se[0][0] = sigma_e[0];
      se[0][2] = se[2][0] = sigma_e[1];
      se[2][2] = sigma_e[4];
    }
  }
  for (int d = 0; d < 3; ++d)
    if (se[d][d] < kstd[d] * kstd[d]) se[d][d] = kstd[d] * kstd[d];
  for (int j = 0; j < 3; ++j)
    for (int k = 0; k < 3; ++k) sigma[j][k] = se[j][k];
  const float *v[3] = {a, b, r};
  for (int d = 0; d < 3; ++d) {
    const float mn = *std::min_element(v[d], v[d] + n), mx = *std::max_element(v[d], v[d] + n);
    lim[2 * d] = mn - 5.0f * std::sqrt(se[d][d]);
    lim[2 * d + 1] = mx + 5.0f * std::sqrt(se[d][d]);
  }
  (void)n1;
  (void)n2;
  (void)n3;
}

}  // namespace crava
