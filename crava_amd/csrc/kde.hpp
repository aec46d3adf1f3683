// Facies density tables from well samples, host side:
// PosteriorElasticPDF3D's histogram and Gaussian smoother
// (src/posteriorelasticpdf3d.cpp:12-138,360-413) and the per-density-index
// setup of FaciesProb::MakePosteriorElasticPDF3D (src/faciesprob.cpp:926-1135,
// CalculateVariances :2332-2377).  The convolution runs on the GPU with the
// library's transforms (capi.hip).
#pragma once

namespace crava {

constexpr int kIMissing = -99999;   // IMISSING, src/definitions.h

// PosteriorElasticPDF::InvertSquareMatrix (src/posteriorelasticpdf.cpp:40-53):
// lib_matrCholR then lib_matrAXeqBMatR on the identity.  False when the
// factorisation fails (a diagonal at or below 1e-20).
bool invert_spd3(const double s[3][3], double inv[3][3]);

// The histogram of the points on the n1 x n2 x n3 lattice over
// lim = {min1, max1, min2, max2, min3, max3}, each point 1/npts (float), in
// the real layout of FFTGrid(n1, n2, n3) [n3][n2][2(n1/2+1)]; points outside
// are dropped.
void kde_histogram(long long npts, const double *d1, const double *d2, const double *d3, int n1, int n2, int n3,
                   const double lim[6], float *grid);

// SetupSmoothingGaussian3D: exp(-x^T sigma^-1 x / 2) at the wrapped lattice
// lags, normalised by its float sum, same layout.  False when sigma is not
// positive definite.
bool kde_smoother(const double sigma[3][3], int n1, int n2, int n3, const double lim[6], float *grid);

// CalculateVariances over the samples with a facies: var[3] of a, b, r.
void facies_sample_variances(int n, const float *a, const float *b, const float *r, const int *facies, float *var);

// One density index of MakePosteriorElasticPDF3D: the smoothing covariance
// (sigma_e [9] row-major upper triangle used, NULL for zero; no_vs: its 2x2
// alpha/rho block in the top-left 2x2) with each diagonal at least
// (hopt std)^2, and the lattice limits min - 5 sd, max + 5 sd.
void facies_density_setup(int n, const float *a, const float *b, const float *r, const int *facies,
                          const double *sigma_e, bool no_vs, int n1, int n2, int n3, double sigma[3][3],
                          double lim[6]);

}  // namespace crava
