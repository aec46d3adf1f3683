// Facies probabilities from the posterior elastic PDF
// (FaciesProb::CalculateFaciesProbFromPosteriorElasticPDF,
// src/faciesprob.cpp:2198-2265; FindDensityFromPosteriorPDF :1277-1317;
// PosteriorElasticPDF3D::FindDensity, src/posteriorelasticpdf3d.cpp:511-604).
//
// One lane per output voxel.  The roundings of the reference are kept bit for
// bit: double interpolation indexes, float weights and products, a double
// accumulator of the eight float terms, float normalisation.  This file is
// compiled with -ffp-contract=off so that no product is fused into an add.
#include "../../include/crava_hip.h"
#include "common.hpp"
#include "facies.hpp"

namespace crava {

constexpr float kRMissing = -99999.0f;  // RMISSING, src/definitions.h:35

struct FaciesArgs {
  const float *alpha, *beta, *rho;
  const float *tables;
  float *prob[kMaxFacies + 1];
  int *argmax;
  unsigned long long *counts;
  long long tsize;      // floats per table
  int nx, ny, nz, rnxp, nyp, smallrnxp;
  int tn1r, tn2;        // table row pitch (2*(n1/2+1)) and n2
  FaciesVolume v0;      // the lattice of density 0
  float prior[kMaxFacies];
  float undef_sum;
  int nfac;
  int ndens, nang;
  const FaciesVolume *vols;
  const float *tgrid;
};

__device__ __forceinline__ float tv(const float *t, int rn, int n2, int a, int b, int c) {
  const float v = t[a + (long long)rn * b + (long long)rn * n2 * c];
  return v > 0.0f ? v : 0.0f;  // std::max<float>(0, v)
}

__device__ __forceinline__ double find_density(const FaciesArgs &f, const FaciesVolume &v, const float *t, double vp,
                                               double vs, double rho) {
  double value = 0.0;
  const double rx = (vp - v.x0) * 1.0 + (vs - v.y0) * 0.0;
  const double ry = -(vp - v.x0) * 0.0 + (vs - v.y0) * 1.0;
  const double jF = rx / v.dx - 0.5;
  const double kF = ry / v.dy - 0.5;
  const double lF = (double)(v.n2) * (rho - v.ztop) / (v.zbot - v.ztop) - 0.5;
  int j1, k1, l1, j2, k2, l2;
  float wj, wk, wl;
  j1 = (int)floor(jF);
  if (j1 < 0) { j1 = 0; j2 = 0; wj = 0; }
  else if (j1 >= v.n0 - 1) { j1 = v.n0 - 1; j2 = j1; wj = 0; }
  else { j2 = j1 + 1; wj = (float)(jF - j1); }
  k1 = (int)floor(kF);
  if (k1 < 0) { k1 = 0; k2 = 0; wk = 0; }
  else if (k1 >= v.n1 - 1) { k1 = v.n1 - 1; k2 = k1; wk = 0; }
  else { k2 = k1 + 1; wk = (float)(kF - k1); }
  l1 = (int)floor(lF);
  if (l1 < 0) { l1 = 0; l2 = 0; wl = 0; }
  else if (l1 >= v.n2 - 1) { l1 = v.n2 - 1; l2 = l1; wl = 0; }
  else { l2 = l1 + 1; wl = (float)(lF - l1); }
  const int rn = f.tn1r, n2 = f.tn2;
  const float v1 = tv(t, rn, n2, j1, k1, l1), v2 = tv(t, rn, n2, j1, k1, l2);
  const float v3 = tv(t, rn, n2, j1, k2, l1), v4 = tv(t, rn, n2, j1, k2, l2);
  const float v5 = tv(t, rn, n2, j2, k1, l1), v6 = tv(t, rn, n2, j2, k1, l2);
  const float v7 = tv(t, rn, n2, j2, k2, l1), v8 = tv(t, rn, n2, j2, k2, l2);
  value += (1.0f - wj) * (1.0f - wk) * (1.0f - wl) * v1;
  value += (1.0f - wj) * (1.0f - wk) * (wl) * v2;
  value += (1.0f - wj) * (wk) * (1.0f - wl) * v3;
  value += (1.0f - wj) * (wk) * (wl) * v4;
  value += (wj) * (1.0f - wk) * (1.0f - wl) * v5;
  value += (wj) * (1.0f - wk) * (wl) * v6;
  value += (wj) * (wk) * (1.0f - wl) * v7;
  value += (wj) * (wk) * (wl) * v8;
  return value;
}

__global__ void __launch_bounds__(256) k_facies(FaciesArgs f) {
  __shared__ unsigned long long cnt[kMaxFacies];
  if (f.counts) {
    for (int l = threadIdx.x; l < f.nfac; l += blockDim.x) cnt[l] = 0;
    __syncthreads();
  }
  const long long nout = (long long)f.nz * f.ny * f.smallrnxp;
  const long long o = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  if (o < nout) {
    const int k = (int)(o % f.smallrnxp);      // x index (the reference's k)
    const long long ij = o / f.smallrnxp;
    const int j = (int)(ij % f.ny);
    const int i = (int)(ij / f.ny);             // z index (the reference's i)
    const long long in = ((long long)i * f.nyp + j) * f.rnxp + k;
    if (k < f.nx) {
      const double a = f.alpha[in], b = f.beta[in], r = f.rho[in];
      float value[kMaxFacies];
      float sum = f.undef_sum;
      for (int l = 0; l < f.nfac; ++l) {
        float vs;
        if (f.ndens == 1) {
          const float v = (float)find_density(f, f.v0, f.tables + l * f.tsize, a, b, r);
          vs = 0.0f + v;
        } else {
          // FindDensityFromPosteriorPDF (src/faciesprob.cpp:1293-1312): the
          // density of set d weighted by t_j or 1 - t_j for bit j of d
          float tt[kMaxDensAngles];
          for (int g = 0; g < f.nang; ++g) tt[g] = f.tgrid[((long long)g * f.ny + j) * f.nx + k];
          vs = 0.0f;
          for (int d = 0; d < f.ndens; ++d) {
            float v = (float)find_density(f, f.vols[d], f.tables + ((long long)d * f.nfac + l) * f.tsize, a, b, r);
            int factor = 1;
            for (int g = 0; g < f.nang; ++g) {
              if (g > 0) factor *= 2;
              if ((d & factor) > 0) v *= tt[g];
              else v *= (1 - tt[g]);
            }
            vs += v;
          }
        }
        const float dens = vs > 0.0f ? vs : 0.0f;
        value[l] = f.prior[l] * dens;
        sum = sum + value[l];
      }
      int best = 0;
      float bestp = 0.0f;
      for (int l = 0; l < f.nfac; ++l) {
        const float h = value[l] / sum;
        f.prob[l][o] = h;
        if (l == 0 || h > bestp) { bestp = h; best = l; }
      }
      f.prob[f.nfac][o] = f.undef_sum / sum;
      if (f.argmax) f.argmax[((long long)i * f.ny + j) * f.nx + k] = best;
      if (f.counts) atomicAdd(&cnt[best], 1ULL);
    } else {
      for (int l = 0; l <= f.nfac; ++l) f.prob[l][o] = kRMissing;
    }
  }
  if (f.counts) {
    __syncthreads();
    for (int l = threadIdx.x; l < f.nfac; l += blockDim.x)
      if (cnt[l]) atomicAdd(&f.counts[l], cnt[l]);
  }
}

int launch_facies(const FaciesCall &c, hipStream_t s) {
  if (c.nfac < 1 || c.nfac > kMaxFacies) {
    set_error("number of facies must be 1.." + std::to_string(kMaxFacies));
    return CRAVA_HIP_E_INVALID;
  }
  FaciesArgs f{};
  f.alpha = c.alpha;
  f.beta = c.beta;
  f.rho = c.rho;
  f.tables = c.tables;
  for (int l = 0; l <= c.nfac; ++l) f.prob[l] = c.prob[l];
  f.argmax = c.argmax;
  f.counts = c.counts;
  f.tn1r = 2 * (c.tn1 / 2 + 1);
  f.tn2 = c.tn2;
  f.tsize = (long long)f.tn1r * c.tn2 * c.tn3;
  f.nx = c.nx; f.ny = c.ny; f.nz = c.nz;
  f.rnxp = 2 * (c.nxp / 2 + 1);
  f.nyp = c.nyp;
  f.smallrnxp = 2 * (c.nx / 2 + 1);
  f.v0.n0 = c.vn[0]; f.v0.n1 = c.vn[1]; f.v0.n2 = c.vn[2];
  f.v0.x0 = c.vol[0]; f.v0.y0 = c.vol[1]; f.v0.dx = c.vol[2]; f.v0.dy = c.vol[3];
  f.v0.ztop = c.vol[4]; f.v0.zbot = c.vol[5];
  if (c.ndens < 1 || c.ndens != (1 << c.nang) || c.nang > kMaxDensAngles ||
      (c.ndens > 1 && (!c.vols || !c.tgrid))) {
    set_error("facies densities: ndens must be 2^nang (nang <= 8) with volumes and angle weights");
    return CRAVA_HIP_E_INVALID;
  }
  f.ndens = c.ndens;
  f.nang = c.nang;
  f.vols = c.vols;
  f.tgrid = c.tgrid;
  for (int l = 0; l < c.nfac; ++l) f.prior[l] = c.prior[l];
  // float / int as the reference (src/faciesprob.cpp:2183-2185)
  f.undef_sum = c.p_undefined / (c.vn[0] * c.vn[1] * c.vn[2]);
  f.nfac = c.nfac;
  const long long nout = (long long)c.nz * c.ny * f.smallrnxp;
  hipLaunchKernelGGL(k_facies, dim3((unsigned)((nout + 255) / 256)), dim3(256), 0, s, f);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

}  // namespace crava
