// Forward modelling: synthetic seismic from the elastic parameters
// (Crava::computeSyntSeismic, src/crava.cpp:1667-1759, with
// Crava::computeSeismicImpedance, :1632-1663).
//
// Per angle l and trace (i, j) inside the volume (i < nx, j < ny):
//   imp(k) = ((0 + A_l0 alpha) + A_l1 beta) + A_l2 rho              k < nz
//   imp(k) = fac ((k - nz) imp(0) + (nzp - k - 1) imp(nz - 1))     k >= nz,
//            fac = 1 / (nzp - nz - 1)                               (taper, :1692-1696)
//   r(k)   = imp(k + 1) - imp(k), cyclic      (Wavelet1D FIRSTORDERFORWARDDIFF,
//                                              src/wavelet1D.cpp:431-436)
//   s(k)   = R(k) conj(w(k)),  R = r2c(r),  w = the trace's local wavelet
//            spectrum: gain exp(-2 pi i shift/dz k/nzp) W(k)   (Wavelet1D::shiftAndScale,
//            src/wavelet1D.cpp:539-561) read through the stretch sf
//            (Wavelet::getCAmp(k, scale), src/wavelet.cpp:328-384)
//   out    = c2r(s) / nzp                     (Wavelet::invFFT1DInPlace, src/wavelet.cpp:264-280)
// Traces outside the volume (the padding) keep imp(k) for every k, as the
// impedance grid the reference writes into does.
//
// GPU form.  The traces at x = 2m and 2m+1 are the real and imaginary parts
// of complex column m of the grid's half-complex view, so the z transforms
// are the z passes of the grid's transform plan over all angles' grids in
// one launch each.  Between them, k_fm_filter separates the two traces'
// spectra from Z(k) and Z(N-k), filters each with its own wavelet and packs
// them again.  Three grid kernels: k_fm_reflect (impedance, taper,
// difference), k_fm_filter, k_fm_padding (exact impedance of the padding
// traces after the transforms).  Built with -ffp-contract=off: the
// impedance, taper and difference follow the reference's float operations.
#include "../../include/crava_hip.h"
#include "forward.hpp"

#include <cmath>

namespace crava {

namespace {

__device__ __forceinline__ float impedance(const float (&A)[3], float a, float b, float c) {
  float v = 0.0f;
  v += a * A[0];
  v += b * A[1];
  v += c * A[2];
  return v;
}

}  // namespace

// One thread per trace (i, j), i < rnxp, j < nyp; loops over k.
template <int NT>
__global__ void __launch_bounds__(256) k_fm_reflect(FwdArgs a) {
  const long long plane = (long long)a.nyp * a.rnxp;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= plane) return;
  const int i = (int)(t % a.rnxp), j = (int)(t / a.rnxp);
  if (i >= a.nx || j >= a.ny) {
    // padding trace: zero through the transforms (its partner trace in the
    // same complex column must not see anything else), the impedance is
    // written afterwards by k_fm_padding
    for (int k = 0; k < a.nzp; ++k) {
#pragma unroll
      for (int l = 0; l < NT; ++l) a.out[l][(long long)k * plane + t] = 0.0f;
    }
    return;
  }
  float imp0[NT], prev[NT];
  for (int k = 0; k < a.nz; ++k) {
    const long long o = (long long)k * plane + t;
    const float va = a.alpha[o], vb = a.beta[o], vc = a.rho[o];
#pragma unroll
    for (int l = 0; l < NT; ++l) {
      const float v = impedance(a.A[l], va, vb, vc);
      if (k == 0) imp0[l] = v;
      else a.out[l][o - plane] = v - prev[l];
      prev[l] = v;
    }
  }
  float last[NT];
#pragma unroll
  for (int l = 0; l < NT; ++l) last[l] = prev[l];
  const float fac = 1.0f / (float)(a.nzp - a.nz - 1);
  for (int k = a.nz; k < a.nzp; ++k) {
    const long long o = (long long)k * plane + t;
#pragma unroll
    for (int l = 0; l < NT; ++l) {
      const float v = fac * ((float)(k - a.nz) * imp0[l] + (float)(a.nzp - k - 1) * last[l]);
      a.out[l][o - plane] = v - prev[l];
      prev[l] = v;
    }
  }
  const long long o = (long long)(a.nzp - 1) * plane + t;
#pragma unroll
  for (int l = 0; l < NT; ++l) a.out[l][o] = imp0[l] - prev[l];
}

// One thread per trace; rewrites the padding traces with the impedance.
template <int NT>
__global__ void __launch_bounds__(256) k_fm_padding(FwdArgs a) {
  const long long plane = (long long)a.nyp * a.rnxp;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= plane) return;
  const int i = (int)(t % a.rnxp), j = (int)(t / a.rnxp);
  if (i < a.nx && j < a.ny) return;
  for (int k = 0; k < a.nzp; ++k) {
    const long long o = (long long)k * plane + t;
    const float va = a.alpha[o], vb = a.beta[o], vc = a.rho[o];
#pragma unroll
    for (int l = 0; l < NT; ++l) a.out[l][o] = impedance(a.A[l], va, vb, vc);
  }
}

// Local wavelet spectrum at bin q: gain exp(-2 pi i ishift q / nzp) W(q), the
// multiplier rounded to float from double as Wavelet1D::shiftAndScale does.
__device__ __forceinline__ cx<float> local_spectrum(const cx<float> *W, int q, float ishift, float gain, int nzp) {
  const cx<float> w = W[q];
  float mr, mi;
  if (ishift == 0.0f) {
    mr = (float)((double)gain * 1.0);
    mi = (float)((double)gain * -0.0);
  } else {
    const double ph = 2.0 * (3.14159265358979323846 * (double)ishift * (double)q) / (double)(float)nzp;
    mr = (float)((double)gain * cos(ph));
    mi = (float)((double)gain * sin(-ph));
  }
  return mk(mr * w.x - mi * w.y, mi * w.x + mr * w.y);
}

// Wavelet::getCAmp(k, scale) of the local spectrum, k < nzp/2 + 1.
__device__ __forceinline__ cx<float> stretched(const cx<float> *W, int k, float sf, float ishift, float gain, int nzp) {
  const int cnz = nzp / 2 + 1;
  const float omega = (float)k / sf;
  cx<float> v = mk(0.0f, 0.0f);
  if (!(omega >= (float)cnz)) {
    const int omL = (int)floorf(omega);
    int omU = omL + 1;
    if (omU >= cnz) omU -= 1;
    const float d = omega - (float)omL;
    const cx<float> a = local_spectrum(W, omL, ishift, gain, nzp);
    const cx<float> b = local_spectrum(W, omU, ishift, gain, nzp);
    v.x = a.x * (1.0f - d) + b.x * d;
    v.y = a.y * (1.0f - d) + b.y * d;
  }
  v.x /= sf;
  v.y /= sf;
  return v;
}

// s = r conj(w) in the reference's order (src/crava.cpp:1711-1714)
__device__ __forceinline__ cx<float> filt(cx<float> r, cx<float> w) {
  return mk(r.x * w.x + r.y * w.y, -r.x * w.y + r.y * w.x);
}

// One thread per (complex column q, bin pair k / N-k), k = 0..N/2; angle =
// blockIdx.y.  Column q = j * cn + m carries traces (2m, j) and (2m+1, j).
__global__ void __launch_bounds__(256) k_fm_filter(FwdArgs a) {
  const int cn = a.rnxp / 2;
  const long long ncol = (long long)cn * a.nyp;
  const int nk = a.nzp / 2 + 1;
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  if (id >= ncol * nk) return;
  const long long q = id % ncol;
  const int k = (int)(id / ncol);
  const int l = blockIdx.y;
  const int N = a.nzp;
  const int km = (N - k) % N;
  cx<float> *g = reinterpret_cast<cx<float> *>(a.out[l]);
  const cx<float> Z = g[(long long)k * ncol + q];
  const cx<float> Zm = g[(long long)km * ncol + q];
  // spectra of the two real traces
  cx<float> Ra = mk(0.5f * (Z.x + Zm.x), 0.5f * (Z.y - Zm.y));
  cx<float> Rb = mk(0.5f * (Z.y + Zm.y), 0.5f * (Zm.x - Z.x));
  const int m = (int)(q % cn), j = (int)(q / cn);
  const cx<float> *W = a.W + (long long)l * nk;
  cx<float> S[2];
  const cx<float> R[2] = {Ra, Rb};
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    const int i = 2 * m + h;
    if (i < a.nx && j < a.ny) {
      const long long tr = (long long)j * a.nx + i;
      const float sf = a.sf[l] ? a.sf[l][tr] : 1.0f;
      const float ishift = a.shift[l] ? a.shift[l][tr] / a.dz : 0.0f;
      const float gain = a.gain[l] ? a.gain[l][tr] : 1.0f;
      const cx<float> s = filt(R[h], stretched(W, k, sf, ishift, gain, N));
      S[h] = scl(s, a.inv_n);
    } else {
      S[h] = scl(R[h], a.inv_n);   // padding: unfiltered (rewritten exactly afterwards)
    }
  }
  if (k == km) {
    // DC, and Nyquist of an even length: a c2r reads only the real part
    g[(long long)k * ncol + q] = mk(S[0].x, S[1].x);
  } else {
    g[(long long)k * ncol + q] = mk(S[0].x - S[1].y, S[0].y + S[1].x);
    g[(long long)km * ncol + q] = mk(S[0].x + S[1].y, S[1].x - S[0].y);
  }
}

int launch_fm_reflect(const FwdArgs &a, hipStream_t s) {
  const long long plane = (long long)a.nyp * a.rnxp;
  const unsigned blocks = (unsigned)((plane + 255) / 256);
  switch (a.nt) {
#define CRAVA_FM_CASE(n) \
  case n: hipLaunchKernelGGL(k_fm_reflect<n>, dim3(blocks), dim3(256), 0, s, a); break;
    CRAVA_FM_CASE(1) CRAVA_FM_CASE(2) CRAVA_FM_CASE(3) CRAVA_FM_CASE(4)
    CRAVA_FM_CASE(5) CRAVA_FM_CASE(6) CRAVA_FM_CASE(7) CRAVA_FM_CASE(8)
#undef CRAVA_FM_CASE
    default: set_error("ntheta must be 1..8"); return CRAVA_HIP_E_INVALID;
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_fm_padding(const FwdArgs &a, hipStream_t s) {
  const long long plane = (long long)a.nyp * a.rnxp;
  const unsigned blocks = (unsigned)((plane + 255) / 256);
  switch (a.nt) {
#define CRAVA_FM_CASE(n) \
  case n: hipLaunchKernelGGL(k_fm_padding<n>, dim3(blocks), dim3(256), 0, s, a); break;
    CRAVA_FM_CASE(1) CRAVA_FM_CASE(2) CRAVA_FM_CASE(3) CRAVA_FM_CASE(4)
    CRAVA_FM_CASE(5) CRAVA_FM_CASE(6) CRAVA_FM_CASE(7) CRAVA_FM_CASE(8)
#undef CRAVA_FM_CASE
    default: set_error("ntheta must be 1..8"); return CRAVA_HIP_E_INVALID;
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_fm_filter(const FwdArgs &a, hipStream_t s) {
  const long long ncol = (long long)(a.rnxp / 2) * a.nyp;
  const long long n = ncol * (a.nzp / 2 + 1);
  const unsigned blocks = (unsigned)((n + 255) / 256);
  hipLaunchKernelGGL(k_fm_filter, dim3(blocks, (unsigned)a.nt), dim3(256), 0, s, a);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

}  // namespace crava
