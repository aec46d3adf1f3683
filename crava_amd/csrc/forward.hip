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
// of complex column m of the grid's half-complex view: one complex z
// transform serves two traces, and the filter separates their spectra from
// Z(k) and Z(N-k), filters each with its own wavelet and packs them again.
//   * nzp with a specialised transform: one fused kernel per (column block,
//     angle) -- parameters in, impedance, taper and difference into LDS, the
//     forward Stockham passes, the filter, the inverse passes, seismic out.
//     One read of the three parameter grids and one write per angle.
//   * other lengths: k_fm_reflect (impedance, taper, difference into the
//     output grids), the plan's z passes over all angles in one launch,
//     k_fm_filter, the inverse z passes, k_fm_padding (exact impedance of the
//     padding traces after the transforms).
// The impedance, taper and difference follow the reference's float
// operations (contraction is off in this file's own code).
#include "../../include/crava_hip.h"
#include "fft_static.hpp"
#include "forward.hpp"
#include "stockham.hpp"

#include <cmath>

// The impedance, taper, difference and filter below keep the reference's
// float operations (no contraction into FMA); the Stockham passes included
// above are contracted.
#pragma clang fp contract(off)

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

// Local wavelet parameters of one trace, staged in LDS by the fused kernel.
struct TraceParams {
  float sf, ishift, gain, inside;
};

// Filter of one bin pair (k, N-k) of a complex column carrying traces pa, pb
// (Re, Im): separate the two traces' spectra, filter each trace inside the
// volume with its local wavelet (spectrum W), scale by 1/N and pack again.
// The separation's factor 1/2 and the 1/N are one factor h = 1/(2N) on the
// (local) wavelet spectrum.  Without local wavelets (LOCAL false: no
// stretch, shift or gain grid) the local spectrum is W(k) itself:
// stretched() at sf = 1, shift 0, gain 1 multiplies by 1, adds 0 and divides
// by 1.
template <bool LOCAL>
__device__ __forceinline__ void filter_pair_h(const cx<float> *W, const TraceParams &pa, const TraceParams &pb, int k,
                                              int N, float h, cx<float> &Zk, cx<float> &Zm) {
#pragma clang fp contract(fast)
  const cx<float> Z = Zk, M = Zm;
  const cx<float> R[2] = {mk(Z.x + M.x, Z.y - M.y), mk(Z.y + M.y, M.x - Z.x)};
  const TraceParams *P[2] = {&pa, &pb};
  cx<float> S[2];
#pragma unroll
  for (int q = 0; q < 2; ++q) {
    const TraceParams &p = *P[q];
    if (p.inside != 0.0f) {
      const cx<float> w = scl(LOCAL ? stretched(W, k, p.sf, p.ishift, p.gain, N) : W[k], h);
      S[q] = filt(R[q], w);
    } else {
      S[q] = scl(R[q], h);
    }
  }
  if (k == (N - k) % N) {
    Zk = mk(S[0].x, S[1].x);
  } else {
    Zk = mk(S[0].x - S[1].y, S[0].y + S[1].x);
    Zm = mk(S[0].x + S[1].y, S[1].x - S[0].y);
  }
}

// Fused forward model of TX complex columns (2 TX traces) for one angle, z
// length N specialised, on the pencil-fastest passes
// (stockham.hpp cpass): thread tid keeps complex column c = tid % TX and its
// butterflies j = tid / TX + t (THREADS / TX) through the whole chain, so
// every LDS offset is a constant from the column's base (pencils N | 1
// apart, no padding inside a pencil).  Per angle:
//   1. impedance of rows e < nz, and the taper to nzp beyond, into LDS;
//   2. the cyclic forward difference imp(e + 1) - imp(e) is the first
//      forward pass's input hook (no separate pass);
//   3. forward passes; the filter in place over slots s (bins s and N - s,
//      slot 0 bins 0 and N/2), each thread on its own column's slots;
//   4. inverse passes, the last storing the seismic (padding traces their
//      impedance).
// The impedance, taper and difference are the reference's float operations;
// the filter's two scalings are folded into one (filter_pair_h).  The
// twiddles, the wavelet spectrum and the trace parameters stay in LDS: read
// from global memory inside the chain they cost more than the fourth
// resident workgroup that the LDS they take would allow.
template <int N, int TX, int THREADS, bool LOCAL>
__global__ void __launch_bounds__(THREADS) k_fm_cols(FwdArgs a, const cx<float> *__restrict__ twg) {
  static_assert(N % 2 == 0, "the specialised lengths are even");
  const unsigned W = xcd_work_id(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const unsigned bx = W / (unsigned)a.nt;
  const int l = (int)(W % (unsigned)a.nt);
  const int cn = a.rnxp / 2;
  const long long ncol = (long long)cn * a.nyp;
  const long long col0 = (long long)bx * TX;
  const int ncb = (int)min((long long)TX, ncol - col0);
  constexpr int NK = N / 2 + 1;
  constexpr int PC = N | 1;
  constexpr int JS = THREADS / TX;
  extern __shared__ unsigned char smem_raw[];
  cx<float> *tw = reinterpret_cast<cx<float> *>(smem_raw);
  cx<float> *wsp = tw + N;
  TraceParams *tp = reinterpret_cast<TraceParams *>(wsp + NK);
  cx<float> *buf = reinterpret_cast<cx<float> *>(tp + 2 * TX);
  const float hn = 0.5f * a.inv_n;
  for (int q = threadIdx.x; q < N; q += THREADS) tw[q] = twg[q];
  for (int q = threadIdx.x; q < NK; q += THREADS) wsp[q] = a.W[(long long)l * NK + q];
  if (threadIdx.x < 2 * TX) {
    const int c = threadIdx.x >> 1, h = threadIdx.x & 1;
    const long long q = col0 + c;
    const int i = (int)(q % cn) * 2 + h, j = (int)(q / cn);
    TraceParams p{1.0f, 0.0f, 1.0f, 0.0f};
    if (c < ncb && i < a.nx && j < a.ny) {
      const long long t = (long long)j * a.nx + i;
      p.inside = 1.0f;
      if (a.sf[l]) p.sf = a.sf[l][t];
      if (a.shift[l]) p.ishift = a.shift[l][t] / a.dz;
      if (a.gain[l]) p.gain = a.gain[l][t];
    }
    tp[threadIdx.x] = p;
  }
  const int c = (int)threadIdx.x % TX, j0 = (int)threadIdx.x / TX;
  cx<float> *col = buf + c * PC;
  const long long gc = col0 + min(c, ncb - 1);   // clamped global column
  const cx<float> *al = reinterpret_cast<const cx<float> *>(a.alpha);
  const cx<float> *be = reinterpret_cast<const cx<float> *>(a.beta);
  const cx<float> *rh = reinterpret_cast<const cx<float> *>(a.rho);
  const auto imp2 = [&](long long o) -> cx<float> {
    const cx<float> x = al[o], y = be[o], z = rh[o];
    return mk(impedance(a.A[l], x.x, y.x, z.x), impedance(a.A[l], x.y, y.y, z.y));
  };
  // 1. impedance of rows e < nz (rows j0, j0 + JS, ...), the taper to nzp
  // in the rows beyond from the column's first and last impedance
  constexpr int ITRIPS = (N + JS - 1) / JS;
  const cx<float> i0 = imp2(gc), il = imp2((long long)(a.nz - 1) * ncol + gc);
  cx<float> r[ITRIPS];
#pragma unroll
  for (int t = 0; t < ITRIPS; ++t) {
    const int e = j0 + t * JS;
    r[t] = mk(0.0f, 0.0f);
    if (e < a.nz) r[t] = imp2((long long)e * ncol + gc);
  }
  const float fac = 1.0f / (float)(a.nzp - a.nz - 1);
#pragma unroll
  for (int t = 0; t < ITRIPS; ++t) {
    const int e = j0 + t * JS;
    if (e < a.nz) {
      col[e] = r[t];
    } else if (N % JS == 0 || e < N) {
      const float u = (float)(e - a.nz), w = (float)(a.nzp - e - 1);
      col[e] = mk(fac * (u * i0.x + w * il.x), fac * (u * i0.y + w * il.y));
    }
  }
  __syncthreads();
  // 2. + 3. the cyclic forward difference as the first pass's input
  const auto diff = [&](int e) -> cx<float> { return sub(col[e == N - 1 ? 0 : e + 1], col[e]); };
  cpasses<N, TX, THREADS, -1, 0, true, 1, float>(col, tw, diff, LdsIo{}, j0);
  // filter over the column's slots
  {
    const TraceParams pa = tp[2 * c], pb = tp[2 * c + 1];
    constexpr int NS = N / 2;
    constexpr int STRIPS = (NS + JS - 1) / JS;
#pragma unroll
    for (int t = 0; t < STRIPS; ++t) {
      const int s = j0 + t * JS;
      if (NS % JS == 0 || s < NS) {
        // slot 0: bin 0 pairs with itself
        cx<float> zk = col[s], zm = col[s == 0 ? 0 : N - s];
        filter_pair_h<LOCAL>(wsp, pa, pb, s, N, hn, zk, zm);
        col[s] = zk;
        if (s != 0) col[N - s] = zm;
      }
    }
    if (j0 == 0) {
      // slot 0 also holds the Nyquist bin
      cx<float> zn = col[N / 2], dummy = zn;
      filter_pair_h<LOCAL>(wsp, pa, pb, N / 2, N, hn, zn, dummy);
      col[N / 2] = zn;
    }
  }
  __syncthreads();
  // 4. inverse passes, the last storing the seismic
  float *out = a.out[l];
  const bool ia = tp[2 * c].inside != 0.0f, ib = tp[2 * c + 1].inside != 0.0f;
  const auto st = [&](int e, cx<float> v) {
    if (c >= ncb) return;
    const long long o = (long long)e * ncol + col0 + c;
    if (!ia || !ib) {
      const cx<float> p = imp2(o);
      if (!ia) v.x = p.x;
      if (!ib) v.y = p.y;
    }
    reinterpret_cast<cx<float> *>(out)[o] = v;
  };
  cpasses<N, TX, THREADS, 1, 0, true, 1, float>(col, tw, LdsIo{}, st, j0);
}

template <int N, int TX, int TH, bool LOCAL>
static int launch_cols_l(const FwdArgs &a, const void *tw, hipStream_t s) {
  const size_t lds = sizeof(cx<float>) * ((size_t)N + (size_t)(N / 2 + 1) + (size_t)TX * (N | 1)) +
                     sizeof(TraceParams) * 2 * TX;
  if (lds > 65536)
    CRAVA_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(k_fm_cols<N, TX, TH, LOCAL>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const long long ncol = (long long)(a.rnxp / 2) * a.nyp;
  const unsigned gx = (unsigned)((ncol + TX - 1) / TX);
  hipLaunchKernelGGL((k_fm_cols<N, TX, TH, LOCAL>), dim3(gx, (unsigned)a.nt), dim3(TH), lds, s, a,
                     reinterpret_cast<const cx<float> *>(tw));
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

// True if any angle has a stretch, shift or gain grid.
static bool has_local_wavelets(const FwdArgs &a) {
  for (int l = 0; l < a.nt; ++l)
    if (a.sf[l] || a.shift[l] || a.gain[l]) return true;
  return false;
}

template <int N, int TX, int TH>
static int launch_fused_v(const FwdArgs &a, const void *tw, hipStream_t s) {
  return has_local_wavelets(a) ? launch_cols_l<N, TX, TH, true>(a, tw, s)
                               : launch_cols_l<N, TX, TH, false>(a, tw, s);
}

// 16 columns (32 traces) per workgroup of 512 threads: the fastest of
// {16, 8, 4} columns x {128, 256, 512} threads measured at 512x512x320.
template <int N>
static int launch_fused_n(const FwdArgs &a, const void *tw, hipStream_t s) {
  return launch_fused_v<N, 16, 512>(a, tw, s);
}

bool fm_fused_available(int nzp) { return static_fft_full(nzp); }

int launch_fm_fused(const FwdArgs &a, const void *tw, hipStream_t s) {
  switch (a.nzp) {
#define CRAVA_FM_N(n) \
  case n: return launch_fused_n<n>(a, tw, s);
    CRAVA_FM_N(64) CRAVA_FM_N(80) CRAVA_FM_N(128) CRAVA_FM_N(160) CRAVA_FM_N(216) CRAVA_FM_N(256)
    CRAVA_FM_N(288) CRAVA_FM_N(320) CRAVA_FM_N(512)
#undef CRAVA_FM_N
    default: set_error("no fused forward-modelling kernel for this nzp"); return CRAVA_HIP_E_INVALID;
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
