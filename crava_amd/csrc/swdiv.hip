// Data preparation of the non-constant-thickness path.
//
// 1. Wavelet division, Crava::divideDataByScaleWavelet (src/crava.cpp:592-714)
//    with Crava::computeAdjustmentFactor (:726-817).  Per angle l and padded
//    trace (i, j), i < nxp, j < nyp:
//      r(k) = d(k) / sqrt(nzp), times max(1 - dist(k)^2, 0) for k > nz
//             (FFTGrid::getDistToBoundary, src/fftgrid.cpp:1374-1412)
//      R    = r2c(r);  S(k) = R(k) f(k) for nz 1000 / (relT lz nzp) k inside
//             (lowCut, highCut), else 0;  d' = c2r(S) / sqrt(nzp)
//    with the adjustment factor f(k) of a trace whose local wavelet is the
//    angle's 1-D wavelet shifted and scaled with the shift / gain of the
//    mirrored trace (iInd, jInd) (:626-642) and read through the stretch
//    sf = relT = Simbox::getRelThick(i, j):
//      wLoc  = getCAmp(k, sf) of the local spectrum L(m) = gain e^{-2 pi i shift/dz m/nzp} W(m)
//      e3    = errorSmooth3 = central(backward(forward(local wavelet))), normalised
//              by its band norm (Wavelet::findNormWithinFrequencyBand)
//      sigma2Loc  = ev (wnc + (1 - wnc) 0.5 (|wLoc|^2/|w|^2 + |e3(k, sf)|^2))
//      sigma2Glob = ev (wnc + (1 - wnc) 0.5 (|L(k)|^2/|w|^2 + |e3(k)|^2))
//      f = (wLoc / (tau2 wLoc2Adj + sigma2Loc)) / (wGlob2Adj / (tau2 |W|^2 + sigma2Glob))
//    with the reference's three tolFac = 0.05 adjustments of |W|^2 and
//    |wLoc|^2 and tau2 = |Re| of the reflection-coefficient covariance
//    spectrum.
//    GPU form.  The reference builds errorSmooth3 per trace by three time-domain
//    differences and an r2c of the local wavelet; a circular difference is a
//    multiplication in frequency, so here e3(m) = i (2 cos t - 2) sin t L(m) /
//    (|gain| |E3|_band), t = 2 pi m / nzp, with the band norm of the global
//    wavelet's E3 computed once per angle (|L| = |gain| |W|).  No per-trace
//    wavelet transform remains.  The bins of e3 that the reference's
//    Wavelet::multiplyRAmpByConstant sets to RMISSING (floats of index >= nzp
//    of the interleaved half spectrum) are RMISSING here too, and the local
//    spectrum's even-length Nyquist bin is real, as after the reference's
//    c2r / r2c round trip.  As in forward.hip, the traces at x = 2m and
//    2m+1 are the two halves of complex column m: one complex z transform
//    serves two traces.
//      * nzp with a specialised transform: one fused kernel per (column
//        block, angle) -- loads with the scaling and taper, forward passes,
//        the filter in LDS, inverse passes storing the scaled result;
//      * other lengths: k_swd_prepare, the plan's z passes, k_swd_filter, the
//        inverse z passes, k_swd_finalize.
//    The slack column of the layout (x >= nxp) holds no data and comes back
//    within rounding of its input.
//
// 2. FFTGrid::interpolateSeismic (src/fftgrid.cpp:2564-2655) with
//    interpolateTrace (:2658-2696) and extrapolateSeismic (:2699-2733).
//    k_isq_energy computes each volume trace's energy; the host derives the
//    whole interpolation sequence from the energies alone (which traces are
//    rewritten and from which neighbours, in the reference's order) and
//    groups it into levels of independent operations; k_isq_interp runs one
//    level per launch and k_isq_extrapolate fills the traces outside the box
//    of good traces from their clamped fill traces with the 3-D taper.
//    Bit-exact: float sums in the reference's order, no contraction.
//
#include "../../include/crava_hip.h"
#include "fft_static.hpp"
#include "stockham.hpp"
#include "swdiv.hpp"
#include "tables.hpp"

#include <algorithm>
#include <cmath>

// This file's own arithmetic keeps the reference's float operations (no
// contraction into FMA); the Stockham passes included above are contracted.
#pragma clang fp contract(off)

namespace crava {

namespace {

constexpr float kRMissing = -99999.0f;   // src/definitions.h:35
constexpr float kTolFac = 0.05f;         // src/crava.cpp:734

// FFTGrid::getDistToBoundary for i < np
__host__ __device__ __forceinline__ float dist_to_boundary(int i, int n, int np) {
  if (i < n) return 0.0f;
  const int mn = n < np - n ? n : np - n;
  const float taper = (float)(mn / 2.1);
  const int below = np - i, above = i - (n - 1);
  return above < below ? (float)above / taper : (float)below / taper;
}

// FFTGrid::getFillNumber for i < np
__host__ __device__ __forceinline__ int fill_number(int i, int n, int np) {
  if (i < n) return i;
  const int below = np - i - 1, above = i - n + 1;
  if (above < below) return (n - above) > n / 2 ? (n - above) : n / 2;
  return below < n / 2 ? below : n / 2;
}

// wavelet index of a padded trace, src/crava.cpp:626-642
__device__ __forceinline__ int mirror(int i, int n, int np) {
  int r = i;
  if (r > 3 * n - 1) r = 0;
  if (r > (np + n) / 2) r = np - r;
  if (r >= n) r = 2 * n - r - 1;
  return r;
}

// the taper factor of sample k > nz (src/crava.cpp:652-656)
__device__ __forceinline__ float prep(float v, int k, float sq, int nz, int nzp) {
  float r = v / sq;
  if (k > nz) {
    const float d = dist_to_boundary(k, nz, nzp);
    const float f = 1 - d * d;
    r *= (f < 0.0f) ? 0.0f : f;
  }
  return r;
}

// prep for the fused kernel: the same scaling and taper with the division
// by sqrt(nzp) and the boundary distance's division as multiplications by
// precomputed reciprocals (isq = 1 / sqrt(nzp), itaper = 1 / taper length).
__device__ __forceinline__ float prep_r(float v, int k, float isq, float itaper, int nz, int nzp) {
  float r = v * isq;
  if (k > nz) {
    const int below = nzp - k, above = k - (nz - 1);
    const float d = (float)(above < below ? above : below) * itaper;
    const float f = 1 - d * d;
    r *= (f < 0.0f) ? 0.0f : f;
  }
  return r;
}

// Local parameters of one trace.
struct SwdTrace {
  double delta_f;   // nz 1000 / (relT lz nzp)
  double ph;        // phase of the shift per bin, in half-turns: 2 ishift / nzp
  float sf, ishift, gain, inside;
  float rsf, iag;   // 1 / sf, 1 / |gain|
  float cs, sn;     // cos, sin of pi ph (one bin's phase step)
};

// sin, cos of pi t for t = ph m: the argument is reduced to [-1, 1]
// half-turns in double, so the float sincospif sees an exact small argument;
// no shift (ph = 0) gives (0, 1) without evaluating it.
__device__ __forceinline__ void phase_at(double ph, int m, float *s, float *c) {
  if (ph == 0.0) {
    *s = 0.0f;
    *c = 1.0f;
    return;
  }
  const double t = ph * (double)m;
  sincospif((float)(t - 2.0 * rint(0.5 * t)), s, c);
}

__device__ __forceinline__ SwdTrace trace_params(const SwdArgs &a, int l, int i, int j) {
  SwdTrace p{0.0, 0.0, 1.0f, 0.0f, 1.0f, 0.0f, 1.0f, 1.0f, 1.0f, 0.0f};
  if (i >= a.nxp) return p;
  p.inside = 1.0f;
  const int ii = mirror(i, a.nx, a.nxp), jj = mirror(j, a.ny, a.nyp);
  float ts = 0.0f;
  if (ii >= 0 && ii < a.nx && jj >= 0 && jj < a.ny) {
    const long long t = (long long)jj * a.nx + ii;
    if (a.shift[l]) ts = a.shift[l][t];
    if (a.gain[l]) p.gain = a.gain[l][t];
  }
  p.ishift = ts / a.dz;
  const double rt = a.rel[(long long)j * a.nxp + i];
  p.sf = (float)(rt * 1.0f);
  p.delta_f = (double)a.nz * 1000.0 / (rt * a.lz * (double)a.nzp);
  p.rsf = 1.0f / p.sf;
  p.iag = 1.0f / fabsf(p.gain);
  p.ph = 2.0 * (double)p.ishift / (double)a.nzp;
  phase_at(p.ph, 1, &p.sn, &p.cs);
  return p;
}

// Local spectrum at bin m: gain e^{-2 pi i ishift m / nzp} W(m)
// (Wavelet1D::shiftAndScale), from the phase's (c, s) at m (phase_at).  An
// even length's Nyquist bin is real (a c2r ignores its imaginary part).
__device__ __forceinline__ cx<float> local_at(const cx<float> *W, int m, float c, float s, float gain, int nzp) {
#pragma clang fp contract(fast)
  const cx<float> w = W[m];
  const float mr = gain * c, mi = -(gain * s);
  cx<float> v = mk(mr * w.x - mi * w.y, mi * w.x + mr * w.y);
  if (2 * m == nzp) v.y = 0.0f;
  return v;
}

// Normalised errorSmooth3 at bin m from the local spectrum v there
// (iag = 1 / |gain|).
__device__ __forceinline__ cx<float> e3_at(const float *t3, int m, cx<float> v, float iag, int nzp) {
#pragma clang fp contract(fast)
  const float t = t3[m] * iag;
  cx<float> e = mk(-t * v.y, t * v.x);
  if (2 * m >= nzp) e.x = kRMissing;
  if (2 * m + 1 >= nzp) e.y = kRMissing;
  return e;
}

// Adjustment factor of one trace at bin k < nzp/2 + 1
// (src/crava.cpp:772-814).  Tables of the angle: W, t3, tau2, wg2.
__device__ __forceinline__ cx<float> adjustment(const cx<float> *W, const float *t3, const float *tau2,
                                                const float *wg2, const SwdTrace &p, int k, float sk, float ck,
                                                int nzp, float mod_w, float ev, float iwn2, float wnc) {
#pragma clang fp contract(fast)
  const int cn = nzp / 2 + 1;
  cx<float> wl = mk(0.0f, 0.0f), el = mk(0.0f, 0.0f);
  const float omega = (float)k * p.rsf;   // k / sf (interpolation is continuous in omega)
  if (!(omega >= (float)cn)) {
    const int omL = (int)floorf(omega);
    int omU = omL + 1;
    if (omU >= cn) omU -= 1;
    const float d = omega - (float)omL;
    // phase at omL, and at omL + 1 by one step's rotation
    float sL, cL;
    phase_at(p.ph, omL, &sL, &cL);
    float cU = cL, sU = sL;
    if (omU != omL) {
      cU = cL * p.cs - sL * p.sn;
      sU = sL * p.cs + cL * p.sn;
    }
    const cx<float> a = local_at(W, omL, cL, sL, p.gain, nzp), b = local_at(W, omU, cU, sU, p.gain, nzp);
    const cx<float> ea = e3_at(t3, omL, a, p.iag, nzp), eb = e3_at(t3, omU, b, p.iag, nzp);
    wl.x = a.x * (1.0f - d) + b.x * d;
    wl.y = (a.y * (1.0f - d) + b.y * d);
    el.x = ea.x * (1.0f - d) + eb.x * d;
    el.y = (ea.y * (1.0f - d) + eb.y * d);
  }
  wl.x *= p.rsf;
  wl.y *= p.rsf;
  el.x *= p.rsf;
  el.y *= p.rsf;
  const float wl2 = (wl.x * wl.x + wl.y * wl.y);
  const float wg2k = wg2[k];
  float s_loc = wl2 * iwn2;
  s_loc += el.x * el.x + el.y * el.y;
  const cx<float> c = local_at(W, k, ck, sk, p.gain, nzp);
  const cx<float> ec = e3_at(t3, k, c, p.iag, nzp);
  float s_glob = (c.x * c.x + c.y * c.y) * iwn2;
  s_glob += ec.x * ec.x + ec.y * ec.y;
  float sig_l = ev * (wnc + (1.0f - wnc) * 0.5f * s_loc);
  const float sig_g = ev * (wnc + (1.0f - wnc) * 0.5f * s_glob);
  const float t2 = tau2[k];
  float wg2a = wg2k, wl2a = wl2;
  if (wg2k < kTolFac * mod_w) wg2a = 0.5f * (wg2k + kTolFac * mod_w);
  if (wl2 < kTolFac * mod_w) wl2a = 0.5f * (wl2 + kTolFac * mod_w);
  if (wl2a < kTolFac * wg2a) {
    wl2a = 0.5f * (wl2 + kTolFac * wg2a);
    sig_l *= wl2a / wl2;
  }
  // (wLoc / den) / (wGlob2Adj / (tau2 |W|^2 + sigma2Glob)) with one
  // reciprocal (the hardware's, within 1 ulp)
  const float r = (t2 * wg2k + sig_g) * __builtin_amdgcn_rcpf((t2 * wl2a + sig_l) * wg2a);
  return mk(wl.x * r, wl.y * r);
}

// Filtered bin k of one trace: R f inside the band, 0 outside; the slack
// traces are only scaled by 1/nzp (their transforms are not normalised).
// (sk, ck): sin and cos of pi ph k, the shift's phase at bin k (phase_at).
__device__ __forceinline__ cx<float> filter_bin(const cx<float> *W, const float *t3, const float *tau2,
                                                const float *wg2, const SwdTrace &p, int k, float sk, float ck,
                                                cx<float> R, const SwdArgs &a, int l) {
#pragma clang fp contract(fast)
  if (p.inside == 0.0f) return scl(R, a.inv_n);
  if ((p.delta_f * k < a.high_cut) && (p.delta_f * k > a.low_cut)) {
    const cx<float> f = adjustment(W, t3, tau2, wg2, p, k, sk, ck, a.nzp, a.mod_w[l], a.ev[l], a.iwn2[l], a.wnc);
    const float tmp = R.x * f.x - R.y * f.y;
    const float im = R.y * f.x + R.x * f.y;
    return mk(tmp, im);
  }
  return mk(0.0f, 0.0f);
}

// Separate the spectra of the two traces of a complex column at bins k and
// N - k, filter each, pack again.  ph_a, ph_b: (sin, cos) of each trace's
// shift phase at bin k.
__device__ __forceinline__ void filter_column_pair(const cx<float> *W, const float *t3, const float *tau2,
                                                   const float *wg2, const SwdTrace &pa, const SwdTrace &pb,
                                                   int k, float2 ph_a, float2 ph_b, const SwdArgs &a, int l,
                                                   cx<float> &Zk, cx<float> &Zm) {
#pragma clang fp contract(fast)
  const int N = a.nzp;
  const cx<float> Z = Zk, M = Zm;
  const cx<float> Ra = mk(0.5f * (Z.x + M.x), 0.5f * (Z.y - M.y));
  const cx<float> Rb = mk(0.5f * (Z.y + M.y), 0.5f * (M.x - Z.x));
  const cx<float> Sa = filter_bin(W, t3, tau2, wg2, pa, k, ph_a.x, ph_a.y, Ra, a, l);
  const cx<float> Sb = filter_bin(W, t3, tau2, wg2, pb, k, ph_b.x, ph_b.y, Rb, a, l);
  if (k == (N - k) % N) {
    // DC, and Nyquist of an even length: a c2r reads only the real part
    Zk = mk(Sa.x, Sb.x);
  } else {
    Zk = mk(Sa.x - Sb.y, Sa.y + Sb.x);
    Zm = mk(Sa.x + Sb.y, Sb.x - Sa.y);
  }
}

}  // namespace

// One thread per trace position (i, j) of the plane; angle = blockIdx.y.
__global__ void __launch_bounds__(256) k_swd_prepare(SwdArgs a) {
  const long long plane = (long long)a.nyp * a.rnxp;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= plane || (int)(t % a.rnxp) >= a.nxp) return;
  float *g = a.g[blockIdx.y];
  for (int k = 0; k < a.nzp; ++k) g[(long long)k * plane + t] = prep(g[(long long)k * plane + t], k, a.sq, a.nz, a.nzp);
}

__global__ void __launch_bounds__(256) k_swd_finalize(SwdArgs a) {
  const long long plane = (long long)a.nyp * a.rnxp;
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= plane || (int)(t % a.rnxp) >= a.nxp) return;
  float *g = a.g[blockIdx.y];
  for (int k = 0; k < a.nzp; ++k) g[(long long)k * plane + t] = g[(long long)k * plane + t] / a.sq;
}

// One thread per (complex column q, bin k < nzp/2 + 1); angle = blockIdx.y.
// Column q = j * cnx + m carries traces (2m, j) and (2m+1, j).
__global__ void __launch_bounds__(256) k_swd_filter(SwdArgs a) {
  const int cnx = a.rnxp / 2;
  const long long ncol = (long long)cnx * a.nyp;
  const int cn = a.nzp / 2 + 1;
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  if (id >= ncol * cn) return;
  const long long q = id % ncol;
  const int k = (int)(id / ncol);
  const int l = blockIdx.y;
  const int km = (a.nzp - k) % a.nzp;
  const int m = (int)(q % cnx), j = (int)(q / cnx);
  const SwdTrace pa = trace_params(a, l, 2 * m, j), pb = trace_params(a, l, 2 * m + 1, j);
  const long long o = (long long)l * cn;
  cx<float> *g = reinterpret_cast<cx<float> *>(a.g[l]);
  cx<float> zk = g[(long long)k * ncol + q], zm = g[(long long)km * ncol + q];
  float2 ha, hb;
  phase_at(pa.ph, k, &ha.x, &ha.y);
  phase_at(pb.ph, k, &hb.x, &hb.y);
  filter_column_pair(a.W + o, a.t3 + o, a.tau2 + o, a.wg2 + o, pa, pb, k, ha, hb, a, l, zk, zm);
  g[(long long)k * ncol + q] = zk;
  if (k != km) g[(long long)km * ncol + q] = zm;
}

// Fused division of TX complex columns (2 TX traces) for one angle, z
// length N specialised, on the pencil-fastest passes (stockham.hpp cpass):
// thread tid keeps complex column c = tid % TX and its butterflies
// j = tid / TX + t (THREADS / TX) through the whole chain, so every LDS
// offset is a constant from the column's base (pencils N | 1 apart).  Grid:
// column blocks x angles.  LDS: the traces' parameters, twiddles, the
// angle's tables, the pencils.
//   1. forward transform, the first pass loading straight from global
//      memory with the scaling and taper;
//   2. the filter in place over slots s (bins s and N - s, slot 0 bins 0
//      and N/2), each thread on its own column's slots;
//   3. inverse transform, the last pass storing the result / sqrt(nzp).
template <int N, int TX, int THREADS>
__global__ void __launch_bounds__(THREADS) k_swd_fused(SwdArgs a, const cx<float> *__restrict__ twg) {
  static_assert(N % 2 == 0, "the specialised lengths are even");
  const unsigned Wid = xcd_work_id(blockIdx.x + gridDim.x * blockIdx.y, gridDim.x * gridDim.y);
  const unsigned bx = Wid / (unsigned)a.nt;
  const int l = (int)(Wid % (unsigned)a.nt);
  const int cnx = a.rnxp / 2;
  const long long ncol = (long long)cnx * a.nyp;
  const long long col0 = (long long)bx * TX;
  const int ncb = (int)min((long long)TX, ncol - col0);
  constexpr int NK = N / 2 + 1;
  constexpr int PC = N | 1;
  constexpr int JS = THREADS / TX;
  extern __shared__ unsigned char smem_raw[];
  SwdTrace *tp = reinterpret_cast<SwdTrace *>(smem_raw);
  cx<float> *tw = reinterpret_cast<cx<float> *>(tp + 2 * TX);
  cx<float> *wsp = tw + N;
  float *t3 = reinterpret_cast<float *>(wsp + NK);
  float *tau2 = t3 + NK;
  float *wg2 = tau2 + NK;
  cx<float> *buf = reinterpret_cast<cx<float> *>(wg2 + NK + (NK & 1));
  for (int q = threadIdx.x; q < N; q += THREADS) tw[q] = twg[q];
  const long long o = (long long)l * NK;
  for (int q = threadIdx.x; q < NK; q += THREADS) {
    wsp[q] = a.W[o + q];
    t3[q] = a.t3[o + q];
    tau2[q] = a.tau2[o + q];
    wg2[q] = a.wg2[o + q];
  }
  if (threadIdx.x < 2 * TX) {
    const int c = threadIdx.x >> 1, h = threadIdx.x & 1;
    const long long q = col0 + min(c, ncb - 1);
    tp[threadIdx.x] = trace_params(a, l, (int)(q % cnx) * 2 + h, (int)(q / cnx));
  }
  const int c = (int)threadIdx.x % TX, j0 = (int)threadIdx.x / TX;
  cx<float> *col = buf + c * PC;
  cx<float> *g = reinterpret_cast<cx<float> *>(a.g[l]);
  const long long gc = col0 + min(c, ncb - 1);   // clamped global column
  const int i0 = (int)(gc % cnx) * 2;
  const bool pa_in = i0 < a.nxp, pb_in = i0 + 1 < a.nxp;
  const float isq = 1.0f / a.sq;
  const int mn = a.nz < a.nzp - a.nz ? a.nz : a.nzp - a.nz;
  const float itaper = 1.0f / (float)(mn / 2.1);
  // 1. forward transform, the first pass loading with the scaling and taper
  // (branch-free: clamped column, then select)
  const auto ld = [&](int e) -> cx<float> {
    const cx<float> v = g[gc + (long long)e * ncol];
    cx<float> r;
    r.x = pa_in ? prep_r(v.x, e, isq, itaper, a.nz, a.nzp) : v.x;
    r.y = pb_in ? prep_r(v.y, e, isq, itaper, a.nz, a.nzp) : v.y;
    return c < ncb ? r : mk(0.0f, 0.0f);
  };
  cpasses<N, TX, THREADS, -1, 0, false, 1, float>(col, tw, ld, LdsIo{}, j0);
  // 2. filter over the column's slots
  {
    const SwdTrace pa = tp[2 * c], pb = tp[2 * c + 1];
    constexpr int NS = N / 2;
    constexpr int STRIPS = (NS + JS - 1) / JS;
    // the shifts' phases at the slots' bins k = j0 + t JS: exact at k = j0
    // (phase_at), then rotated by the phase of JS bins
    float2 ha, hb, da, db;
    phase_at(pa.ph, j0, &ha.x, &ha.y);
    phase_at(pb.ph, j0, &hb.x, &hb.y);
    phase_at(pa.ph, JS, &da.x, &da.y);
    phase_at(pb.ph, JS, &db.x, &db.y);
#pragma unroll
    for (int t = 0; t < STRIPS; ++t) {
      const int s = j0 + t * JS;
      if (NS % JS == 0 || s < NS) {
        // slot 0: bin 0 pairs with itself
        cx<float> zk = col[s], zm = col[s == 0 ? 0 : N - s];
        filter_column_pair(wsp, t3, tau2, wg2, pa, pb, s, ha, hb, a, l, zk, zm);
        col[s] = zk;
        if (s != 0) col[N - s] = zm;
      }
      ha = make_float2(ha.x * da.y + ha.y * da.x, ha.y * da.y - ha.x * da.x);
      hb = make_float2(hb.x * db.y + hb.y * db.x, hb.y * db.y - hb.x * db.x);
    }
    if (j0 == 0) {
      // slot 0 also holds the Nyquist bin
      cx<float> zn = col[N / 2], dummy = zn;
      phase_at(pa.ph, N / 2, &ha.x, &ha.y);
      phase_at(pb.ph, N / 2, &hb.x, &hb.y);
      filter_column_pair(wsp, t3, tau2, wg2, pa, pb, N / 2, ha, hb, a, l, zn, dummy);
      col[N / 2] = zn;
    }
  }
  __syncthreads();
  // 3. inverse transform, the last pass storing the result / sqrt(nzp)
  const bool ia = tp[2 * c].inside != 0.0f, ib = tp[2 * c + 1].inside != 0.0f;
  const auto st = [&](int e, cx<float> v) {
    if (c >= ncb) return;
    if (ia) v.x = v.x * isq;
    if (ib) v.y = v.y * isq;
    g[col0 + c + (long long)e * ncol] = v;
  };
  cpasses<N, TX, THREADS, 1, 0, true, 1, float>(col, tw, LdsIo{}, st, j0);
}

template <int N, int TX, int TH>
static int launch_fused_v(const SwdArgs &a, const void *tw, hipStream_t s) {
  constexpr int NK = N / 2 + 1;
  const size_t lds = sizeof(SwdTrace) * 2 * TX + sizeof(cx<float>) * ((size_t)N + NK) +
                     sizeof(float) * (3 * NK + (NK & 1)) + sizeof(cx<float>) * (size_t)TX * (N | 1);
  if (lds > 65536)
    CRAVA_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(k_swd_fused<N, TX, TH>),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds));
  const long long ncol = (long long)(a.rnxp / 2) * a.nyp;
  const unsigned gx = (unsigned)((ncol + TX - 1) / TX);
  hipLaunchKernelGGL((k_swd_fused<N, TX, TH>), dim3(gx, (unsigned)a.nt), dim3(TH), lds, s, a,
                     reinterpret_cast<const cx<float> *>(tw));
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

bool swd_fused_available(int nzp) { return static_fft_full(nzp); }

int launch_swd_fused(const SwdArgs &a, const void *tw, hipStream_t s) {
  switch (a.nzp) {
#define CRAVA_SWD_N(n) \
  case n: return launch_fused_v<n, 16, 512>(a, tw, s);
    CRAVA_SWD_N(64) CRAVA_SWD_N(80) CRAVA_SWD_N(128) CRAVA_SWD_N(160) CRAVA_SWD_N(216) CRAVA_SWD_N(256)
    CRAVA_SWD_N(288) CRAVA_SWD_N(320) CRAVA_SWD_N(512)
#undef CRAVA_SWD_N
    default: set_error("no fused wavelet-division kernel for this nzp"); return CRAVA_HIP_E_INVALID;
  }
}

int launch_swd_prepare(const SwdArgs &a, hipStream_t s) {
  const long long plane = (long long)a.nyp * a.rnxp;
  hipLaunchKernelGGL(k_swd_prepare, dim3((unsigned)((plane + 255) / 256), (unsigned)a.nt), dim3(256), 0, s, a);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_swd_finalize(const SwdArgs &a, hipStream_t s) {
  const long long plane = (long long)a.nyp * a.rnxp;
  hipLaunchKernelGGL(k_swd_finalize, dim3((unsigned)((plane + 255) / 256), (unsigned)a.nt), dim3(256), 0, s, a);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_swd_filter(const SwdArgs &a, hipStream_t s) {
  const long long n = (long long)(a.rnxp / 2) * a.nyp * (a.nzp / 2 + 1);
  hipLaunchKernelGGL(k_swd_filter, dim3((unsigned)((n + 255) / 256), (unsigned)a.nt), dim3(256), 0, s, a);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

void swd_angle_tables(int nz, int nzp, const float *wavelet, float wnorm, const float *A3, const double *var0,
                      const float *corr_t, float dz, double lz, float low_cut, float high_cut, float *W_re_im,
                      float *t3, float *tau2, float *wg2, float *mod_w) {
  const int cn = nzp / 2 + 1;
  wavelet_half_spectrum(wavelet, nzp, W_re_im);
  // computeElasticImpedanceTimeCovariance + computeReflectionCoefficientTimeCovariance
  // (src/crava.cpp:476-504), float in the reference's order
  double ei_var = 0.0;
  for (int a = 0; a < 3; ++a)
    for (int b = 0; b < 3; ++b) ei_var += (double)(A3[a] * (float)var0[3 * a + b] * A3[b]);
  std::vector<float> rc(nzp), spec(2 * cn);
  for (int k = 0; k < nzp; ++k) rc[k] = (float)(ei_var * corr_t[k]);
  const float first = rc[0];
  float prev = rc[nzp - 1];
  for (int t = 0; t < nzp - 1; ++t) {
    const float curr = rc[t], next = rc[t + 1];
    rc[t] = 2 * curr - next - prev;
    prev = curr;
  }
  rc[nzp - 1] = 2 * rc[nzp - 1] - first - prev;
  wavelet_half_spectrum(rc.data(), nzp, spec.data());
  for (int k = 0; k < cn; ++k) {
    tau2[k] = std::fabs(spec[2 * k]);
    wg2[k] = (W_re_im[2 * k] * W_re_im[2 * k] + W_re_im[2 * k + 1] * W_re_im[2 * k + 1]);
  }
  // band norm of E3 = D3 W, D3(m) = i (2 cos t - 2) sin t (Wavelet::findNormWithinFrequencyBand)
  const float d_om = (float)(1.0 / ((double)((float)nzp * dz) * 0.001));
  int lo = (int)std::floor(low_cut / d_om), hi = (int)std::ceil(high_cut / d_om);
  lo = std::max(lo, 0);
  hi = std::min(hi, cn);
  // D3(m) depends on nzp alone: kept per length (and per thread) across
  // angles and calls
  thread_local int d3_n = 0;
  thread_local std::vector<double> d3;
  if (d3_n != nzp) {
    d3.assign(cn, 0.0);
    for (int m = 0; m < cn; ++m) {
      const long double t = 2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)nzp;
      d3[m] = (double)((2.0L * cosl(t) - 2.0L) * sinl(t));
    }
    d3_n = nzp;
  }
  const double fac = 1.0 / (double)nzp;
  double n2 = 0.0;
  for (int q = lo; q < hi; ++q)
    n2 += 2.0 * fac * d3[q] * d3[q] *
          ((double)W_re_im[2 * q] * W_re_im[2 * q] + (double)W_re_im[2 * q + 1] * W_re_im[2 * q + 1]);
  if (lo == 0) n2 -= fac * d3[0] * d3[0] * ((double)W_re_im[0] * W_re_im[0] + (double)W_re_im[1] * W_re_im[1]);
  const double nrm = std::sqrt(n2);
  for (int m = 0; m < cn; ++m) t3[m] = (float)(d3[m] / nrm);
  float mw = wnorm;                                     // src/crava.cpp:765-768
  mw *= mw;
  const float maxfreq = (float)((nzp / 2) * 1000.0 * nz) / (float)(lz * nzp);
  mw *= maxfreq / high_cut;
  *mod_w = mw;
}

// ---------------------------------------------------------------- interpolateSeismic

// One thread per volume trace (i < nx, j < ny): float energy over k < nz.
__global__ void __launch_bounds__(256) k_isq_energy(const float *__restrict__ g, int nx, int ny, int nz,
                                                    long long plane, int rnxp, float *__restrict__ energy) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)nx * ny) return;
  const int i = (int)(t % nx), j = (int)(t / nx);
  float e = 0.0f;
  for (int k = 0; k < nz; ++k) {
    const float v = g[(long long)k * plane + (long long)j * rnxp + i];
    e += v * v;
  }
  energy[t] = e;
}

// One thread per (operation, sample); the operations of one level write
// distinct traces and read only traces no operation of the level writes.
__global__ void __launch_bounds__(256) k_isq_interp(float *__restrict__ g, const IsqOp *__restrict__ ops, int nops,
                                                    int nx, long long plane, int rnxp, int nzp) {
  const long long id = (long long)blockIdx.x * 256 + threadIdx.x;
  if (id >= (long long)nops * nzp) return;
  const IsqOp op = ops[id % nops];
  const int k = (int)(id / nops);
  const int i = op.trace % nx, j = op.trace / nx;
  const long long o = (long long)k * plane + (long long)j * rnxp + i;
  float m = (op.mask & 1) ? g[o - 1] : 0.0f;
  if (op.mask & 2) m += g[o + 1];
  if (op.mask & 4) m += g[o - rnxp];
  if (op.mask & 8) m += g[o + rnxp];
  g[o] = m / (float)__popc(op.mask);
}

// b^3 rounded once (the reference's pow(b, 3)): b^3 = c + ce + pe b exactly
// up to the last product, with b b = p + pe and p b = c + ce.
__device__ __forceinline__ double cube_rn(double b) {
  const double p = b * b;
  const double pe = fma(b, b, -p);
  const double c = p * b;
  const double ce = fma(p, b, -c);
  return c + (ce + pe * b);
}

// One thread per padded trace (i < nxp, j < nyp) outside the box of good
// traces; its fill trace lies inside the box, which no thread writes.
__global__ void __launch_bounds__(256) k_isq_extrapolate(float *__restrict__ g, int nx, int ny, int nz, int nxp,
                                                         int nyp, int nzp, long long plane, int rnxp, int imin,
                                                         int imax, int jmin, int jmax) {
  const long long t = (long long)blockIdx.x * 256 + threadIdx.x;
  if (t >= (long long)nxp * nyp) return;
  const int i = (int)(t % nxp), j = (int)(t / nxp);
  if (!(i < imin || i > imax || j < jmin || j > jmax)) return;
  int refj = fill_number(j, ny, nyp);
  if (refj < jmin) refj = jmin;
  else if (refj > jmax) refj = jmax;
  int refi = fill_number(i, nx, nxp);
  if (refi < imin) refi = imin;
  else if (refi > imax) refi = imax;
  const float dx = dist_to_boundary(i, nx, nxp), dy = dist_to_boundary(j, ny, nyp);
  const float dx2 = dx * dx, dy2 = dy * dy;
  for (int k = 0; k < nzp; ++k) {
    const float value = g[(long long)k * plane + (long long)refj * rnxp + refi];
    const float dzk = dist_to_boundary(k, nz, nzp);
    double b = 1.0 - (double)dx2 - (double)dy2 - (double)(dzk * dzk);
    if (b < 0.0) b = 0.0;
    const float mult = (float)cube_rn(b);
    g[(long long)k * plane + (long long)j * rnxp + i] = mult * value;
  }
}

void isq_schedule(const float *energy, int nx, int ny, float threshold, IsqSchedule *out) {
  const int n = nx * ny;
  std::vector<short> flags(n);
  float total = 0.0f;
  for (int t = 0; t < n; ++t) total += energy[t];
  const float limit = threshold * total / (float)(nx * ny);
  int flag = 0, imin = nx, imax = 0, jmin = ny, jmax = 0, n_flagged = 0, n_zero = 0;
  for (int j = 0, t = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i, ++t) {
      int cur = 0;
      if (energy[t] <= limit) {
        cur = 1;
        ++n_flagged;
        if (energy[t] == 0.0f) ++n_zero;
      }
      flags[t] = (short)(flag + cur);
      if (cur == 1) {
        flag = 2;
      } else {
        imin = std::min(imin, i);
        imax = std::max(imax, i);
        jmin = std::min(jmin, j);
        jmax = std::max(jmax, j);
      }
    }
  // interpolateTrace on flags only; level = 1 + the latest level among the
  // traces read (0 for traces never rewritten)
  std::vector<int> level(n, 0);
  std::vector<std::pair<int, IsqOp>> seq;
  const auto interp = [&](int index, int i, int j) -> int {
    const bool left = i > 0 && flags[index - 1] % 2 == 0;
    const bool right = i < nx - 1 && flags[index + 1] % 2 == 0;
    const bool up = j > 0 && flags[index - nx] % 2 == 0;
    const bool down = j < ny - 1 && flags[index + nx] % 2 == 0;
    const int nt = (int)left + (int)right + (int)up + (int)down;
    if (nt > 1) {
      int lv = 0;
      if (left) lv = std::max(lv, level[index - 1]);
      if (right) lv = std::max(lv, level[index + 1]);
      if (up) lv = std::max(lv, level[index - nx]);
      if (down) lv = std::max(lv, level[index + nx]);
      level[index] = lv + 1;
      seq.push_back({lv + 1, IsqOp{index, (left ? 1 : 0) | (right ? 2 : 0) | (up ? 4 : 0) | (down ? 8 : 0)}});
      flags[index]--;
    }
    return flags[index];
  };
  for (int j = 0, cur = 0; j < ny; ++j)
    for (int i = 0; i < nx; ++i, ++cur) {
      if (flags[cur] % 2 != 1) continue;
      if (interp(cur, i, j) % 2 != 0) continue;
      int index = cur - 1;
      while (index >= 0 && flags[index] > 1) {
        if (flags[index] % 2 == 1) interp(index, index % nx, index / nx);
        --index;
      }
      if (index < 0) index = 0;
      ++index;
      while (flags[index - 1] == 0 && index <= cur) {
        if (flags[index] > 1) flags[index] -= 2;
        ++index;
      }
    }
  std::stable_sort(seq.begin(), seq.end(),
                   [](const std::pair<int, IsqOp> &x, const std::pair<int, IsqOp> &y) { return x.first < y.first; });
  out->ops.clear();
  out->level_start.clear();
  int cur_level = 0;
  for (size_t q = 0; q < seq.size(); ++q) {
    if (seq[q].first != cur_level) {
      out->level_start.push_back((int)q);
      cur_level = seq[q].first;
    }
    out->ops.push_back(seq[q].second);
  }
  out->level_start.push_back((int)seq.size());
  out->imin = imin;
  out->imax = imax;
  out->jmin = jmin;
  out->jmax = jmax;
  out->n_flagged = n_flagged;
  out->n_zero = n_zero;
}

int launch_isq_energy(const float *g, int nx, int ny, int nz, int nyp, int rnxp, float *energy, hipStream_t s) {
  const long long n = (long long)nx * ny;
  hipLaunchKernelGGL(k_isq_energy, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, nx, ny, nz,
                     (long long)nyp * rnxp, rnxp, energy);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_isq_interp(float *g, const IsqOp *ops, int nops, int nx, int nyp, int nzp, int rnxp, hipStream_t s) {
  if (nops <= 0) return CRAVA_HIP_OK;
  const long long n = (long long)nops * nzp;
  hipLaunchKernelGGL(k_isq_interp, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, ops, nops, nx,
                     (long long)nyp * rnxp, rnxp, nzp);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int launch_isq_extrapolate(float *g, int nx, int ny, int nz, int nxp, int nyp, int nzp, int rnxp, int imin,
                           int imax, int jmin, int jmax, hipStream_t s) {
  const long long n = (long long)nxp * nyp;
  hipLaunchKernelGGL(k_isq_extrapolate, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, s, g, nx, ny, nz, nxp,
                     nyp, nzp, (long long)nyp * rnxp, rnxp, imin, imax, jmin, jmax);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

}  // namespace crava
