// Compile-time specialised Stockham passes for gfx950.
//
// Same algorithm as the generic kernels in fft.hip, but the length N, the
// radix sequence, the pencils per workgroup TX and the workgroup size are
// template constants, so every index division folds into shifts or
// multiply-highs and the passes unroll.  Each pencil sits in LDS with one pad
// slot every 32 complex floats or 16 complex doubles and an odd pencil pitch,
// which spreads the strided early-pass writes and the column-major tile
// loads over the banks.
#include "../../include/crava_hip.h"
#include "fft_dft.hpp"
#include "fft_static.hpp"
#include "stockham.hpp"

#include <cmath>

namespace crava {

// Strided complex pass: columns contiguous, elements `es` apart.
// DIO: the first pass loads its butterflies straight from global memory and
// the last stores straight to it (pencil-fastest lanes, so each row segment
// of TX columns is one coalesced access), which saves two LDS round trips;
// otherwise the pencils are staged through LDS before and after the passes.
template <int N, int TX, int THREADS, int DIR, bool DIO, typename T, typename SI, typename SO>
__global__ void __launch_bounds__(THREADS) k_sfft_strided(GridSet gs, long long es, long long ncols, long long bstride,
                                                          const cx<T> *__restrict__ twg) {
  // (column block, batch, grid) of this workgroup, XCD-contiguous
  const unsigned W = xcd_work_id(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                                 gridDim.x * gridDim.y * gridDim.z);
  const unsigned bx = W % gridDim.x, by = (W / gridDim.x) % gridDim.y, bz = W / (gridDim.x * gridDim.y);
  const cx<SI> *src = reinterpret_cast<const cx<SI> *>(gs.src[bz]);
  cx<SO> *dst = reinterpret_cast<cx<SO> *>(gs.dst[bz]);
  extern __shared__ unsigned char smem_raw[];
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + N;
  constexpr int PITCH = SP<N, T>::pitch;
  const long long col0 = (long long)bx * TX;
  const int ncol = (int)min((long long)TX, ncols - col0);
  const long long off = (long long)by * bstride + col0;
  const cx<SI> *ibase = src + off;
  cx<SO> *obase = dst + off;
  for (int q = threadIdx.x; q < N; q += THREADS) tw[q] = twg[q];
  if constexpr (DIO) {
    // branch-free loads (clamped address, then select) so the compiler can
    // issue a thread's loads back to back before converting any of them
    const auto ld = [&](int c, int e) -> cx<T> {
      const cx<SI> v = ibase[min(c, ncol - 1) + (long long)e * es];
      return c < ncol ? mk((T)v.x, (T)v.y) : mk((T)0, (T)0);
    };
    const auto st = [&](int c, int e, cx<T> v) {
      if (c < ncol) obase[c + (long long)e * es] = mk((SO)v.x, (SO)v.y);
    };
    spasses_io<N, TX, THREADS, DIR, 0, true, true, T>(buf, tw, ld, st);
    return;
  }
  constexpr int TOT = TX * N;
  constexpr int TRIPS = (TOT + THREADS - 1) / THREADS;
#pragma unroll
  for (int t = 0; t < TRIPS; ++t) {
    const int q = (int)threadIdx.x + t * THREADS;
    if (TOT % THREADS == 0 || q < TOT) {
      const int e = q / TX, c = q % TX;
      if (c < ncol) {
        const cx<SI> v = ibase[c + (long long)e * es];
        buf[c * PITCH + padx<T>(e)] = mk((T)v.x, (T)v.y);
      }
    }
  }
  __syncthreads();
  spasses<N, TX, THREADS, DIR, 0>(buf, tw);
#pragma unroll
  for (int t = 0; t < TRIPS; ++t) {
    const int q = (int)threadIdx.x + t * THREADS;
    if (TOT % THREADS == 0 || q < TOT) {
      const int e = q / TX, c = q % TX;
      if (c < ncol) {
        const cx<T> v = buf[c * PITCH + padx<T>(e)];
        obase[c + (long long)e * es] = mk((SO)v.x, (SO)v.y);
      }
    }
  }
}

// x pass forward: real rows -> half spectra (two rows per complex pencil).
// DIO: the first pass loads rows 2c and 2c+1 straight from global memory.
template <int N, int TX, int THREADS, bool DIO, typename T, typename SI, typename SO>
__global__ void __launch_bounds__(THREADS) k_srows_r2c(GridSet gs, long long nrows, int rnxp,
                                                       const cx<T> *__restrict__ twg) {
  // (row block, grid) of this workgroup, XCD-contiguous
  const unsigned W = xcd_work_id(blockIdx.x + gridDim.x * blockIdx.z, gridDim.x * gridDim.z);
  const unsigned bx = W % gridDim.x, bz = W / gridDim.x;
  const SI *src = reinterpret_cast<const SI *>(gs.src[bz]);
  SO *dst = reinterpret_cast<SO *>(gs.dst[bz]);
  const T scale = (T)gs.scale[bz];
  // mean deflation: subtract c in storage precision (exact near c)
  const SI off = (gs.dc && ((gs.dcmask >> bz) & 1u)) ? (SI)gs.dc[bz] : (SI)0;
  extern __shared__ unsigned char smem_raw[];
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + N;
  constexpr int PITCH = SP<N, T>::pitch;
  const long long row0 = (long long)bx * 2 * TX;
  for (int q = threadIdx.x; q < N; q += THREADS) tw[q] = twg[q];
  static_assert(N % 2 == 0, "row kernels need an even length");
  if constexpr (DIO) {
    // branch-free loads (clamped rows, then select), as in the strided kernel
    const auto ld = [&](int c, int e) -> cx<T> {
      const long long ra = row0 + 2 * c;
      const SI a = src[min(ra, nrows - 1) * rnxp + e] - off;
      const SI b = src[min(ra + 1, nrows - 1) * rnxp + e] - off;
      return mk(ra < nrows ? (T)a : (T)0, ra + 1 < nrows ? (T)b : (T)0);
    };
    spasses_io<N, TX, THREADS, -1, 0, false, false, T>(buf, tw, ld, LdsIo{});
  } else {
    // lane (c, e2) reads the value pair e2 of rows 2c and 2c+1 and stores the
    // two whole complex elements 2e2, 2e2+1 of pencil c (a = Re, b = Im)
    constexpr int H = N / 2;
    constexpr int TOT = TX * H;
    constexpr int TRIPS = (TOT + THREADS - 1) / THREADS;
#pragma unroll
    for (int t = 0; t < TRIPS; ++t) {
      const int q = (int)threadIdx.x + t * THREADS;
      if (TOT % THREADS == 0 || q < TOT) {
        const int c = q / H, e2 = q % H;
        const long long ra = row0 + 2 * c;
        cx<SI> a = mk((SI)0, (SI)0), b = mk((SI)0, (SI)0);
        if (ra < nrows) {
          a = reinterpret_cast<const cx<SI> *>(src + ra * rnxp)[e2];
          a = mk((SI)(a.x - off), (SI)(a.y - off));
        }
        if (ra + 1 < nrows) {
          b = reinterpret_cast<const cx<SI> *>(src + (ra + 1) * rnxp)[e2];
          b = mk((SI)(b.x - off), (SI)(b.y - off));
        }
        cx<T> *d = buf + c * PITCH;
        d[padx<T>(2 * e2)] = mk((T)a.x, (T)b.x);
        d[padx<T>(2 * e2 + 1)] = mk((T)a.y, (T)b.y);
      }
    }
    __syncthreads();
    spasses<N, TX, THREADS, -1, 0>(buf, tw);
  }
  constexpr int CN = N / 2 + 1;
  constexpr int TOT2 = TX * CN;
  constexpr int TRIPS2 = (TOT2 + THREADS - 1) / THREADS;
  const T hs = (T)0.5 * scale;
#pragma unroll
  for (int t = 0; t < TRIPS2; ++t) {
    const int q = (int)threadIdx.x + t * THREADS;
    if (TOT2 % THREADS == 0 || q < TOT2) {
      const int c = q / CN, k = q % CN;
      const cx<T> z = buf[c * PITCH + padx<T>(k)];
      const cx<T> zm = buf[c * PITCH + padx<T>(k == 0 ? 0 : N - k)];
      const long long ra = row0 + 2 * c;
      if (ra < nrows)
        reinterpret_cast<cx<SO> *>(dst + ra * rnxp)[k] = mk((SO)(hs * (z.x + zm.x)), (SO)(hs * (z.y - zm.y)));
      if (ra + 1 < nrows)
        reinterpret_cast<cx<SO> *>(dst + (ra + 1) * rnxp)[k] = mk((SO)(hs * (z.y + zm.y)), (SO)(hs * (zm.x - z.x)));
    }
  }
}

// x pass inverse: half spectra of two rows -> Z = A + iB -> inverse -> rows.
// DIO: the first pass builds Z straight from global memory and the last
// stores the two rows straight to it.
template <int N, int TX, int THREADS, bool DIO, typename T, typename SI, typename SO>
__global__ void __launch_bounds__(THREADS) k_srows_c2r(GridSet gs, long long nrows, int rnxp,
                                                       const cx<T> *__restrict__ twg) {
  // (row block, grid) of this workgroup, XCD-contiguous
  const unsigned W = xcd_work_id(blockIdx.x + gridDim.x * blockIdx.z, gridDim.x * gridDim.z);
  const unsigned bx = W % gridDim.x, bz = W / gridDim.x;
  const SI *src = reinterpret_cast<const SI *>(gs.src[bz]);
  SO *dst = reinterpret_cast<SO *>(gs.dst[bz]);
  const T scale = (T)gs.scale[bz];
  // mean deflation: the level removed from the DC bin returns here
  const T add = (gs.dc && ((gs.dcmask >> bz) & 1u)) ? (T)gs.dc[bz] : (T)0;
  extern __shared__ unsigned char smem_raw[];
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + N;
  constexpr int PITCH = SP<N, T>::pitch;
  constexpr int CN = N / 2 + 1;
  const long long row0 = (long long)bx * 2 * TX;
  for (int q = threadIdx.x; q < N; q += THREADS) tw[q] = twg[q];
  if constexpr (DIO) {
    const auto ld = [&](int c, int k) -> cx<T> {
      const bool mirror = k >= CN;
      const int kk = mirror ? N - k : k;
      const bool selfc = (kk == 0) || (kk == N / 2);
      const long long ra = row0 + 2 * c;
      cx<SI> a = reinterpret_cast<const cx<SI> *>(src + min(ra, nrows - 1) * rnxp)[kk];
      cx<SI> b = reinterpret_cast<const cx<SI> *>(src + min(ra + 1, nrows - 1) * rnxp)[kk];
      if (ra >= nrows) a = mk((SI)0, (SI)0);
      if (ra + 1 >= nrows) b = mk((SI)0, (SI)0);
      if (selfc) { a.y = (SI)0; b.y = (SI)0; }
      if (mirror) { a.y = -a.y; b.y = -b.y; }
      return mk((T)a.x - (T)b.y, (T)a.y + (T)b.x);
    };
    const auto st = [&](int c, int e, cx<T> v) {
      const long long ra = row0 + 2 * c;
      if (ra < nrows) dst[ra * rnxp + e] = (SO)(v.x * scale + add);
      if (ra + 1 < nrows) dst[(ra + 1) * rnxp + e] = (SO)(v.y * scale + add);
    };
    spasses_io<N, TX, THREADS, 1, 0, false, false, T>(buf, tw, ld, st);
    return;
  }
  constexpr int TOT = TX * N;
  constexpr int TRIPS = (TOT + THREADS - 1) / THREADS;
#pragma unroll
  for (int t = 0; t < TRIPS; ++t) {
    const int q = (int)threadIdx.x + t * THREADS;
    if (TOT % THREADS == 0 || q < TOT) {
      const int c = q / N, k = q % N;
      const bool mirror = k >= CN;
      const int kk = mirror ? N - k : k;
      const bool selfc = (kk == 0) || (kk == N / 2);
      const long long ra = row0 + 2 * c;
      cx<SI> a = mk((SI)0, (SI)0), b = mk((SI)0, (SI)0);
      if (ra < nrows) a = reinterpret_cast<const cx<SI> *>(src + ra * rnxp)[kk];
      if (ra + 1 < nrows) b = reinterpret_cast<const cx<SI> *>(src + (ra + 1) * rnxp)[kk];
      if (selfc) { a.y = (SI)0; b.y = (SI)0; }
      if (mirror) { a.y = -a.y; b.y = -b.y; }
      buf[c * PITCH + padx<T>(k)] = mk((T)a.x - (T)b.y, (T)a.y + (T)b.x);
    }
  }
  __syncthreads();
  spasses<N, TX, THREADS, 1, 0>(buf, tw);
  // elements 2e2, 2e2+1 of pencil c -> value pair e2 of rows 2c (Re) and 2c+1 (Im)
  constexpr int H = N / 2;
  constexpr int TOT2 = TX * H;
  constexpr int TRIPS2 = (TOT2 + THREADS - 1) / THREADS;
#pragma unroll
  for (int t = 0; t < TRIPS2; ++t) {
    const int q = (int)threadIdx.x + t * THREADS;
    if (TOT2 % THREADS == 0 || q < TOT2) {
      const int c = q / H, e2 = q % H;
      const long long ra = row0 + 2 * c;
      const cx<T> *s0 = buf + c * PITCH;
      const cx<T> v0 = s0[padx<T>(2 * e2)], v1 = s0[padx<T>(2 * e2 + 1)];
      if (ra < nrows)
        reinterpret_cast<cx<SO> *>(dst + ra * rnxp)[e2] = mk((SO)(v0.x * scale + add), (SO)(v1.x * scale + add));
      if (ra + 1 < nrows)
        reinterpret_cast<cx<SO> *>(dst + (ra + 1) * rnxp)[e2] =
            mk((SO)(v0.y * scale + add), (SO)(v1.y * scale + add));
    }
  }
}

// ---------------------------------------------------------------------------
// Variant tables.  Strided: (TX, THREADS); rows: (row pairs, THREADS).
// ---------------------------------------------------------------------------
// Variants v and v + half share a shape; the upper half loads the first
// pass and stores the last straight from and to global memory.
constexpr int kSV[kNumStridedVariants][2] = {{16, 256}, {16, 512}, {8, 256}, {32, 512},
                                             {16, 256}, {16, 512}, {8, 256}, {32, 512}};
constexpr int kRV[kNumRowVariants][2] = {{8, 256}, {16, 512}, {4, 256}, {8, 256}, {16, 512}, {4, 256}};

template <typename F>
static int set_lds_attr(F kernel, size_t bytes) {
  if (bytes > 65536)
    CRAVA_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return CRAVA_HIP_OK;
}

template <int N, int V, int DIR, typename T, typename SI, typename SO>
static int launch_strided_v(const GridSet &gs, long long es, long long ncols, long long nbatch, long long bstride,
                            const void *tw, hipStream_t s) {
  constexpr int TX = kSV[V][0], TH = kSV[V][1];
  constexpr bool DIO = V >= kNumStridedVariants / 2;
  const size_t lds = lds_bytes<N, T>(TX);
  const unsigned gx = (unsigned)((ncols + TX - 1) / TX);
  int rc;
  if ((rc = set_lds_attr(k_sfft_strided<N, TX, TH, DIR, DIO, T, SI, SO>, lds))) return rc;
  hipLaunchKernelGGL((k_sfft_strided<N, TX, TH, DIR, DIO, T, SI, SO>), dim3(gx, (unsigned)nbatch, (unsigned)gs.n), dim3(TH),
                     lds, s, gs, es, ncols, bstride, reinterpret_cast<const cx<T> *>(tw));
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

template <int N, int V, bool INV, typename T, typename SI, typename SO>
static int launch_rows_v(const GridSet &gs, long long nrows, int rnxp, const void *tw, hipStream_t s) {
  constexpr int TX = kRV[V][0], TH = kRV[V][1];
  constexpr bool DIO = V >= kNumRowVariants / 2;
  const size_t lds = lds_bytes<N, T>(TX);
  const unsigned gx = (unsigned)((nrows + 2LL * TX - 1) / (2LL * TX));
  const cx<T> *twp = reinterpret_cast<const cx<T> *>(tw);
  int rc;
  if constexpr (!INV) {
    if ((rc = set_lds_attr(k_srows_r2c<N, TX, TH, DIO, T, SI, SO>, lds))) return rc;
    hipLaunchKernelGGL((k_srows_r2c<N, TX, TH, DIO, T, SI, SO>), dim3(gx, 1, (unsigned)gs.n), dim3(TH), lds, s, gs, nrows,
                       rnxp, twp);
  } else {
    if ((rc = set_lds_attr(k_srows_c2r<N, TX, TH, DIO, T, SI, SO>, lds))) return rc;
    hipLaunchKernelGGL((k_srows_c2r<N, TX, TH, DIO, T, SI, SO>), dim3(gx, 1, (unsigned)gs.n), dim3(TH), lds, s, gs, nrows,
                       rnxp, twp);
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

template <int N, int DIR, typename T, typename SI, typename SO>
static int strided_any(int v, const GridSet &gs, long long es, long long ncols, long long nbatch, long long bstride,
                       const void *tw, hipStream_t s) {
  switch (v) {
    case 0: return launch_strided_v<N, 0, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    case 1: return launch_strided_v<N, 1, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    case 2: return launch_strided_v<N, 2, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    case 3: return launch_strided_v<N, 3, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    case 4: return launch_strided_v<N, 4, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    case 5: return launch_strided_v<N, 5, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    case 6: return launch_strided_v<N, 6, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    default: return launch_strided_v<N, 7, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
  }
}

template <int N, bool INV, typename T, typename SI, typename SO>
static int rows_any(int v, const GridSet &gs, long long nrows, int rnxp, const void *tw, hipStream_t s) {
  switch (v) {
    case 0: return launch_rows_v<N, 0, INV, T, SI, SO>(gs, nrows, rnxp, tw, s);
    case 1: return launch_rows_v<N, 1, INV, T, SI, SO>(gs, nrows, rnxp, tw, s);
    case 2: return launch_rows_v<N, 2, INV, T, SI, SO>(gs, nrows, rnxp, tw, s);
    case 3: return launch_rows_v<N, 3, INV, T, SI, SO>(gs, nrows, rnxp, tw, s);
    case 4: return launch_rows_v<N, 4, INV, T, SI, SO>(gs, nrows, rnxp, tw, s);
    default: return launch_rows_v<N, 5, INV, T, SI, SO>(gs, nrows, rnxp, tw, s);
  }
}

template <int N, typename T, typename S>
static int all_kinds(FftKind kind, int v, const GridSet &gs, long long es, long long ncols, long long nbatch,
                     long long bstride, long long nrows, int rnxp, const void *tw, hipStream_t s) {
  switch (kind) {
    case kStridedFwd: return strided_any<N, -1, T, S, S>(v, gs, es, ncols, nbatch, bstride, tw, s);
    case kStridedInv: return strided_any<N, 1, T, S, S>(v, gs, es, ncols, nbatch, bstride, tw, s);
    case kRowsR2C: return rows_any<N, false, T, S, S>(v, gs, nrows, rnxp, tw, s);
    default: return rows_any<N, true, T, S, S>(v, gs, nrows, rnxp, tw, s);
  }
}

// Storage x arithmetic: float, double, and float storage with double
// arithmetic (pencils, twiddles and butterflies in double inside the pass;
// one rounding to float per output).
template <int N>
static int launch_store(int io, FftKind kind, int v, const GridSet &gs, long long es, long long ncols,
                        long long nbatch, long long bstride, long long nrows, int rnxp, const void *tw, hipStream_t s) {
  if (io == kIoF32) return all_kinds<N, float, float>(kind, v, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
  if (io == kIoF64) return all_kinds<N, double, double>(kind, v, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
  if (io == kIoF32MathF64)
    return all_kinds<N, double, float>(kind, v, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
  set_error("unsupported storage combination for this pass");
  return CRAVA_HIP_E_INVALID;
}

// Lengths with specialised kernels: the x, y and z lengths of the standard
// workloads (80/160 plain CPU grid, 288 plain GPU grid, 512/320 large grid,
// 216 differentiation grid) plus common powers of two.
#define CRAVA_STATIC_LENGTHS(X) X(64) X(80) X(128) X(160) X(216) X(256) X(288) X(320) X(512)

bool static_fft_available(int n) {
#define CRAVA_CASE(L) if (n == L) return true;
  CRAVA_STATIC_LENGTHS(CRAVA_CASE)
#undef CRAVA_CASE
  return false;
}

static size_t max_lds() {
  static int v = -1;
  if (v < 0) {
    int dev = 0, val = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&val, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && val > 0)
      v = val;
    else
      v = 65536;
  }
  return (size_t)v;
}

// Twiddles of the specialised kernels, laid out per radix pass: pass p
// (radix R, Ns = product of the earlier radices) holds W_n^(r k n/(Ns R)) for
// r = 1..R-1, k = 0..Ns-1 at off(p) + (r-1) Ns + k.  n entries in total
// (the last is unused).  re_im receives 2n values.
bool static_twiddle_table(int n, double *re_im) {
  if (!static_fft_available(n)) return false;
  const RadixList rl = make_radices(n);
  int o = 0, ns = 1;
  for (int i = 0; i < 2 * n; ++i) re_im[i] = 0.0;
  for (int p = 0; p < rl.cnt; ++p) {
    const int R = rl.r[p];
    for (int r = 1; r < R; ++r)
      for (int k = 0; k < ns; ++k) {
        const long long m = ((long long)r * k * (n / (ns * R))) % n;
        const long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
        re_im[2 * (o + (r - 1) * ns + k)] = (double)cosl(a);
        re_im[2 * (o + (r - 1) * ns + k) + 1] = (double)sinl(a);
      }
    o += ns * (R - 1);
    ns *= R;
  }
  return true;
}

bool static_fft_geometry(int n, bool dbl, FftKind kind, int variant, StaticLaunch *out) {
  if (!static_fft_available(n)) return false;
  const bool strided = kind == kStridedFwd || kind == kStridedInv;
  if (strided && (variant < 0 || variant >= kNumStridedVariants)) return false;
  if (!strided && (variant < 0 || variant >= kNumRowVariants)) return false;
  const int tx = strided ? kSV[variant][0] : kRV[variant][0];
  const int th = strided ? kSV[variant][1] : kRV[variant][1];
  const size_t eb = dbl ? 16 : 8;
  const size_t len = (size_t)n + (size_t)((n - 1) >> (dbl ? 4 : 5)) + 1;
  const size_t pitch = len | 1;
  const size_t lds = eb * ((size_t)n + (size_t)tx * pitch);
  if (lds > max_lds()) return false;
  if (out) { out->tx = tx; out->threads = th; out->lds = lds; }
  return true;
}

int static_fft_launch(int n, int io, FftKind kind, int variant, const GridSet &gs, long long es, long long ncols,
                      long long nbatch, long long bstride, long long nrows, int rnxp, const void *tw, hipStream_t s) {
  if (!static_fft_geometry(n, io != kIoF32, kind, variant, nullptr)) {
    set_error("no specialised transform for length " + std::to_string(n) + " / variant " + std::to_string(variant));
    return CRAVA_HIP_E_INVALID;
  }
#define CRAVA_CASE(L)                                                                                          \
  if (n == L)                                                                                                  \
    return launch_store<L>(io, kind, variant, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
  CRAVA_STATIC_LENGTHS(CRAVA_CASE)
#undef CRAVA_CASE
  return CRAVA_HIP_E_INVALID;
}

}  // namespace crava
