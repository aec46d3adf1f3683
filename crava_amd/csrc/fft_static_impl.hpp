// Compile-time specialised Stockham passes for gfx950: the kernel and launch
// templates, instantiated per length by fft_static.hip (the main lengths,
// every launch variant) and fft_static_ext*.hip (the other 7-smooth lengths,
// the direct-I/O launch shapes).
//
// Same algorithm as the generic kernels in fft.hip, but the length N, the
// radix sequence, the pencils per workgroup TX and the workgroup size are
// template constants, so every index division folds into shifts or
// multiply-highs and the passes unroll.  Each pencil sits in LDS with one pad
// slot every 32 complex floats or 16 complex doubles and an odd pencil pitch,
// which spreads the strided early-pass writes and the column-major tile
// loads over the banks.
#pragma once
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

// Direct-I/O strided pass on the pencil-fastest passes (stockham.hpp cpass):
// thread tid keeps column tid % TX through every pass, the pencils sit N | 1
// apart in LDS with no padding, and every LDS offset a thread uses is a
// compile-time constant from its column's base.  The first pass loads its
// butterflies straight from global memory (row segments of TX columns), the
// last stores straight to it.
template <int N, int TX, int THREADS, int DIR, typename T, typename SI, typename SO>
__global__ void __launch_bounds__(THREADS) k_sfft_cols(GridSet gs, long long es, long long ncols, long long bstride,
                                                       const cx<T> *__restrict__ twg) {
  // (column block, batch, grid) of this workgroup, XCD-contiguous
  const unsigned W = xcd_work_id(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                                 gridDim.x * gridDim.y * gridDim.z);
  const unsigned bx = W % gridDim.x, by = (W / gridDim.x) % gridDim.y, bz = W / (gridDim.x * gridDim.y);
  extern __shared__ unsigned char smem_raw[];
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + N;
  const long long col0 = (long long)bx * TX;
  const int ncol = (int)min((long long)TX, ncols - col0);
  const long long off = (long long)by * bstride + col0;
  // the first pass reads no twiddles; the barrier that ends it orders these
  // writes before the later passes read them
  for (int q = threadIdx.x; q < N; q += THREADS) tw[q] = twg[q];
  const int c = (int)threadIdx.x % TX, j0 = (int)threadIdx.x / TX;
  const bool ok = c < ncol;
  const cx<SI> *ib = reinterpret_cast<const cx<SI> *>(gs.src[bz]) + off + min(c, ncol - 1);
  cx<SO> *ob = reinterpret_cast<cx<SO> *>(gs.dst[bz]) + off + c;
  // branch-free loads (clamped column, then select)
  const auto ld = [&](int e) -> cx<T> {
    const cx<SI> v = ib[(long long)e * es];
    return ok ? mk((T)v.x, (T)v.y) : mk((T)0, (T)0);
  };
  const auto st = [&](int e, cx<T> v) {
    if (ok) ob[(long long)e * es] = mk((SO)v.x, (SO)v.y);
  };
  cpasses<N, TX, THREADS, DIR, 0, false, 1, T>(buf + c * (N | 1), tw, ld, st, j0);
}

// Persistent form of the direct-I/O strided pass.  The tiles (column block,
// batch, grid) are split into 8 contiguous ranges, one per XCD (workgroups
// are dispatched round-robin over the XCDs); the gridDim.x / 8 workgroups of
// an XCD walk its range together, workgroup i taking tiles i, i + P/8, ...,
// so at any moment they work on neighbouring column blocks.  Each workgroup
// issues the next tile's first-pass loads into registers before it runs the
// current tile's passes, so those loads are in flight during the LDS work.
// Every workgroup's trip count depends on blockIdx only: all its threads
// reach the same barriers, and every workgroup ends after its last tile.
template <int N, int TX, int THREADS, int DIR, typename T, typename SI, typename SO>
__global__ void __launch_bounds__(THREADS) k_sfft_pipe(GridSet gs, long long es, long long ncols, long long bstride,
                                                       unsigned gx, unsigned nb, const cx<T> *__restrict__ twg) {
  constexpr int R0 = SP<N, T>::rl.r[0];
  constexpr int NB0 = N / R0;
  constexpr int TOT0 = TX * NB0;
  constexpr int TRIPS0 = (TOT0 + THREADS - 1) / THREADS;
  constexpr int PITCH = SP<N, T>::pitch;
  const unsigned total = gx * nb * (unsigned)gs.n;
  const unsigned xcd = blockIdx.x % 8u, px = gridDim.x / 8u;
  const unsigned q8 = total / 8u, r8 = total % 8u;
  const unsigned lo = xcd * q8 + min(xcd, r8), hi = lo + q8 + (xcd < r8 ? 1u : 0u);
  unsigned W = lo + blockIdx.x / 8u;
  if (W >= hi) return;
  extern __shared__ unsigned char smem_raw[];
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + N;
  for (int q = threadIdx.x; q < N; q += THREADS) tw[q] = twg[q];
  const cx<SI> *ib;
  cx<SO> *ob;
  int ncol;
  const auto tile = [&](unsigned w) {
    const unsigned bx = w % gx, by = (w / gx) % nb, bz = w / (gx * nb);
    const long long col0 = (long long)bx * TX;
    ncol = (int)min((long long)TX, ncols - col0);
    const long long off = (long long)by * bstride + col0;
    ib = reinterpret_cast<const cx<SI> *>(gs.src[bz]) + off;
    ob = reinterpret_cast<cx<SO> *>(gs.dst[bz]) + off;
  };
  // first-pass butterfly inputs, pencil-fastest lanes (coalesced row
  // segments), clamped columns (the select happens at use)
  cx<SI> pre[TRIPS0][R0];
  const auto load = [&]() {
#pragma unroll
    for (int t = 0; t < TRIPS0; ++t) {
      const int b = (int)threadIdx.x + t * THREADS;
      if (TOT0 % THREADS == 0 || b < TOT0) {
        const int c = b % TX, j = b / TX;
#pragma unroll
        for (int r = 0; r < R0; ++r) pre[t][r] = ib[min(c, ncol - 1) + (long long)(j + r * NB0) * es];
      }
    }
  };
  tile(W);
  load();
  __syncthreads();  // twiddles
  for (;;) {
    cx<T> v[TRIPS0][R0];
#pragma unroll
    for (int t = 0; t < TRIPS0; ++t) {
      const int b = (int)threadIdx.x + t * THREADS;
      if (TOT0 % THREADS == 0 || b < TOT0) {
        const bool ok = b % TX < ncol;
#pragma unroll
        for (int r = 0; r < R0; ++r) v[t][r] = ok ? mk((T)pre[t][r].x, (T)pre[t][r].y) : mk((T)0, (T)0);
        Dft<R0, DIR, T>::run(v[t]);
      }
    }
    cx<SO> *const cur = ob;
    const int cur_ncol = ncol;
    const unsigned Wn = W + px;
    if (Wn < hi) {
      tile(Wn);
      load();
    }
#pragma unroll
    for (int t = 0; t < TRIPS0; ++t) {
      const int b = (int)threadIdx.x + t * THREADS;
      if (TOT0 % THREADS == 0 || b < TOT0) {
        const int c = b % TX, j = b / TX;
#pragma unroll
        for (int r = 0; r < R0; ++r) buf[c * PITCH + padx<T>(j * R0 + r)] = v[t][r];
      }
    }
    __syncthreads();
    const auto st = [&](int c, int e, cx<T> x) {
      if (c < cur_ncol) cur[c + (long long)e * es] = mk((SO)x.x, (SO)x.y);
    };
    // the thread index, opaque to the optimiser on every tile: the passes'
    // per-thread LDS addresses are then computed where they are used instead
    // of being hoisted out of the loop and held in registers for all of it
    int tid = (int)threadIdx.x;
    asm volatile("" : "+v"(tid));
    spasses_io<N, TX, THREADS, DIR, 1, true, true, T>(buf, tw, LdsIo{}, st, tid);
    if (Wn >= hi) break;
    W = Wn;
    __syncthreads();  // the last pass has read the pencils before the next tile's first pass writes them
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
  // the LDS-staged form reads value pairs; an odd length needs direct I/O
  static_assert(DIO || N % 2 == 0, "LDS-staged row kernels need an even length");
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
  static_assert(DIO || N % 2 == 0, "LDS-staged row kernels need an even length");
  const long long row0 = (long long)bx * 2 * TX;
  for (int q = threadIdx.x; q < N; q += THREADS) tw[q] = twg[q];
  if constexpr (DIO) {
    const auto ld = [&](int c, int k) -> cx<T> {
      const bool mirror = k >= CN;
      const int kk = mirror ? N - k : k;
      const bool selfc = (kk == 0) || (N % 2 == 0 && kk == N / 2);
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
      const bool selfc = (kk == 0) || (N % 2 == 0 && kk == N / 2);
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
// Strided (TX, THREADS, form): form 0 stages the pencils through LDS before
// and after the passes, form 1 loads the first pass and stores the last
// straight from and to global memory, form 2 is form 1 in persistent
// workgroups that prefetch the next tile (k_sfft_pipe), form 3 is form 1 on
// the pencil-fastest passes (k_sfft_cols).
constexpr int kSV[kNumStridedVariants][3] = {{16, 256, 0}, {16, 512, 0}, {8, 256, 0},  {32, 512, 0},
                                             {16, 256, 1}, {16, 512, 1}, {8, 256, 1},  {32, 512, 1},
                                             {16, 256, 2}, {16, 512, 2}, {16, 512, 3}, {16, 256, 3}};

// LDS bytes of a strided variant: the twiddles and TX pencils (padded
// pitch, or N | 1 for the pencil-fastest form).
template <int N, typename T>
constexpr size_t strided_lds(int tx, int form) {
  return form == 3 ? sizeof(cx<T>) * ((size_t)N + (size_t)tx * (size_t)(N | 1)) : lds_bytes<N, T>(tx);
}
constexpr int kRV[kNumRowVariants][2] = {{8, 256}, {16, 512}, {4, 256}, {8, 256}, {16, 512}, {4, 256}};

template <typename F>
static int set_lds_attr(F kernel, size_t bytes) {
  if (bytes > 65536)
    CRAVA_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  return CRAVA_HIP_OK;
}

// Workgroups of a persistent launch: the resident count per CU times the CU
// count, a multiple of 8 (one share per XCD), at most one per tile rounded
// up to 8.
template <typename K>
static int persistent_grid(K kernel, int threads, size_t lds, unsigned tiles, unsigned *grid) {
  static int ncu = 0;
  if (ncu <= 0) {
    int dev = 0;
    CRAVA_HIP_CHECK(hipGetDevice(&dev));
    CRAVA_HIP_CHECK(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  }
  int per_cu = 0;
  CRAVA_HIP_CHECK(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kernel, threads, lds));
  unsigned g = (unsigned)(max(per_cu, 1) * max(ncu, 1));
  g = max(8u, g / 8u * 8u);
  g = min(g, (tiles + 7u) / 8u * 8u);
  *grid = g;
  return CRAVA_HIP_OK;
}

template <int N, int V, int DIR, typename T, typename SI, typename SO>
static int launch_strided_v(const GridSet &gs, long long es, long long ncols, long long nbatch, long long bstride,
                            const void *tw, hipStream_t s) {
  constexpr int TX = kSV[V][0], TH = kSV[V][1], FORM = kSV[V][2];
  const size_t lds = strided_lds<N, T>(TX, FORM);
  const unsigned gx = (unsigned)((ncols + TX - 1) / TX);
  int rc;
  if constexpr (FORM == 2) {
    const auto k = k_sfft_pipe<N, TX, TH, DIR, T, SI, SO>;
    if ((rc = set_lds_attr(k, lds))) return rc;
    const unsigned tiles = gx * (unsigned)nbatch * (unsigned)gs.n;
    unsigned grid = 0;
    if ((rc = persistent_grid(k, TH, lds, tiles, &grid))) return rc;
    hipLaunchKernelGGL(k, dim3(grid), dim3(TH), lds, s, gs, es, ncols, bstride, gx, (unsigned)nbatch,
                       reinterpret_cast<const cx<T> *>(tw));
  } else if constexpr (FORM == 3) {
    if ((rc = set_lds_attr(k_sfft_cols<N, TX, TH, DIR, T, SI, SO>, lds))) return rc;
    hipLaunchKernelGGL((k_sfft_cols<N, TX, TH, DIR, T, SI, SO>), dim3(gx, (unsigned)nbatch, (unsigned)gs.n),
                       dim3(TH), lds, s, gs, es, ncols, bstride, reinterpret_cast<const cx<T> *>(tw));
  } else {
    constexpr bool DIO = FORM == 1;
    if ((rc = set_lds_attr(k_sfft_strided<N, TX, TH, DIR, DIO, T, SI, SO>, lds))) return rc;
    hipLaunchKernelGGL((k_sfft_strided<N, TX, TH, DIR, DIO, T, SI, SO>), dim3(gx, (unsigned)nbatch, (unsigned)gs.n),
                       dim3(TH), lds, s, gs, es, ncols, bstride, reinterpret_cast<const cx<T> *>(tw));
  }
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
    case 7: return launch_strided_v<N, 7, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    case 8: return launch_strided_v<N, 8, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    case 9: return launch_strided_v<N, 9, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    case 10: return launch_strided_v<N, 10, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
    default: return launch_strided_v<N, 11, DIR, T, SI, SO>(gs, es, ncols, nbatch, bstride, tw, s);
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

// The direct-I/O launch shapes only: strided variants 4 and 5 (16 columns,
// 256 or 512 threads) and 10 and 11 (the same on pencil-fastest passes),
// row variants 3 and 5 (8 or 4 row pairs, 256 threads).
template <int N, typename T, typename S>
static int ext_kinds(FftKind kind, int v, const GridSet &gs, long long es, long long ncols, long long nbatch,
                     long long bstride, long long nrows, int rnxp, const void *tw, hipStream_t s) {
  switch (kind) {
    case kStridedFwd:
      if (v == 10) return launch_strided_v<N, 10, -1, T, S, S>(gs, es, ncols, nbatch, bstride, tw, s);
      if (v == 11) return launch_strided_v<N, 11, -1, T, S, S>(gs, es, ncols, nbatch, bstride, tw, s);
      return v == 4 ? launch_strided_v<N, 4, -1, T, S, S>(gs, es, ncols, nbatch, bstride, tw, s)
                    : launch_strided_v<N, 5, -1, T, S, S>(gs, es, ncols, nbatch, bstride, tw, s);
    case kStridedInv:
      if (v == 10) return launch_strided_v<N, 10, 1, T, S, S>(gs, es, ncols, nbatch, bstride, tw, s);
      if (v == 11) return launch_strided_v<N, 11, 1, T, S, S>(gs, es, ncols, nbatch, bstride, tw, s);
      return v == 4 ? launch_strided_v<N, 4, 1, T, S, S>(gs, es, ncols, nbatch, bstride, tw, s)
                    : launch_strided_v<N, 5, 1, T, S, S>(gs, es, ncols, nbatch, bstride, tw, s);
    case kRowsR2C:
      return v == 3 ? launch_rows_v<N, 3, false, T, S, S>(gs, nrows, rnxp, tw, s)
                    : launch_rows_v<N, 5, false, T, S, S>(gs, nrows, rnxp, tw, s);
    default:
      return v == 3 ? launch_rows_v<N, 3, true, T, S, S>(gs, nrows, rnxp, tw, s)
                    : launch_rows_v<N, 5, true, T, S, S>(gs, nrows, rnxp, tw, s);
  }
}

template <int N>
static int ext_store(int io, FftKind kind, int v, const GridSet &gs, long long es, long long ncols, long long nbatch,
                     long long bstride, long long nrows, int rnxp, const void *tw, hipStream_t s) {
  if (io == kIoF32) return ext_kinds<N, float, float>(kind, v, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
  if (io == kIoF64) return ext_kinds<N, double, double>(kind, v, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
  if (io == kIoF32MathF64)
    return ext_kinds<N, double, float>(kind, v, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
  set_error("unsupported storage combination for this pass");
  return CRAVA_HIP_E_INVALID;
}

// The other 7-smooth lengths from 32 to 320, six groups (one translation
// unit each, fft_static_ext<k>.hip).
#define CRAVA_EXT_LENGTHS_0(X) X(32) X(35) X(36) X(40) X(42) X(45) X(48) X(49) X(50)
#define CRAVA_EXT_LENGTHS_1(X) X(54) X(56) X(60) X(63) X(70) X(72) X(75) X(81) X(84)
#define CRAVA_EXT_LENGTHS_2(X) X(90) X(96) X(98) X(100) X(105) X(108) X(112) X(120) X(125)
#define CRAVA_EXT_LENGTHS_3(X) X(126) X(135) X(140) X(144) X(147) X(150) X(162) X(168) X(175)
#define CRAVA_EXT_LENGTHS_4(X) X(180) X(189) X(192) X(196) X(200) X(210) X(224) X(225) X(240)
#define CRAVA_EXT_LENGTHS_5(X) X(243) X(245) X(250) X(252) X(270) X(280) X(294) X(300) X(315)

// Launch of an extended length (in group k) with an ext variant; E_INVALID
// for a length outside the group.
#define CRAVA_EXT_LAUNCH_DECL(k)                                                                                 \
  int static_ext_launch_##k(int n, int io, FftKind kind, int variant, const GridSet &gs, long long es,          \
                            long long ncols, long long nbatch, long long bstride, long long nrows, int rnxp,    \
                            const void *tw, hipStream_t s);
CRAVA_EXT_LAUNCH_DECL(0)
CRAVA_EXT_LAUNCH_DECL(1)
CRAVA_EXT_LAUNCH_DECL(2)
CRAVA_EXT_LAUNCH_DECL(3)
CRAVA_EXT_LAUNCH_DECL(4)
CRAVA_EXT_LAUNCH_DECL(5)
#undef CRAVA_EXT_LAUNCH_DECL

#define CRAVA_EXT_LAUNCH_DEF(k)                                                                                  \
  int static_ext_launch_##k(int n, int io, FftKind kind, int variant, const GridSet &gs, long long es,          \
                            long long ncols, long long nbatch, long long bstride, long long nrows, int rnxp,    \
                            const void *tw, hipStream_t s) {                                                     \
    CRAVA_EXT_LENGTHS_##k(CRAVA_EXT_CASE) return CRAVA_HIP_E_INVALID;                                            \
  }
#define CRAVA_EXT_CASE(L) \
  if (n == L) return ext_store<L>(io, kind, variant, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);

}  // namespace crava
