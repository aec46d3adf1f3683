// Compile-time Stockham passes over pencils staged in LDS, shared by the
// specialised transform kernels (fft_static.hip) and the fused forward
// modelling kernel (forward.hip).
#pragma once
#include <type_traits>

#include "fft_dft.hpp"

namespace crava {

// One pad slot every 32 complex floats, or every 16 complex doubles (a
// 16-byte element spans four banks; the denser padding keeps the strided
// Stockham writes spread over the banks).
template <typename T>
constexpr int kPadShift = sizeof(T) == 8 ? 4 : 5;
template <typename T>
__device__ __forceinline__ constexpr int padx(int x) { return x + (x >> kPadShift<T>); }

template <int N, typename T>
struct SP {
  static constexpr RadixList rl = make_radices(N);
  static constexpr int len = N + ((N - 1) >> kPadShift<T>) + 1;
  static constexpr int pitch = len | 1;
  static constexpr int ns(int p) {
    int s = 1;
    for (int i = 0; i < p; ++i) s *= rl.r[i];
    return s;
  }
  // start of pass p's twiddles in the per-pass table (static_twiddle_table)
  static constexpr int off(int p) {
    int o = 0;
    for (int i = 0; i < p; ++i) o += ns(i) * (rl.r[i] - 1);
    return o;
  }
};

// Marks a pass that reads or writes its pencils in LDS (no global hook).
struct LdsIo {};

template <typename F>
constexpr bool kLds = std::is_same<F, LdsIo>::value;

// Pass P of the Stockham sequence over TX pencils of length N.
//   inputs:  element j + r*NB of pencil c, from LDS or from ld(c, e);
//   outputs: element (j / Ns) * Ns * R + k + r*Ns, to LDS or to st(c, e, v).
// CF picks the lane mapping: pencil-fastest (c = b % TX, for columns that
// sit side by side in global memory) or element-fastest (c = b / NB).
// A pass whose inputs come from global memory needs no barrier before its
// LDS writes (nothing else reads LDS then); a pass writing global memory
// needs none after.  tid is the thread's index in the workgroup.
template <int N, int TX, int THREADS, int DIR, int P, bool CF, typename T, typename Ld, typename St>
__device__ __forceinline__ void spass(cx<T> *__restrict__ buf, const cx<T> *__restrict__ tw, const Ld &ld,
                                      const St &st, int tid) {
  constexpr int R = SP<N, T>::rl.r[P];
  constexpr int Ns = SP<N, T>::ns(P);
  constexpr int NB = N / R;
  constexpr int TOT = TX * NB;
  constexpr int TRIPS = (TOT + THREADS - 1) / THREADS;
  constexpr int PITCH = SP<N, T>::pitch;
  cx<T> v[TRIPS][R];
#pragma unroll
  for (int t = 0; t < TRIPS; ++t) {
    const int b = tid + t * THREADS;
    if (TOT % THREADS == 0 || b < TOT) {
      const int c = CF ? b % TX : b / NB, j = CF ? b / TX : b % NB;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (kLds<Ld>) v[t][r] = buf[c * PITCH + padx<T>(j + r * NB)];
        else v[t][r] = ld(c, j + r * NB);
      }
      if constexpr (Ns > 1) {
        // per-pass table [r-1][k]
        const int k = j % Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cx<T> w = tw[SP<N, T>::off(P) + (r - 1) * Ns + k];
          if (DIR > 0) w.y = -w.y;
          v[t][r] = mul(v[t][r], w);
        }
      }
      Dft<R, DIR, T>::run(v[t]);
    }
  }
  if constexpr (kLds<Ld> && kLds<St>) __syncthreads();
#pragma unroll
  for (int t = 0; t < TRIPS; ++t) {
    const int b = tid + t * THREADS;
    if (TOT % THREADS == 0 || b < TOT) {
      const int c = CF ? b % TX : b / NB, j = CF ? b / TX : b % NB;
      const int k = j % Ns;
      const int base = (j / Ns) * Ns * R + k;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (kLds<St>) buf[c * PITCH + padx<T>(base + r * Ns)] = v[t][r];
        else st(c, base + r * Ns, v[t][r]);
      }
    }
  }
  if constexpr (kLds<St>) __syncthreads();
}

// All passes.  The first pass takes its inputs from ld and the last writes
// through st (either may be LdsIo); CF0 / CFL are their lane mappings, the
// passes between are element-fastest.
template <int N, int TX, int THREADS, int DIR, int P, bool CF0, bool CFL, typename T, typename Ld, typename St>
__device__ __forceinline__ void spasses_io(cx<T> *__restrict__ buf, const cx<T> *__restrict__ tw, const Ld &ld,
                                           const St &st, int tid) {
  constexpr int CNT = SP<N, T>::rl.cnt;
  static_assert(CNT >= 2, "the specialised lengths have at least two passes");
  if constexpr (P < CNT) {
    if constexpr (P == 0) spass<N, TX, THREADS, DIR, P, CF0, T>(buf, tw, ld, LdsIo{}, tid);
    else if constexpr (P == CNT - 1) spass<N, TX, THREADS, DIR, P, CFL, T>(buf, tw, LdsIo{}, st, tid);
    else spass<N, TX, THREADS, DIR, P, false, T>(buf, tw, LdsIo{}, LdsIo{}, tid);
    spasses_io<N, TX, THREADS, DIR, P + 1, CF0, CFL, T>(buf, tw, ld, st, tid);
  }
}

template <int N, int TX, int THREADS, int DIR, int P, bool CF0, bool CFL, typename T, typename Ld, typename St>
__device__ __forceinline__ void spasses_io(cx<T> *__restrict__ buf, const cx<T> *__restrict__ tw, const Ld &ld,
                                           const St &st) {
  spasses_io<N, TX, THREADS, DIR, P, CF0, CFL, T>(buf, tw, ld, st, (int)threadIdx.x);
}

template <int N, int TX, int THREADS, int DIR, int P, typename T>
__device__ __forceinline__ void spasses(cx<T> *__restrict__ buf, const cx<T> *__restrict__ tw) {
  spasses_io<N, TX, THREADS, DIR, P, false, false, T>(buf, tw, LdsIo{}, LdsIo{});
}

template <int N, typename T>
constexpr size_t lds_bytes(int tx) {
  return sizeof(cx<T>) * ((size_t)N + (size_t)tx * SP<N, T>::pitch);
}

// Pencil-fastest passes without padding, for the fused per-column kernels.
// Thread tid keeps pencil c = tid % TX through every pass, working on its
// butterflies j = tid / TX + t (THREADS / TX); element e of its pencil is
// col[e ES] in LDS.  With ES = TX (element-major tile, buf[e TX + c]) the
// TX lanes that share a butterfly index touch TX consecutive elements,
// which spreads them over the banks with no padding, and every LDS offset a
// thread uses is a compile-time constant from its pencil's base `col` (no
// per-element index arithmetic).  Twiddles come from the same per-pass table
// as spass, in LDS or global memory.

// Pass P over the thread's pencil.  Inputs element j + r NB, from col or
// ld(e); outputs element (j / Ns) Ns R + j % Ns + r Ns, to col or st(e, v).
// SYNC puts a workgroup barrier between the reads and the writes (needed
// whenever both touch the pencils in LDS); a pass writing LDS ends with one.
template <int N, int TX, int THREADS, int DIR, int P, bool SYNC, int ES, typename T, typename Ld, typename St>
__device__ __forceinline__ void cpass(cx<T> *__restrict__ col, const cx<T> *__restrict__ tw, const Ld &ld,
                                      const St &st, int j0) {
  static_assert(THREADS % TX == 0, "a thread keeps one pencil");
  constexpr int R = SP<N, T>::rl.r[P];
  constexpr int Ns = SP<N, T>::ns(P);
  constexpr int NB = N / R;
  constexpr int JS = THREADS / TX;
  constexpr int TRIPS = (NB + JS - 1) / JS;
  cx<T> v[TRIPS][R];
#pragma unroll
  for (int t = 0; t < TRIPS; ++t) {
    const int j = j0 + t * JS;
    if (NB % JS == 0 || j < NB) {
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (kLds<Ld>) v[t][r] = col[(j + r * NB) * ES];
        else v[t][r] = ld(j + r * NB);
      }
      if constexpr (Ns > 1) {
        const int k = j % Ns;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          const cx<T> w = tw[SP<N, T>::off(P) + (r - 1) * Ns + k];
          v[t][r] = DIR > 0 ? mulc(v[t][r], w) : mul(v[t][r], w);
        }
      }
      Dft<R, DIR, T>::run(v[t]);
    }
  }
  if constexpr (SYNC) __syncthreads();
#pragma unroll
  for (int t = 0; t < TRIPS; ++t) {
    const int j = j0 + t * JS;
    if (NB % JS == 0 || j < NB) {
      const int base = (j / Ns) * Ns * R + j % Ns;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (kLds<St>) col[(base + r * Ns) * ES] = v[t][r];
        else st(base + r * Ns, v[t][r]);
      }
    }
  }
  if constexpr (kLds<St>) __syncthreads();
}

// All passes of one direction.  The first pass reads through ld (SYNC0: it
// reads LDS, so its writes wait for a barrier), the last writes through st;
// either may be LdsIo.
template <int N, int TX, int THREADS, int DIR, int P, bool SYNC0, int ES, typename T, typename Ld, typename St>
__device__ __forceinline__ void cpasses(cx<T> *__restrict__ col, const cx<T> *__restrict__ tw, const Ld &ld,
                                        const St &st, int j0) {
  constexpr int CNT = SP<N, T>::rl.cnt;
  static_assert(CNT >= 2, "the specialised lengths have at least two passes");
  if constexpr (P < CNT) {
    if constexpr (P == 0) cpass<N, TX, THREADS, DIR, P, SYNC0 || kLds<Ld>, ES, T>(col, tw, ld, LdsIo{}, j0);
    else if constexpr (P == CNT - 1) cpass<N, TX, THREADS, DIR, P, kLds<St>, ES, T>(col, tw, LdsIo{}, st, j0);
    else cpass<N, TX, THREADS, DIR, P, true, ES, T>(col, tw, LdsIo{}, LdsIo{}, j0);
    cpasses<N, TX, THREADS, DIR, P + 1, SYNC0, ES, T>(col, tw, ld, st, j0);
  }
}

}  // namespace crava
