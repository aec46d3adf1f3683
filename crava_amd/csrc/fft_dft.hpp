// Small in-register DFTs (radix 2, 3, 4, 5, 7, 8) shared by the Stockham
// kernels.  DIR = -1 is FFTW's forward sign, +1 the backward one.
#pragma once
#include "common.hpp"

namespace crava {

template <int DIR, typename T>
__device__ __forceinline__ cx<T> rot_i(cx<T> a) {  // multiply by DIR * i
  return DIR < 0 ? mk(a.y, -a.x) : mk(-a.y, a.x);
}

template <int R, int DIR, typename T>
struct Dft;

template <int DIR, typename T>
struct Dft<2, DIR, T> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    cx<T> a = v[0], b = v[1];
    v[0] = add(a, b);
    v[1] = sub(a, b);
  }
};

template <int DIR, typename T>
struct Dft<4, DIR, T> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    cx<T> t0 = add(v[0], v[2]), t1 = sub(v[0], v[2]);
    cx<T> t2 = add(v[1], v[3]), t3 = rot_i<DIR>(sub(v[1], v[3]));
    v[0] = add(t0, t2);
    v[2] = sub(t0, t2);
    v[1] = add(t1, t3);
    v[3] = sub(t1, t3);
  }
};

template <int DIR, typename T>
struct Dft<8, DIR, T> {
  static __device__ __forceinline__ void run(cx<T> *v) {
    const T h = (T)0.70710678118654752440084436210484903928;
    cx<T> e[4] = {v[0], v[2], v[4], v[6]};
    cx<T> o[4] = {v[1], v[3], v[5], v[7]};
    Dft<4, DIR, T>::run(e);
    Dft<4, DIR, T>::run(o);
    // W8^1 = (1 + DIR i)/sqrt2, W8^2 = DIR i, W8^3 = (-1 + DIR i)/sqrt2
    cx<T> o1 = DIR < 0 ? mk(h * (o[1].x + o[1].y), h * (o[1].y - o[1].x))
                       : mk(h * (o[1].x - o[1].y), h * (o[1].y + o[1].x));
    cx<T> o2 = rot_i<DIR>(o[2]);
    cx<T> o3 = DIR < 0 ? mk(h * (o[3].y - o[3].x), -h * (o[3].x + o[3].y))
                       : mk(-h * (o[3].x + o[3].y), h * (o[3].x - o[3].y));
    v[0] = add(e[0], o[0]);
    v[4] = sub(e[0], o[0]);
    v[1] = add(e[1], o1);
    v[5] = sub(e[1], o1);
    v[2] = add(e[2], o2);
    v[6] = sub(e[2], o2);
    v[3] = add(e[3], o3);
    v[7] = sub(e[3], o3);
  }
};

// cos(2 pi q / R), sin(2 pi q / R) for the odd prime radices
template <int R>
struct Roots;
template <>
struct Roots<3> {
  static constexpr double c[3] = {1.0, -0.5, -0.5};
  static constexpr double s[3] = {0.0, 0.86602540378443864676372317075293618, -0.86602540378443864676372317075293618};
};
template <>
struct Roots<5> {
  static constexpr double c[5] = {1.0, 0.30901699437494742410229341718281906, -0.80901699437494742410229341718281906,
                                  -0.80901699437494742410229341718281906, 0.30901699437494742410229341718281906};
  static constexpr double s[5] = {0.0, 0.95105651629515357211643933337938214, 0.58778525229247312916870595463907277,
                                  -0.58778525229247312916870595463907277, -0.95105651629515357211643933337938214};
};
template <>
struct Roots<7> {
  static constexpr double c[7] = {1.0, 0.62348980185873353052500488400423981, -0.22252093395631440428890256449679476,
                                  -0.90096886790241912623610231950744505, -0.90096886790241912623610231950744505,
                                  -0.22252093395631440428890256449679476, 0.62348980185873353052500488400423981};
  static constexpr double s[7] = {0.0, 0.78183148246802980870844452667405775, 0.97492791218182360701813168299393122,
                                  0.43388373911755812047576833284835875, -0.43388373911755812047576833284835875,
                                  -0.97492791218182360701813168299393122, -0.78183148246802980870844452667405775};
};

// Odd prime radix by symmetric pairs a_r = v_r + v_{R-r}, b_r = v_r - v_{R-r}:
// X_m = v0 + sum_r a_r cos(2 pi r m / R) + DIR i sum_r b_r sin(2 pi r m / R).
template <int R, int DIR, typename T>
struct DftOdd {
  static __device__ __forceinline__ void run(cx<T> *v) {
    constexpr int H = (R - 1) / 2;
    cx<T> a[H], b[H];
#pragma unroll
    for (int r = 1; r <= H; ++r) {
      a[r - 1] = add(v[r], v[R - r]);
      b[r - 1] = sub(v[r], v[R - r]);
    }
    const cx<T> x0 = v[0];
    cx<T> s0 = x0;
#pragma unroll
    for (int r = 0; r < H; ++r) s0 = add(s0, a[r]);
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      cx<T> re = x0, im = mk((T)0, (T)0);
#pragma unroll
      for (int r = 1; r <= H; ++r) {
        const int q = (r * m) % R;
        re = add(re, scl(a[r - 1], (T)Roots<R>::c[q]));
        im = add(im, scl(b[r - 1], (T)Roots<R>::s[q]));
      }
      const cx<T> t = rot_i<DIR>(im);
      v[m] = add(re, t);
      v[R - m] = sub(re, t);
    }
    v[0] = s0;
  }
};
template <int DIR, typename T>
struct Dft<3, DIR, T> : DftOdd<3, DIR, T> {};
template <int DIR, typename T>
struct Dft<5, DIR, T> : DftOdd<5, DIR, T> {};
template <int DIR, typename T>
struct Dft<7, DIR, T> : DftOdd<7, DIR, T> {};

// Radix sequence of a 7-smooth length: radix 8 first, the power-of-two rest
// as 4s (never a lone 2 unless n = 2), then 7, 5, 3.
struct RadixList {
  int n = 1;
  int cnt = 0;
  int r[24] = {0};
  int rmin = 1;
  bool ok = true;
};

constexpr RadixList make_radices(int n) {
  RadixList L{};
  L.n = n;
  if (n < 1) { L.ok = false; return L; }
  int m = n, a = 0;
  while (m % 2 == 0) { m /= 2; ++a; }
  const int q = a / 3, rem = a % 3;
  if (rem == 0) {
    for (int i = 0; i < q; ++i) L.r[L.cnt++] = 8;
  } else if (rem == 1) {
    if (q >= 1) { for (int i = 0; i < q - 1; ++i) L.r[L.cnt++] = 8; L.r[L.cnt++] = 4; L.r[L.cnt++] = 4; }
    else L.r[L.cnt++] = 2;
  } else {
    for (int i = 0; i < q; ++i) L.r[L.cnt++] = 8;
    L.r[L.cnt++] = 4;
  }
  const int pf[3] = {7, 5, 3};
  for (int t = 0; t < 3; ++t)
    while (m % pf[t] == 0) { m /= pf[t]; L.r[L.cnt++] = pf[t]; }
  if (m != 1) L.ok = false;
  L.rmin = 8;
  for (int i = 0; i < L.cnt; ++i) L.rmin = L.r[i] < L.rmin ? L.r[i] : L.rmin;
  if (L.cnt == 0) L.rmin = 1;
  return L;
}

}  // namespace crava
