// Stockham mixed-radix 3-D real FFT for gfx950 (see fft.hpp for the layout).
//
// One workgroup owns `tx` whole pencils in LDS.  Each radix-R pass reads R
// values per butterfly (stride n/R), applies the twiddles W_n^(r*k*n/(Ns*R)),
// runs the R-point DFT in registers and writes to (j/Ns)*Ns*R + k + r*Ns
// (self-sorting Stockham).  A workgroup has tx * (n / rmin) threads, so every
// thread owns at most one butterfly per pass.  Global loads/stores move tx
// contiguous columns per pencil element for the strided (y, z) passes and
// whole contiguous rows for the x pass.
#include "fft.hpp"
#include "../../include/crava_hip.h"

#include <cmath>
#include <vector>

namespace crava {

// ---------------------------------------------------------------------------
// Small DFTs.  DIR = -1 forward (FFTW_FORWARD), +1 backward.
// ---------------------------------------------------------------------------
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

// Odd prime radix by symmetric pairs: a_r = v_r + v_{R-r}, b_r = v_r - v_{R-r}.
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
    cx<T> x0 = v[0];
    cx<T> s0 = x0;
#pragma unroll
    for (int r = 0; r < H; ++r) s0 = add(s0, a[r]);
#pragma unroll
    for (int m = 1; m <= H; ++m) {
      cx<T> re = x0, im = mk((T)0, (T)0);
#pragma unroll
      for (int r = 1; r <= H; ++r) {
        const int q = (r * m) % R;
        const T c = (T)cos(2.0 * M_PI * q / R);
        const T s = (T)sin(2.0 * M_PI * q / R);
        re = add(re, scl(a[r - 1], c));
        im = add(im, scl(b[r - 1], s));
      }
      // X_m = re + DIR*i*im, X_{R-m} = re - DIR*i*im
      cx<T> t = rot_i<DIR>(im);
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

// One radix-R Stockham pass over the tx pencils held in LDS.
template <int R, int DIR, typename T>
__device__ __forceinline__ void stockham_pass(cx<T> *buf, const cx<T> *tw, int n, int ldsp, int Ns,
                                              int c, int j, bool in_range) {
  const int nb = n / R;
  const bool act = in_range && j < nb;
  cx<T> v[R];
  int k = 0;
  if (act) {
    k = j % Ns;
    const int tstep = k * (n / (Ns * R));
    const cx<T> *src = buf + c * ldsp + j;
#pragma unroll
    for (int r = 0; r < R; ++r) v[r] = src[r * nb];
#pragma unroll
    for (int r = 1; r < R; ++r) {
      cx<T> w = tw[r * tstep];
      if (DIR > 0) w.y = -w.y;
      v[r] = mul(v[r], w);
    }
    Dft<R, DIR, T>::run(v);
  }
  __syncthreads();
  if (act) {
    cx<T> *dst = buf + c * ldsp + (j / Ns) * Ns * R + k;
#pragma unroll
    for (int r = 0; r < R; ++r) dst[r * Ns] = v[r];
  }
  __syncthreads();
}

template <int DIR, typename T>
__device__ void run_passes(cx<T> *buf, const cx<T> *tw, const PassPlan &pl) {
  const int tpp = pl.n / pl.rmin;
  const int c = threadIdx.x / tpp;
  const int j = threadIdx.x - c * tpp;
  const bool in_range = c < pl.tx;
  int Ns = 1;
  for (int p = 0; p < pl.nrad; ++p) {
    const int R = pl.rad[p];
    switch (R) {
      case 2: stockham_pass<2, DIR>(buf, tw, pl.n, pl.ldsp, Ns, c, j, in_range); break;
      case 3: stockham_pass<3, DIR>(buf, tw, pl.n, pl.ldsp, Ns, c, j, in_range); break;
      case 4: stockham_pass<4, DIR>(buf, tw, pl.n, pl.ldsp, Ns, c, j, in_range); break;
      case 5: stockham_pass<5, DIR>(buf, tw, pl.n, pl.ldsp, Ns, c, j, in_range); break;
      case 7: stockham_pass<7, DIR>(buf, tw, pl.n, pl.ldsp, Ns, c, j, in_range); break;
      default: stockham_pass<8, DIR>(buf, tw, pl.n, pl.ldsp, Ns, c, j, in_range); break;
    }
    Ns *= R;
  }
}

// ---------------------------------------------------------------------------
// Strided complex pass (y or z).  Columns are contiguous; elements `es`
// apart; blockIdx.y selects a batch `bstride` apart.
// ---------------------------------------------------------------------------
template <int DIR, typename T>
__global__ void __launch_bounds__(1024) k_fft_strided(cx<T> *__restrict__ data, long long es, long long ncols,
                                                      long long bstride, PassPlan pl,
                                                      const cx<T> *__restrict__ twg) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + pl.n;
  const int nthr = blockDim.x;
  const long long col0 = (long long)blockIdx.x * pl.tx;
  const int ncol = (int)min((long long)pl.tx, ncols - col0);
  cx<T> *base = data + (long long)blockIdx.y * bstride + col0;
  for (int q = threadIdx.x; q < pl.n; q += nthr) tw[q] = twg[q];
  const int tot = pl.tx * pl.n;
  for (int q = threadIdx.x; q < tot; q += nthr) {
    const int e = q / pl.tx;
    const int c = q - e * pl.tx;
    if (c < ncol) buf[c * pl.ldsp + e] = base[c + e * es];
  }
  __syncthreads();
  run_passes<DIR>(buf, tw, pl);
  for (int q = threadIdx.x; q < tot; q += nthr) {
    const int e = q / pl.tx;
    const int c = q - e * pl.tx;
    if (c < ncol) base[c + e * es] = buf[c * pl.ldsp + e];
  }
}

// ---------------------------------------------------------------------------
// x pass, forward: real rows (nxp values, row pitch rnxp reals) -> half
// spectra (cnxp complex) in place.  Pencil c holds rows 2c (real part) and
// 2c+1 (imaginary part); the two spectra are separated afterwards with
// A_k = (Z_k + conj Z_{n-k})/2, B_k = (Z_k - conj Z_{n-k})/(2i).
// ---------------------------------------------------------------------------
template <typename T>
__global__ void __launch_bounds__(1024) k_fft_rows_r2c(T *__restrict__ data, long long nrows, int rnxp,
                                                       PassPlan pl, const cx<T> *__restrict__ twg, T scale) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + pl.n;
  const int n = pl.n;
  const int nthr = blockDim.x;
  const long long row0 = (long long)blockIdx.x * 2 * pl.tx;
  for (int q = threadIdx.x; q < n; q += nthr) tw[q] = twg[q];
  const int tot = 2 * pl.tx * n;
  for (int q = threadIdx.x; q < tot; q += nthr) {
    const int rl = q / n;
    const int e = q - rl * n;
    const long long row = row0 + rl;
    T v = row < nrows ? data[row * rnxp + e] : (T)0;
    T *dst = reinterpret_cast<T *>(buf + (rl >> 1) * pl.ldsp + e);
    dst[rl & 1] = v;
  }
  __syncthreads();
  run_passes<-1>(buf, tw, pl);
  const int cn = n / 2 + 1;
  const T hs = (T)0.5 * scale;
  const int tot2 = pl.tx * cn;
  for (int q = threadIdx.x; q < tot2; q += nthr) {
    const int c = q / cn;
    const int k = q - c * cn;
    const cx<T> z = buf[c * pl.ldsp + k];
    const cx<T> zm = buf[c * pl.ldsp + (k == 0 ? 0 : n - k)];
    const long long ra = row0 + 2 * c;
    if (ra < nrows) {
      cx<T> *pa = reinterpret_cast<cx<T> *>(data + ra * rnxp);
      pa[k] = mk(hs * (z.x + zm.x), hs * (z.y - zm.y));
    }
    if (ra + 1 < nrows) {
      cx<T> *pb = reinterpret_cast<cx<T> *>(data + (ra + 1) * rnxp);
      pb[k] = mk(hs * (z.y + zm.y), hs * (zm.x - z.x));
    }
  }
}

// x pass, inverse: half spectra of rows a, b -> Z = A + iB on the full length
// (Hermitian extension, Im of bins 0 and n/2 ignored as in a halfcomplex
// c2r), inverse DFT, real part -> row a, imaginary part -> row b.
template <typename T>
__global__ void __launch_bounds__(1024) k_fft_rows_c2r(T *__restrict__ data, long long nrows, int rnxp,
                                                       PassPlan pl, const cx<T> *__restrict__ twg, T scale) {
  extern __shared__ unsigned char smem_raw[];
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + pl.n;
  const int n = pl.n;
  const int cn = n / 2 + 1;
  const int nthr = blockDim.x;
  const long long row0 = (long long)blockIdx.x * 2 * pl.tx;
  for (int q = threadIdx.x; q < n; q += nthr) tw[q] = twg[q];
  const int tot = pl.tx * n;
  for (int q = threadIdx.x; q < tot; q += nthr) {
    const int c = q / n;
    const int k = q - c * n;
    const bool mirror = k >= cn;
    const int kk = mirror ? n - k : k;
    const bool selfc = (kk == 0) || ((n & 1) == 0 && kk == n / 2);
    const long long ra = row0 + 2 * c;
    cx<T> a = mk((T)0, (T)0), b = mk((T)0, (T)0);
    if (ra < nrows) a = reinterpret_cast<const cx<T> *>(data + ra * rnxp)[kk];
    if (ra + 1 < nrows) b = reinterpret_cast<const cx<T> *>(data + (ra + 1) * rnxp)[kk];
    if (selfc) { a.y = (T)0; b.y = (T)0; }
    if (mirror) { a.y = -a.y; b.y = -b.y; }
    buf[c * pl.ldsp + k] = mk(a.x - b.y, a.y + b.x);
  }
  __syncthreads();
  run_passes<+1>(buf, tw, pl);
  const int tot2 = 2 * pl.tx * n;
  for (int q = threadIdx.x; q < tot2; q += nthr) {
    const int rl = q / n;
    const int e = q - rl * n;
    const long long row = row0 + rl;
    if (row < nrows) {
      const T *src = reinterpret_cast<const T *>(buf + (rl >> 1) * pl.ldsp + e);
      data[row * rnxp + e] = src[rl & 1] * scale;
    }
  }
}

// ---------------------------------------------------------------------------
// Host side: plans and launches.
// ---------------------------------------------------------------------------
bool factor_radices(int n, PassPlan &p) {
  p.n = n;
  p.nrad = 0;
  if (n < 1) return false;
  int m = n, a = 0;
  while (m % 2 == 0) { m /= 2; ++a; }
  std::vector<int> r;
  const int q = a / 3, rem = a % 3;
  if (rem == 0) { for (int i = 0; i < q; ++i) r.push_back(8); }
  else if (rem == 1) {
    if (q >= 1) { for (int i = 0; i < q - 1; ++i) r.push_back(8); r.push_back(4); r.push_back(4); }
    else r.push_back(2);
  } else { for (int i = 0; i < q; ++i) r.push_back(8); r.push_back(4); }
  for (int pf : {7, 5, 3})
    while (m % pf == 0) { m /= pf; r.push_back(pf); }
  if (m != 1 || r.size() > 24) return false;
  p.nrad = (int)r.size();
  p.rmin = 8;
  for (int i = 0; i < p.nrad; ++i) { p.rad[i] = r[i]; p.rmin = std::min(p.rmin, r[i]); }
  if (p.nrad == 0) p.rmin = 1;
  return true;
}

static int max_lds_per_block() {
  static int v = -1;
  if (v < 0) {
    int dev = 0, val = 0;
    if (hipGetDevice(&dev) == hipSuccess &&
        hipDeviceGetAttribute(&val, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && val > 0)
      v = val;
    else
      v = 65536;
  }
  return v;
}

Fft3D::Fft3D(int nx, int ny, int nz, bool d) : nxp(nx), nyp(ny), nzp(nz), dbl(d) {}

Fft3D::~Fft3D() {
  if (twx) (void)hipFree(twx);
  if (twy) (void)hipFree(twy);
  if (twz) (void)hipFree(twz);
}

int Fft3D::build_plan(int n, PassPlan &p, int want_tx) {
  if (!factor_radices(n, p)) {
    set_error("FFT length " + std::to_string(n) + " is not 7-smooth");
    return CRAVA_HIP_E_INVALID;
  }
  const size_t eb = dbl ? 16 : 8;
  const int tpp = std::max(1, n / p.rmin);
  if (tpp > 1024) {
    set_error("FFT length " + std::to_string(n) + " too long for one workgroup");
    return CRAVA_HIP_E_INVALID;
  }
  p.ldsp = (n % 2 == 0) ? n + 1 : n;
  // keep two workgroups per CU: at most 80 KiB (or the device limit)
  const size_t budget = std::min<size_t>((size_t)max_lds_per_block(), 80 * 1024);
  int tx_lds = (int)((budget / eb - (size_t)n) / (size_t)p.ldsp);
  int tx_thr = 1024 / tpp;
  int tx = std::min(tx_lds, tx_thr);
  if (want_tx > 0) tx = std::min(tx, want_tx);
  tx = std::min(tx, 64);
  if (tx < 1) {
    set_error("FFT length " + std::to_string(n) + " does not fit in LDS");
    return CRAVA_HIP_E_INVALID;
  }
  p.tx = tx;
  int threads = tx * tpp;
  p.threads = ((threads + 63) / 64) * 64;
  p.lds_bytes = eb * ((size_t)n + (size_t)tx * p.ldsp);
  return CRAVA_HIP_OK;
}

template <typename T>
static int upload_twiddles(int n, void **dst) {
  std::vector<cx<T>> h(n);
  for (int m = 0; m < n; ++m) {
    // exp(-2 pi i m / n), argument reduced to the first octant for accuracy
    long double a = -2.0L * 3.14159265358979323846264338327950288L * (long double)m / (long double)n;
    h[m].x = (T)cosl(a);
    h[m].y = (T)sinl(a);
  }
  if (*dst) (void)hipFree(*dst);
  *dst = nullptr;
  CRAVA_HIP_CHECK(hipMalloc(dst, sizeof(cx<T>) * n));
  CRAVA_HIP_CHECK(hipMemcpy(*dst, h.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
  return CRAVA_HIP_OK;
}

int Fft3D::init() {
  if (nxp < 2 || nyp < 1 || nzp < 1) {
    set_error("invalid FFT grid size");
    return CRAVA_HIP_E_INVALID;
  }
  int rc;
  if ((rc = build_plan(nxp, px, 0)) || (rc = build_plan(nyp, py, 0)) || (rc = build_plan(nzp, pz, 0))) return rc;
  if (dbl) {
    if ((rc = upload_twiddles<double>(nxp, &twx)) || (rc = upload_twiddles<double>(nyp, &twy)) ||
        (rc = upload_twiddles<double>(nzp, &twz)))
      return rc;
  } else {
    if ((rc = upload_twiddles<float>(nxp, &twx)) || (rc = upload_twiddles<float>(nyp, &twy)) ||
        (rc = upload_twiddles<float>(nzp, &twz)))
      return rc;
  }
  return CRAVA_HIP_OK;
}

int Fft3D::set_tx(int tx_x, int tx_y, int tx_z) {
  int rc;
  if ((rc = build_plan(nxp, px, tx_x)) || (rc = build_plan(nyp, py, tx_y)) || (rc = build_plan(nzp, pz, tx_z))) return rc;
  return CRAVA_HIP_OK;
}

template <typename F>
static int set_lds(F kernel, size_t bytes) {
  if (bytes > 65536) {
    CRAVA_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  }
  return CRAVA_HIP_OK;
}

template <typename T>
static int launch_x(const Fft3D &f, void *grid, hipStream_t s, bool inverse, double scale) {
  const long long nrows = (long long)f.nyp * f.nzp;
  const int rnxp = 2 * (f.nxp / 2 + 1);
  const PassPlan &p = f.px;
  const unsigned blocks = (unsigned)((nrows + 2LL * p.tx - 1) / (2LL * p.tx));
  int rc;
  if (!inverse) {
    if ((rc = set_lds(k_fft_rows_r2c<T>, p.lds_bytes))) return rc;
    hipLaunchKernelGGL(k_fft_rows_r2c<T>, dim3(blocks), dim3(p.threads), p.lds_bytes, s,
                       reinterpret_cast<T *>(grid), nrows, rnxp, p, reinterpret_cast<const cx<T> *>(f.twx), (T)scale);
  } else {
    if ((rc = set_lds(k_fft_rows_c2r<T>, p.lds_bytes))) return rc;
    hipLaunchKernelGGL(k_fft_rows_c2r<T>, dim3(blocks), dim3(p.threads), p.lds_bytes, s,
                       reinterpret_cast<T *>(grid), nrows, rnxp, p, reinterpret_cast<const cx<T> *>(f.twx), (T)scale);
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

template <typename T>
static int launch_strided(const PassPlan &p, const void *tw, void *grid, hipStream_t s, bool inverse,
                          long long es, long long ncols, long long nbatch, long long bstride) {
  if (p.n <= 1) return CRAVA_HIP_OK;
  const unsigned gx = (unsigned)((ncols + p.tx - 1) / p.tx);
  int rc;
  if (!inverse) {
    if ((rc = set_lds(k_fft_strided<-1, T>, p.lds_bytes))) return rc;
    hipLaunchKernelGGL((k_fft_strided<-1, T>), dim3(gx, (unsigned)nbatch), dim3(p.threads), p.lds_bytes, s,
                       reinterpret_cast<cx<T> *>(grid), es, ncols, bstride, p, reinterpret_cast<const cx<T> *>(tw));
  } else {
    if ((rc = set_lds(k_fft_strided<1, T>, p.lds_bytes))) return rc;
    hipLaunchKernelGGL((k_fft_strided<1, T>), dim3(gx, (unsigned)nbatch), dim3(p.threads), p.lds_bytes, s,
                       reinterpret_cast<cx<T> *>(grid), es, ncols, bstride, p, reinterpret_cast<const cx<T> *>(tw));
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int Fft3D::pass_x(void *grid, hipStream_t s, bool inverse, double scale) {
  return dbl ? launch_x<double>(*this, grid, s, inverse, scale) : launch_x<float>(*this, grid, s, inverse, scale);
}

int Fft3D::pass_y(void *grid, hipStream_t s, bool inverse) {
  const long long cnxp = nxp / 2 + 1;
  return dbl ? launch_strided<double>(py, twy, grid, s, inverse, cnxp, cnxp, nzp, cnxp * nyp)
             : launch_strided<float>(py, twy, grid, s, inverse, cnxp, cnxp, nzp, cnxp * nyp);
}

int Fft3D::pass_z(void *grid, hipStream_t s, bool inverse) {
  const long long cnxp = nxp / 2 + 1;
  const long long plane = cnxp * nyp;
  return dbl ? launch_strided<double>(pz, twz, grid, s, inverse, plane, plane, 1, 0)
             : launch_strided<float>(pz, twz, grid, s, inverse, plane, plane, 1, 0);
}

int Fft3D::r2c(void *grid, hipStream_t s, double scale) {
  int rc;
  if ((rc = pass_x(grid, s, false, scale))) return rc;
  if ((rc = pass_y(grid, s, false))) return rc;
  return pass_z(grid, s, false);
}

int Fft3D::c2r(void *grid, hipStream_t s, double scale) {
  int rc;
  if ((rc = pass_z(grid, s, true))) return rc;
  if ((rc = pass_y(grid, s, true))) return rc;
  return pass_x(grid, s, true, scale);
}

}  // namespace crava
