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
#include "fft_dft.hpp"
#include "fft_static.hpp"

#include <cmath>
#include <type_traits>
#include <vector>

namespace crava {

__device__ __forceinline__ int gpadx(int x) { return x + (x >> 5); }

// Global loads issued back to back per thread in the staging loops.
constexpr int kUnroll = 8;

// x / d with the plan's multiplier (d == 1 when mul == 0).
__device__ __forceinline__ int fdiv(int x, unsigned mul) { return mul ? (int)__umulhi((unsigned)x, mul) : x; }

// Marks a pass that reads its butterflies from LDS.
struct FromLds {};

// One radix-R Stockham pass over the tx pencils held in LDS.  Butterfly b
// (pencil b / NB, index j = b mod NB; with CF pencil b mod tx, index b / tx,
// txm the multiplier of tx) runs on thread b mod nthr; a workgroup has half
// as many threads as the widest pass has butterflies, so each thread runs up
// to two independent butterflies and the active ones fill whole waves.
// ld: where the pass reads its inputs -- FromLds, or a functor ld(c, e)
// giving element e of pencil c (the first pass of a kernel that loads its
// butterflies straight from global memory; no barrier is then needed before
// the LDS writes).
template <int R, int DIR, typename T, typename Ld = FromLds, bool CF = false>
__device__ __forceinline__ void stockham_pass(cx<T> *buf, const cx<T> *tw, const PassStep &st, int tx, int ldsp,
                                              const Ld &ld = Ld{}, unsigned txm = 0) {
  constexpr bool kFromLds = std::is_same<Ld, FromLds>::value;
  const int NB = st.NB, Ns = st.Ns;
  const int nthr = blockDim.x;
  const int tot = tx * NB;
  cx<T> v[2][R];
  int base[2] = {0, 0}, cb[2] = {0, 0};
  bool act[2];
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    const int b = (int)threadIdx.x + t * nthr;
    act[t] = b < tot;
    if (act[t]) {
      int c, j;
      if constexpr (CF) {
        j = fdiv(b, txm);
        c = b - j * tx;
      } else {
        c = fdiv(b, st.nb_mul);
        j = b - c * NB;
      }
      const cx<T> *src = buf + c * ldsp;
#pragma unroll
      for (int r = 0; r < R; ++r) {
        if constexpr (kFromLds) v[t][r] = src[gpadx(j + r * NB)];
        else v[t][r] = ld(c, j + r * NB);
      }
      int k = 0;
      if (Ns > 1) {
        k = j - fdiv(j, st.ns_mul) * Ns;
        const int ts = k * st.tstep;
#pragma unroll
        for (int r = 1; r < R; ++r) {
          cx<T> w = tw[r * ts];
          if (DIR > 0) w.y = -w.y;
          v[t][r] = mul(v[t][r], w);
        }
      }
      Dft<R, DIR, T>::run(v[t]);
      base[t] = (j - k) * R + k;   // (j / Ns) * Ns * R + k
      cb[t] = c * ldsp;
    }
  }
  if constexpr (kFromLds) __syncthreads();
#pragma unroll
  for (int t = 0; t < 2; ++t) {
    if (act[t]) {
      cx<T> *dst = buf + cb[t];
#pragma unroll
      for (int r = 0; r < R; ++r) dst[gpadx(base[t] + r * Ns)] = v[t][r];
    }
  }
  __syncthreads();
}

// All passes; the first takes its butterflies from ld unless that is
// FromLds, with pencil-fastest lanes when CF (strided columns side by side).
template <int DIR, typename T, typename Ld = FromLds, bool CF = false>
__device__ void run_passes(cx<T> *buf, const cx<T> *tw, const PassPlan &pl, const Ld &ld = Ld{}) {
  int p = 0;
  if constexpr (!std::is_same<Ld, FromLds>::value) {
    const PassStep &st = pl.step[0];
    const unsigned txm = pl.tx_mul;
    switch (st.R) {
      case 2: stockham_pass<2, DIR, T, Ld, CF>(buf, tw, st, pl.tx, pl.ldsp, ld, txm); break;
      case 3: stockham_pass<3, DIR, T, Ld, CF>(buf, tw, st, pl.tx, pl.ldsp, ld, txm); break;
      case 4: stockham_pass<4, DIR, T, Ld, CF>(buf, tw, st, pl.tx, pl.ldsp, ld, txm); break;
      case 5: stockham_pass<5, DIR, T, Ld, CF>(buf, tw, st, pl.tx, pl.ldsp, ld, txm); break;
      case 7: stockham_pass<7, DIR, T, Ld, CF>(buf, tw, st, pl.tx, pl.ldsp, ld, txm); break;
      default: stockham_pass<8, DIR, T, Ld, CF>(buf, tw, st, pl.tx, pl.ldsp, ld, txm); break;
    }
    p = 1;
  }
  for (; p < pl.nrad; ++p) {
    const PassStep &st = pl.step[p];
    switch (st.R) {
      case 2: stockham_pass<2, DIR>(buf, tw, st, pl.tx, pl.ldsp); break;
      case 3: stockham_pass<3, DIR>(buf, tw, st, pl.tx, pl.ldsp); break;
      case 4: stockham_pass<4, DIR>(buf, tw, st, pl.tx, pl.ldsp); break;
      case 5: stockham_pass<5, DIR>(buf, tw, st, pl.tx, pl.ldsp); break;
      case 7: stockham_pass<7, DIR>(buf, tw, st, pl.tx, pl.ldsp); break;
      default: stockham_pass<8, DIR>(buf, tw, st, pl.tx, pl.ldsp); break;
    }
  }
}

// ---------------------------------------------------------------------------
// Strided complex pass (y or z).  Columns are contiguous; elements `es`
// apart; a batch dimension `bstride` apart; blockIdx.z selects the grid.
// ---------------------------------------------------------------------------
// DIO: the first radix pass loads its butterflies straight from the columns
// (pencil-fastest lanes, branch-free clamped loads) instead of staging the
// pencils through LDS.
template <int DIR, typename T, typename SI, typename SO, bool DIO>
__global__ void __launch_bounds__(1024) k_fft_strided(GridSet gs, long long es, long long ncols, long long bstride,
                                                      PassPlan pl, const cx<T> *__restrict__ twg) {
  extern __shared__ unsigned char smem_raw[];
  // (column block, batch, grid) of this workgroup, XCD-contiguous
  const unsigned W = xcd_work_id(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z),
                                 gridDim.x * gridDim.y * gridDim.z);
  const unsigned bx = W % gridDim.x, by = (W / gridDim.x) % gridDim.y, bz = W / (gridDim.x * gridDim.y);
  const cx<SI> *src = reinterpret_cast<const cx<SI> *>(gs.src[bz]);
  cx<SO> *dst = reinterpret_cast<cx<SO> *>(gs.dst[bz]);
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + pl.n;
  const int nthr = blockDim.x;
  const long long col0 = (long long)bx * pl.tx;
  const int ncol = (int)min((long long)pl.tx, ncols - col0);
  const long long off = (long long)by * bstride + col0;
  const cx<SI> *ibase = src + off;
  cx<SO> *obase = dst + off;
  for (int q = threadIdx.x; q < pl.n; q += nthr) tw[q] = twg[q];
  const int tot = pl.tx * pl.n;
  if constexpr (DIO) {
    const auto ld = [&](int c, int e) -> cx<T> {
      const cx<SI> v = ibase[min(c, ncol - 1) + e * es];
      return c < ncol ? mk((T)v.x, (T)v.y) : mk((T)0, (T)0);
    };
    run_passes<DIR, T, decltype(ld), true>(buf, tw, pl, ld);
  } else {
  // kUnroll loads in flight per thread before they are stored to LDS: the
  // loads are branch-free (clamped addresses, validity kept aside) and stay
  // in their storage type until the LDS store, so none waits on another
  for (int q0 = threadIdx.x; q0 < tot; q0 += kUnroll * nthr) {
    cx<SI> v[kUnroll];
    int slot[kUnroll];
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int q = min(q0 + u * nthr, tot - 1);
      const int e = fdiv(q, pl.tx_mul);
      const int c = q - e * pl.tx;
      v[u] = ibase[min(c, ncol - 1) + e * es];
      slot[u] = (q0 + u * nthr < tot && c < ncol) ? c * pl.ldsp + gpadx(e) : -1;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      if (slot[u] >= 0) buf[slot[u]] = mk((T)v[u].x, (T)v[u].y);
  }
  __syncthreads();
  run_passes<DIR>(buf, tw, pl);
  }
#pragma unroll 4
  for (int q = threadIdx.x; q < tot; q += nthr) {
    const int e = fdiv(q, pl.tx_mul);
    const int c = q - e * pl.tx;
    if (c < ncol) {
      const cx<T> v = buf[c * pl.ldsp + gpadx(e)];
      obase[c + e * es] = mk((SO)v.x, (SO)v.y);
    }
  }
}

// ---------------------------------------------------------------------------
// x pass, forward: real rows (nxp values, row pitch rnxp reals) -> half
// spectra (cnxp complex) in place.  Pencil c holds rows 2c (real part) and
// 2c+1 (imaginary part); the two spectra are separated afterwards with
// A_k = (Z_k + conj Z_{n-k})/2, B_k = (Z_k - conj Z_{n-k})/(2i).
// ---------------------------------------------------------------------------
// DIO: the first radix pass loads its butterflies straight from the rows
// (branch-free, clamped rows) instead of staging the pencils through LDS.
template <typename T, typename SI, typename SO, bool DIO>
__global__ void __launch_bounds__(1024) k_fft_rows_r2c(GridSet gs, long long nrows, int rnxp, PassPlan pl,
                                                       const cx<T> *__restrict__ twg) {
  extern __shared__ unsigned char smem_raw[];
  // (row block, grid) of this workgroup, XCD-contiguous
  const unsigned W = xcd_work_id(blockIdx.x + gridDim.x * blockIdx.z, gridDim.x * gridDim.z);
  const unsigned bx = W % gridDim.x, bz = W / gridDim.x;
  const SI *src = reinterpret_cast<const SI *>(gs.src[bz]);
  SO *dst = reinterpret_cast<SO *>(gs.dst[bz]);
  const T scale = (T)gs.scale[bz];
  // mean deflation: subtract c in storage precision (exact near c)
  const SI off = (gs.dc && ((gs.dcmask >> bz) & 1u)) ? (SI)gs.dc[bz] : (SI)0;
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + pl.n;
  const int n = pl.n;
  const int nthr = blockDim.x;
  const long long row0 = (long long)bx * 2 * pl.tx;
  for (int q = threadIdx.x; q < n; q += nthr) tw[q] = twg[q];
  if constexpr (DIO) {
    const auto ld = [&](int c, int e) -> cx<T> {
      const long long ra = row0 + 2 * c;
      const SI a = src[min(ra, nrows - 1) * rnxp + e] - off;
      const SI b = src[min(ra + 1, nrows - 1) * rnxp + e] - off;
      return mk(ra < nrows ? (T)a : (T)0, ra + 1 < nrows ? (T)b : (T)0);
    };
    run_passes<-1>(buf, tw, pl, ld);
  } else {
  // element e of pencil c is (row 2c, row 2c+1) at e: one whole complex LDS
  // store per lane
  const int tot = pl.tx * n;
  for (int q0 = threadIdx.x; q0 < tot; q0 += kUnroll * nthr) {
    SI va[kUnroll], vb[kUnroll];
    int slot[kUnroll];
    unsigned okb = 0;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int q = min(q0 + u * nthr, tot - 1);
      const int c = fdiv(q, pl.n_mul);
      const int e = q - c * n;
      const long long ra = row0 + 2 * c;
      va[u] = src[min(ra, nrows - 1) * rnxp + e] - off;
      vb[u] = src[min(ra + 1, nrows - 1) * rnxp + e] - off;
      okb |= (ra < nrows ? 1u : 0u) << (2 * u);
      okb |= (ra + 1 < nrows ? 1u : 0u) << (2 * u + 1);
      slot[u] = q0 + u * nthr < tot ? c * pl.ldsp + gpadx(e) : -1;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u)
      if (slot[u] >= 0)
        buf[slot[u]] = mk((okb >> (2 * u)) & 1u ? (T)va[u] : (T)0, (okb >> (2 * u + 1)) & 1u ? (T)vb[u] : (T)0);
  }
  __syncthreads();
  run_passes<-1>(buf, tw, pl);
  }
  const int cn = n / 2 + 1;
  const T hs = (T)0.5 * scale;
  const int tot2 = pl.tx * cn;
  for (int q = threadIdx.x; q < tot2; q += nthr) {
    const int c = fdiv(q, pl.cn_mul);
    const int k = q - c * cn;
    const cx<T> z = buf[c * pl.ldsp + gpadx(k)];
    const cx<T> zm = buf[c * pl.ldsp + gpadx(k == 0 ? 0 : n - k)];
    const long long ra = row0 + 2 * c;
    if (ra < nrows) {
      cx<SO> *pa = reinterpret_cast<cx<SO> *>(dst + ra * rnxp);
      pa[k] = mk((SO)(hs * (z.x + zm.x)), (SO)(hs * (z.y - zm.y)));
    }
    if (ra + 1 < nrows) {
      cx<SO> *pb = reinterpret_cast<cx<SO> *>(dst + (ra + 1) * rnxp);
      pb[k] = mk((SO)(hs * (z.y + zm.y)), (SO)(hs * (zm.x - z.x)));
    }
  }
}

// x pass, inverse: half spectra of rows a, b -> Z = A + iB on the full length
// (Hermitian extension, Im of bins 0 and n/2 ignored as in a halfcomplex
// c2r), inverse DFT, real part -> row a, imaginary part -> row b.
template <typename T, typename SI, typename SO, bool DIO>
__global__ void __launch_bounds__(1024) k_fft_rows_c2r(GridSet gs, long long nrows, int rnxp, PassPlan pl,
                                                       const cx<T> *__restrict__ twg) {
  extern __shared__ unsigned char smem_raw[];
  // (row block, grid) of this workgroup, XCD-contiguous
  const unsigned W = xcd_work_id(blockIdx.x + gridDim.x * blockIdx.z, gridDim.x * gridDim.z);
  const unsigned bx = W % gridDim.x, bz = W / gridDim.x;
  const SI *src = reinterpret_cast<const SI *>(gs.src[bz]);
  SO *dst = reinterpret_cast<SO *>(gs.dst[bz]);
  const T scale = (T)gs.scale[bz];
  // mean deflation: the level removed from the DC bin returns here
  const T add = (gs.dc && ((gs.dcmask >> bz) & 1u)) ? (T)gs.dc[bz] : (T)0;
  cx<T> *tw = reinterpret_cast<cx<T> *>(smem_raw);
  cx<T> *buf = tw + pl.n;
  const int n = pl.n;
  const int cn = n / 2 + 1;
  const int nthr = blockDim.x;
  const long long row0 = (long long)bx * 2 * pl.tx;
  for (int q = threadIdx.x; q < n; q += nthr) tw[q] = twg[q];
  if constexpr (DIO) {
    const auto ld = [&](int c, int k) -> cx<T> {
      const bool mirror = k >= cn;
      const int kk = mirror ? n - k : k;
      const bool selfc = (kk == 0) || ((n & 1) == 0 && kk == n / 2);
      const long long ra = row0 + 2 * c;
      cx<SI> x = reinterpret_cast<const cx<SI> *>(src + min(ra, nrows - 1) * rnxp)[kk];
      cx<SI> y = reinterpret_cast<const cx<SI> *>(src + min(ra + 1, nrows - 1) * rnxp)[kk];
      if (ra >= nrows) x = mk((SI)0, (SI)0);
      if (ra + 1 >= nrows) y = mk((SI)0, (SI)0);
      if (selfc) { x.y = (SI)0; y.y = (SI)0; }
      if (mirror) { x.y = -x.y; y.y = -y.y; }
      return mk((T)x.x - (T)y.y, (T)x.y + (T)y.x);
    };
    run_passes<+1>(buf, tw, pl, ld);
  } else {
  const int tot = pl.tx * n;
  for (int q0 = threadIdx.x; q0 < tot; q0 += kUnroll * nthr) {
    cx<SI> a[kUnroll], b[kUnroll];
    int slot[kUnroll];
    unsigned okb = 0;
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      const int q = min(q0 + u * nthr, tot - 1);
      const int c = fdiv(q, pl.n_mul);
      const int k = q - c * n;
      const bool mirror = k >= cn;
      const int kk = mirror ? n - k : k;
      const long long ra = row0 + 2 * c;
      a[u] = reinterpret_cast<const cx<SI> *>(src + min(ra, nrows - 1) * rnxp)[kk];
      b[u] = reinterpret_cast<const cx<SI> *>(src + min(ra + 1, nrows - 1) * rnxp)[kk];
      okb |= (ra < nrows ? 1u : 0u) << (2 * u);
      okb |= (ra + 1 < nrows ? 1u : 0u) << (2 * u + 1);
      slot[u] = q0 + u * nthr < tot ? c * pl.ldsp + gpadx(k) : -1;
    }
#pragma unroll
    for (int u = 0; u < kUnroll; ++u) {
      if (slot[u] < 0) continue;
      const int k = (q0 + u * nthr) - fdiv(q0 + u * nthr, pl.n_mul) * n;
      const bool mirror = k >= cn;
      const int kk = mirror ? n - k : k;
      const bool selfc = (kk == 0) || ((n & 1) == 0 && kk == n / 2);
      cx<SI> x = a[u], y = b[u];
      if (!((okb >> (2 * u)) & 1u)) x = mk((SI)0, (SI)0);
      if (!((okb >> (2 * u + 1)) & 1u)) y = mk((SI)0, (SI)0);
      if (selfc) { x.y = (SI)0; y.y = (SI)0; }
      if (mirror) { x.y = -x.y; y.y = -y.y; }
      buf[slot[u]] = mk((T)x.x - (T)y.y, (T)x.y + (T)y.x);
    }
  }
  __syncthreads();
  run_passes<+1>(buf, tw, pl);
  }
  // element e of pencil c -> real part to row 2c, imaginary part to row 2c+1
  const int tot2 = pl.tx * n;
  for (int q = threadIdx.x; q < tot2; q += nthr) {
    const int c = fdiv(q, pl.n_mul);
    const int e = q - c * n;
    const long long ra = row0 + 2 * c;
    const cx<T> v = buf[c * pl.ldsp + gpadx(e)];
    if (ra < nrows) dst[ra * rnxp + e] = (SO)(v.x * scale + add);
    if (ra + 1 < nrows) dst[(ra + 1) * rnxp + e] = (SO)(v.y * scale + add);
  }
}

// ---------------------------------------------------------------------------
// Host side: plans and launches.
// ---------------------------------------------------------------------------
bool factor_radices(int n, PassPlan &p, int order) {
  const RadixList L = make_radices(n);
  p.n = n;
  p.nrad = L.cnt;
  p.rmin = L.rmin;
  for (int i = 0; i < L.cnt; ++i) p.rad[i] = order ? L.r[L.cnt - 1 - i] : L.r[i];
  return L.ok;
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

// Generic-kernel variant g: radix order (g / 3: 0 large radices first, 1
// small first) x cap on pencils per workgroup (g % 3: 64, 16, 8).
static const int kGenericTxCap[3] = {64, 16, 8};

Fft3D::Fft3D(int nx, int ny, int nz, int p) : nxp(nx), nyp(ny), nzp(nz), prec(p), dbl(p >= 1) {}

Fft3D::~Fft3D() {
  if (twx) (void)hipFree(twx);
  if (twy) (void)hipFree(twy);
  if (twz) (void)hipFree(twz);
  for (void *t : stw)
    if (t) (void)hipFree(t);
}

int Fft3D::build_plan(int n, PassPlan &p, int g) {
  if (!factor_radices(n, p, g / 3)) {
    set_error("FFT length " + std::to_string(n) + " is not 7-smooth");
    return CRAVA_HIP_E_INVALID;
  }
  const size_t eb = dbl ? 16 : 8;
  const int tpp = std::max(1, n / p.rmin);   // butterflies of the widest pass, per pencil
  if (tpp > 1024) {
    set_error("FFT length " + std::to_string(n) + " too long for one workgroup");
    return CRAVA_HIP_E_INVALID;
  }
  // one pad slot every 32 elements, odd pitch (as the specialised kernels)
  p.ldsp = (n + (n - 1) / 32 + 1) | 1;
  // keep two workgroups per CU: at most 80 KiB (or the device limit); up to
  // two butterflies per thread; a power-of-two pencil count keeps the
  // strided passes' row segments aligned
  const size_t budget = std::min<size_t>((size_t)max_lds_per_block(), 80 * 1024);
  const int tx_lds = (int)((budget / eb - (size_t)n) / (size_t)p.ldsp);
  const int tx_thr = 2048 / tpp;
  int tx = std::min(std::min(tx_lds, tx_thr), kGenericTxCap[g % 3]);
  if (tx < 1) {
    set_error("FFT length " + std::to_string(n) + " does not fit in LDS");
    return CRAVA_HIP_E_INVALID;
  }
  while (tx & (tx - 1)) tx &= tx - 1;
  p.tx = tx;
  const int threads = (tx * tpp + 1) / 2;
  p.threads = std::max(64, ((threads + 63) / 64) * 64);
  p.lds_bytes = eb * ((size_t)n + (size_t)tx * p.ldsp);
  p.tx_mul = div_mul(tx);
  p.n_mul = div_mul(n);
  p.cn_mul = div_mul(n / 2 + 1);
  int ns = 1;
  for (int i = 0; i < p.nrad; ++i) {
    PassStep &st = p.step[i];
    st.R = p.rad[i];
    st.Ns = ns;
    st.NB = n / st.R;
    st.tstep = n / (ns * st.R);
    st.ns_mul = div_mul(ns);
    st.nb_mul = div_mul(st.NB);
    ns *= st.R;
  }
  return CRAVA_HIP_OK;
}

template <typename T>
static int upload_static_twiddles(int n, void **dst) {
  std::vector<double> h(2 * (size_t)n);
  if (!static_twiddle_table(n, h.data())) return CRAVA_HIP_OK;
  std::vector<cx<T>> t(n);
  for (int i = 0; i < n; ++i) t[i] = mk((T)h[2 * i], (T)h[2 * i + 1]);
  CRAVA_HIP_CHECK(hipMalloc(dst, sizeof(cx<T>) * n));
  CRAVA_HIP_CHECK(hipMemcpy(*dst, t.data(), sizeof(cx<T>) * n, hipMemcpyHostToDevice));
  return CRAVA_HIP_OK;
}

template <typename T>
static int upload_twiddles(int n, void **dst) {
  std::vector<cx<T>> h(n);
  for (int m = 0; m < n; ++m) {
    // exp(-2 pi i m / n) evaluated in long double
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
        (rc = upload_twiddles<double>(nzp, &twz)) || (rc = upload_static_twiddles<double>(nxp, &stw[0])) ||
        (rc = upload_static_twiddles<double>(nyp, &stw[1])) || (rc = upload_static_twiddles<double>(nzp, &stw[2])))
      return rc;
  } else {
    if ((rc = upload_twiddles<float>(nxp, &twx)) || (rc = upload_twiddles<float>(nyp, &twy)) ||
        (rc = upload_twiddles<float>(nzp, &twz)) || (rc = upload_static_twiddles<float>(nxp, &stw[0])) ||
        (rc = upload_static_twiddles<float>(nyp, &stw[1])) || (rc = upload_static_twiddles<float>(nzp, &stw[2])))
      return rc;
  }
  // specialised kernels where a length has them (until tuned: variant 1,
  // else the first of 0, 5, 4, 3 the length has)
  for (int axis = 0; axis < 3; ++axis) {
    var[axis] = -1;
    if (num_variants(axis) < 1) continue;
    for (int v : {1, 0, 5, 4, 3})
      if (set_variant(axis, v) == CRAVA_HIP_OK) break;
  }
  set_error("");
  return CRAVA_HIP_OK;
}

long long Fft3D::complex_elems() const { return (long long)(nxp / 2 + 1) * nyp * nzp; }

template <typename F>
static int set_lds(F kernel, size_t bytes) {
  if (bytes > 65536) {
    CRAVA_HIP_CHECK(hipFuncSetAttribute(reinterpret_cast<const void *>(kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)bytes));
  }
  return CRAVA_HIP_OK;
}

template <typename T, typename SI, typename SO, bool DIO>
static int launch_x_v(const Fft3D &f, const GridSet &gs, hipStream_t s, bool inverse) {
  const long long nrows = (long long)f.nyp * f.nzp;
  const int rnxp = 2 * (f.nxp / 2 + 1);
  const PassPlan &p = f.px;
  const unsigned blocks = (unsigned)((nrows + 2LL * p.tx - 1) / (2LL * p.tx));
  const cx<T> *tw = reinterpret_cast<const cx<T> *>(f.twx);
  int rc;
  if (!inverse) {
    if ((rc = set_lds(k_fft_rows_r2c<T, SI, SO, DIO>, p.lds_bytes))) return rc;
    hipLaunchKernelGGL((k_fft_rows_r2c<T, SI, SO, DIO>), dim3(blocks, 1, (unsigned)gs.n), dim3(p.threads), p.lds_bytes,
                       s, gs, nrows, rnxp, p, tw);
  } else {
    if ((rc = set_lds(k_fft_rows_c2r<T, SI, SO, DIO>, p.lds_bytes))) return rc;
    hipLaunchKernelGGL((k_fft_rows_c2r<T, SI, SO, DIO>), dim3(blocks, 1, (unsigned)gs.n), dim3(p.threads), p.lds_bytes,
                       s, gs, nrows, rnxp, p, tw);
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

template <typename T, typename SI, typename SO, bool DIO>
static int launch_strided_g(const PassPlan &p, const void *tw, const GridSet &gs, hipStream_t s, bool inverse,
                            long long es, long long ncols, long long nbatch, long long bstride) {
  const unsigned gx = (unsigned)((ncols + p.tx - 1) / p.tx);
  const dim3 grid(gx, (unsigned)nbatch, (unsigned)gs.n);
  const cx<T> *twp = reinterpret_cast<const cx<T> *>(tw);
  int rc;
  if (!inverse) {
    if ((rc = set_lds(k_fft_strided<-1, T, SI, SO, DIO>, p.lds_bytes))) return rc;
    hipLaunchKernelGGL((k_fft_strided<-1, T, SI, SO, DIO>), grid, dim3(p.threads), p.lds_bytes, s, gs, es, ncols,
                       bstride, p, twp);
  } else {
    if ((rc = set_lds(k_fft_strided<1, T, SI, SO, DIO>, p.lds_bytes))) return rc;
    hipLaunchKernelGGL((k_fft_strided<1, T, SI, SO, DIO>), grid, dim3(p.threads), p.lds_bytes, s, gs, es, ncols,
                       bstride, p, twp);
  }
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

template <typename T, typename SI, typename SO>
static int launch_x(const Fft3D &f, const GridSet &gs, hipStream_t s, bool inverse) {
  return f.gdio[0] ? launch_x_v<T, SI, SO, true>(f, gs, s, inverse) : launch_x_v<T, SI, SO, false>(f, gs, s, inverse);
}

// Generic-kernel dispatch on the storage of one pass.
static int generic_x(const Fft3D &f, int io, const GridSet &gs, hipStream_t s, bool inv) {
  switch (io) {
    case kIoF32: return launch_x<float, float, float>(f, gs, s, inv);
    case kIoF64: return launch_x<double, double, double>(f, gs, s, inv);
    default: return launch_x<double, float, float>(f, gs, s, inv);
  }
}

template <typename T, typename SI, typename SO>
static int launch_strided(const PassPlan &p, const void *tw, const GridSet &gs, hipStream_t s, bool inverse,
                          long long es, long long ncols, long long nbatch, long long bstride, bool dio) {
  return dio ? launch_strided_g<T, SI, SO, true>(p, tw, gs, s, inverse, es, ncols, nbatch, bstride)
             : launch_strided_g<T, SI, SO, false>(p, tw, gs, s, inverse, es, ncols, nbatch, bstride);
}

static int generic_strided(const PassPlan &p, const void *tw, int io, const GridSet &gs, hipStream_t s, bool inv,
                           long long es, long long ncols, long long nbatch, long long bstride, bool dio) {
  switch (io) {
    case kIoF32: return launch_strided<float, float, float>(p, tw, gs, s, inv, es, ncols, nbatch, bstride, dio);
    case kIoF64: return launch_strided<double, double, double>(p, tw, gs, s, inv, es, ncols, nbatch, bstride, dio);
    default: return launch_strided<double, float, float>(p, tw, gs, s, inv, es, ncols, nbatch, bstride, dio);
  }
}

// Every pass runs in place.  A plan of precision 1 (float storage, double
// arithmetic) reads float, transforms in double and rounds each output once
// to float, so the grids stay float between passes and no scratch is needed.
int Fft3D::make_set(int axis, bool inverse, void *const *grids, int ng, const double *scale, GridSet &gs,
                    int &io) const {
  (void)axis;
  (void)inverse;
  if (ng < 1 || ng > kMaxGrids) {
    set_error("a transform set holds 1.." + std::to_string(kMaxGrids) + " grids");
    return CRAVA_HIP_E_INVALID;
  }
  io = prec == 0 ? kIoF32 : prec == 2 ? kIoF64 : kIoF32MathF64;
  gs.n = ng;
  gs.dc = nullptr;
  gs.dcmask = 0;
  for (int i = 0; i < ng; ++i) {
    if (!grids[i]) {
      set_error("NULL grid in a transform set");
      return CRAVA_HIP_E_INVALID;
    }
    gs.src[i] = grids[i];
    gs.dst[i] = grids[i];
    gs.scale[i] = scale ? scale[i] : 1.0;
  }
  for (int i = ng; i < kMaxGrids; ++i) {
    gs.src[i] = nullptr;
    gs.dst[i] = nullptr;
    gs.scale[i] = 0.0;
  }
  return CRAVA_HIP_OK;
}

int Fft3D::pass(int axis, void *const *grids, int ng, bool inverse, const double *scale, hipStream_t s,
                 const double *dc, unsigned dcmask) {
  GridSet gs;
  int io, rc;
  if ((rc = make_set(axis, inverse, grids, ng, scale, gs, io))) return rc;
  if (axis == 0) {
    gs.dc = dc;
    gs.dcmask = dc ? dcmask : 0u;
  }
  const long long cnxp = nxp / 2 + 1;
  if (axis == 0) {
    if (var[0] >= 0)
      return static_fft_launch(nxp, io, inverse ? kRowsC2R : kRowsR2C, var[0], gs, 0, 0, 0, 0, (long long)nyp * nzp,
                               (int)(2 * cnxp), stw[0], s);
    return generic_x(*this, io, gs, s, inverse);
  }
  const FftKind kind = inverse ? kStridedInv : kStridedFwd;
  if (axis == 1) {
    if (nyp <= 1) return CRAVA_HIP_OK;  // a length-1 axis is the identity
    if (var[1] >= 0) return static_fft_launch(nyp, io, kind, var[1], gs, cnxp, cnxp, nzp, cnxp * nyp, 0, 0, stw[1], s);
    return generic_strided(py, twy, io, gs, s, inverse, cnxp, cnxp, nzp, cnxp * nyp, gdio[1]);
  }
  const long long plane = cnxp * nyp;
  if (nzp <= 1) return CRAVA_HIP_OK;
  if (var[2] >= 0) return static_fft_launch(nzp, io, kind, var[2], gs, plane, plane, 1, 0, 0, 0, stw[2], s);
  return generic_strided(pz, twz, io, gs, s, inverse, plane, plane, 1, 0, gdio[2]);
}

int Fft3D::set_variant(int axis, int v) {
  if (axis < 0 || axis > 2) return CRAVA_HIP_E_INVALID;
  const int n = axis == 0 ? nxp : (axis == 1 ? nyp : nzp);
  PassPlan &p = axis == 0 ? px : (axis == 1 ? py : pz);
  if (v < 0) {
    int g = -1 - v;
    // -7..-12 are -1..-6 with the first radix pass loading its butterflies
    // straight from global memory
    const bool dio = g >= kNumGenericVariants;
    if (dio) g -= kNumGenericVariants;
    if (g >= kNumGenericVariants) {
      set_error("generic variant " + std::to_string(v) + " does not exist");
      return CRAVA_HIP_E_INVALID;
    }
    PassPlan q;
    int rc;
    if ((rc = build_plan(n, q, g))) return rc;
    p = q;
    var[axis] = v;
    gdio[axis] = dio;
    return CRAVA_HIP_OK;
  }
  const FftKind kind = axis == 0 ? kRowsR2C : kStridedFwd;
  if (!static_fft_geometry(n, dbl, kind, v, nullptr)) {
    set_error("variant " + std::to_string(v) + " unavailable for length " + std::to_string(n));
    return CRAVA_HIP_E_INVALID;
  }
  var[axis] = v;
  return CRAVA_HIP_OK;
}

int Fft3D::num_variants(int axis) const {
  const int n = axis == 0 ? nxp : (axis == 1 ? nyp : nzp);
  // odd x lengths: the direct-I/O row variants only (static_fft_geometry)
  if (!static_fft_available(n)) return 0;
  return axis == 0 ? kNumRowVariants : kNumStridedVariants;
}

// Time every launch variant of every axis (generic radix orders and pencil
// counts, and the specialised kernels where the length has them) on the set
// `grids` (contents are destroyed) and keep the fastest.  Forward and inverse
// share a variant.
int Fft3D::autotune(void *const *grids, int ng, hipStream_t s, int reps, float *ms_out) {
  int rc;
  hipEvent_t e0, e1;
  CRAVA_HIP_CHECK(hipEventCreate(&e0));
  CRAVA_HIP_CHECK(hipEventCreate(&e1));
  rc = CRAVA_HIP_OK;
  for (int axis = 0; axis < 3 && rc == CRAVA_HIP_OK; ++axis) {
    const int nv = num_variants(axis);
    float best = 1e30f;
    int bestv = var[axis];
    for (int v = -2 * kNumGenericVariants; v < nv; ++v) {
      if (set_variant(axis, v) != CRAVA_HIP_OK) continue;
      // forward and inverse pass together: one variant serves both directions
      auto run = [&]() {
        int r = pass(axis, grids, ng, false, nullptr, s);
        return r ? r : pass(axis, grids, ng, true, nullptr, s);
      };
      if ((rc = run())) break;  // warm-up
      if (hipEventRecord(e0, s) != hipSuccess) { rc = CRAVA_HIP_E_HIP; break; }
      for (int r = 0; r < reps && rc == CRAVA_HIP_OK; ++r) rc = run();
      if (rc) break;
      if (hipEventRecord(e1, s) != hipSuccess || hipEventSynchronize(e1) != hipSuccess) { rc = CRAVA_HIP_E_HIP; break; }
      float ms = 0.0f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      ms /= (float)reps;
      if (ms < best) { best = ms; bestv = v; }
    }
    if (rc == CRAVA_HIP_OK) {
      rc = set_variant(axis, bestv);
      if (ms_out) ms_out[axis] = best;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  set_error("");
  return rc;
}

// Mean deflation around the passes (one thread per grid of the set; the
// set's scale holds the exact level scales):
// 0 before a forward transform: dc[g] = the grid's first real value;
// 1 after it: DC bin += c * N * scale[g] in double, c = dc[g] as the x
//   pass subtracted it (rounded to the storage type);
// 2 before an inverse transform: dc[g] = DC bin * scale[g], DC bin = 0.
template <typename S>
__global__ void k_fft_dc(GridSet gs, int which, double n, double *dc, unsigned dcmask) {
  const int g = threadIdx.x;
  if (g >= gs.n || !((dcmask >> g) & 1u)) return;
  S *p = static_cast<S *>(gs.dst[g]);
  if (which == 0) {
    dc[g] = (double)p[0];
  } else if (which == 1) {
    p[0] = (S)((double)p[0] + (double)(S)dc[g] * n * gs.scale[g]);
  } else {
    dc[g] = (double)p[0] * gs.scale[g];
    p[0] = (S)0;
    p[1] = (S)0;
  }
}

int Fft3D::dc_kernel(int which, void *const *grids, int ng, const double *dscale, double *dc, unsigned dcmask,
                     hipStream_t s) const {
  GridSet gs;
  int io, rc;
  if ((rc = make_set(0, false, grids, ng, dscale, gs, io))) return rc;
  const double n = (double)nxp * nyp * nzp;
  if (prec == 2)
    hipLaunchKernelGGL(k_fft_dc<double>, dim3(1), dim3(kMaxGrids), 0, s, gs, which, n, dc, dcmask);
  else
    hipLaunchKernelGGL(k_fft_dc<float>, dim3(1), dim3(kMaxGrids), 0, s, gs, which, n, dc, dcmask);
  CRAVA_HIP_CHECK(hipGetLastError());
  return CRAVA_HIP_OK;
}

int Fft3D::r2c(void *const *grids, int ng, const double *scale, hipStream_t s, double *dc, unsigned dcmask,
               const double *dscale) {
  int rc;
  if (!dc) dcmask = 0;
  if (!dscale) dscale = scale;
  if (dcmask && (rc = dc_kernel(0, grids, ng, dscale, dc, dcmask, s))) return rc;
  if ((rc = pass(0, grids, ng, false, scale, s, dc, dcmask))) return rc;
  if ((rc = pass(1, grids, ng, false, nullptr, s))) return rc;
  if ((rc = pass(2, grids, ng, false, nullptr, s))) return rc;
  return dcmask ? dc_kernel(1, grids, ng, dscale, dc, dcmask, s) : CRAVA_HIP_OK;
}

int Fft3D::c2r(void *const *grids, int ng, const double *scale, hipStream_t s, double *dc, unsigned dcmask,
               const double *dscale) {
  int rc;
  if (!dc) dcmask = 0;
  if (!dscale) dscale = scale;
  if (dcmask && (rc = dc_kernel(2, grids, ng, dscale, dc, dcmask, s))) return rc;
  if ((rc = pass(2, grids, ng, true, nullptr, s))) return rc;
  if ((rc = pass(1, grids, ng, true, nullptr, s))) return rc;
  return pass(0, grids, ng, true, scale, s, dc, dcmask);
}

}  // namespace crava
