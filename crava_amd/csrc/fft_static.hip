// Specialised transform kernels of the main lengths (every launch variant),
// the availability and geometry of every specialised length, and the
// dispatch to the extended lengths' translation units.
#include "fft_static_impl.hpp"

namespace crava {

// Lengths with specialised kernels in every launch variant: the x, y and z
// lengths of the standard workloads (80/160 plain CPU grid, 288 plain GPU
// grid, 512/320 large grid, 216 differentiation grid) plus common powers of
// two.  The other 7-smooth lengths from 32 to 320 (the padding rule's
// outputs) have the direct-I/O shapes (fft_static_impl.hpp).
#define CRAVA_STATIC_LENGTHS(X) X(64) X(80) X(128) X(160) X(216) X(256) X(288) X(320) X(512)

bool static_fft_full(int n) {
#define CRAVA_CASE(L) if (n == L) return true;
  CRAVA_STATIC_LENGTHS(CRAVA_CASE)
#undef CRAVA_CASE
  return false;
}

// Group of an extended length (0..5), or -1.
static int ext_group(int n) {
#define CRAVA_CASE(L) if (n == L) return g;
  int g = 0;
  CRAVA_EXT_LENGTHS_0(CRAVA_CASE)
  g = 1;
  CRAVA_EXT_LENGTHS_1(CRAVA_CASE)
  g = 2;
  CRAVA_EXT_LENGTHS_2(CRAVA_CASE)
  g = 3;
  CRAVA_EXT_LENGTHS_3(CRAVA_CASE)
  g = 4;
  CRAVA_EXT_LENGTHS_4(CRAVA_CASE)
  g = 5;
  CRAVA_EXT_LENGTHS_5(CRAVA_CASE)
#undef CRAVA_CASE
  return -1;
}

bool static_fft_available(int n) { return static_fft_full(n) || ext_group(n) >= 0; }

bool static_variant_available(int n, FftKind kind, int variant) {
  const bool strided = kind == kStridedFwd || kind == kStridedInv;
  if (strided && (variant < 0 || variant >= kNumStridedVariants)) return false;
  if (!strided && (variant < 0 || variant >= kNumRowVariants)) return false;
  if (static_fft_full(n)) return true;
  if (ext_group(n) < 0) return false;
  return strided ? (variant == 4 || variant == 5 || variant >= 10) : (variant == 3 || variant == 5);
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
  if (!static_variant_available(n, kind, variant)) return false;
  const bool strided = kind == kStridedFwd || kind == kStridedInv;
  const int tx = strided ? kSV[variant][0] : kRV[variant][0];
  const int th = strided ? kSV[variant][1] : kRV[variant][1];
  const size_t eb = dbl ? 16 : 8;
  const size_t len = (size_t)n + (size_t)((n - 1) >> (dbl ? 4 : 5)) + 1;
  // padded pencils, or N | 1 apart for the pencil-fastest form
  const size_t pitch = (strided && kSV[variant][2] == 3) ? (size_t)(n | 1) : (len | 1);
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
  switch (ext_group(n)) {
    case 0: return static_ext_launch_0(n, io, kind, variant, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
    case 1: return static_ext_launch_1(n, io, kind, variant, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
    case 2: return static_ext_launch_2(n, io, kind, variant, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
    case 3: return static_ext_launch_3(n, io, kind, variant, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
    case 4: return static_ext_launch_4(n, io, kind, variant, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
    case 5: return static_ext_launch_5(n, io, kind, variant, gs, es, ncols, nbatch, bstride, nrows, rnxp, tw, s);
    default: return CRAVA_HIP_E_INVALID;
  }
}

}  // namespace crava
