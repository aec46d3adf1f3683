// Compile-time specialised Stockham kernels for the transform lengths of the
// standard workloads (see fft_static.hip).  Each length has a few launch
// variants (pencils per workgroup x threads); the autotuner picks one.
#pragma once
#include "common.hpp"

namespace crava {

// Kinds of pass.
enum FftKind { kStridedFwd = 0, kStridedInv = 1, kRowsR2C = 2, kRowsC2R = 3 };

constexpr int kNumStridedVariants = 12;
constexpr int kNumRowVariants = 6;

struct StaticLaunch {
  int tx = 0;        // pencils (strided) or row pairs (rows) per workgroup
  int threads = 0;
  size_t lds = 0;
};

// True if length n has specialised kernels (some launch variants).
bool static_fft_available(int n);
// True if length n has specialised kernels in every launch variant (the
// main lengths; the fused forward-modelling and wavelet-division kernels
// exist for these).
bool static_fft_full(int n);
// True if launch variant `variant` of `kind` exists for length n.
bool static_variant_available(int n, FftKind kind, int variant);
// The per-pass twiddle table the specialised kernels of length n read
// (2n doubles, {re, im} pairs); false if n is not specialised.
bool static_twiddle_table(int n, double *re_im);
// Launch geometry of a variant; false if the variant does not fit.
bool static_fft_geometry(int n, bool dbl_compute, FftKind kind, int variant, StaticLaunch *out);
// Up to kMaxGrids same-size grids transformed by one launch (blockIdx.z).
constexpr int kMaxGrids = 16;
struct GridSet {
  const void *src[kMaxGrids];
  void *dst[kMaxGrids];
  double scale[kMaxGrids];
  int n;
  // Mean deflation of the x pass (grids whose bit is set in dcmask): the
  // forward pass transforms x - c (c = dc[g], a value near the grid's
  // level, exact by Sterbenz's lemma for values within a factor 2 of it),
  // the inverse pass adds dc[g] to every output; the DC bin carries c
  // (see Fft3D::r2c / c2r).  dc is device memory or NULL.
  const double *dc;
  unsigned dcmask;
};

// Storage and arithmetic of a pass: float, double, or float storage with
// double arithmetic inside the pass (each pass rounds its outputs once).
enum FftIo { kIoF32 = 0, kIoF64 = 1, kIoF32MathF64 = 2 };

// Launch a specialised pass over every grid of `gs` (src -> dst, may alias
// for in place).  Strided: complex, `es` element stride, `ncols` contiguous
// columns, `nbatch` batches `bstride` apart.  Rows: `nrows` rows of pitch
// `rnxp` reals, gs.scale applied to the output.
int static_fft_launch(int n, int io, FftKind kind, int variant, const GridSet &gs, long long es, long long ncols,
                      long long nbatch, long long bstride, long long nrows, int rnxp, const void *twiddles,
                      hipStream_t s);

}  // namespace crava
