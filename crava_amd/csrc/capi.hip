// C ABI of libcrava_hip.so (declared in include/crava_hip.h).
#include "../../include/crava_hip.h"
#include "batch.hpp"
#include "common.hpp"
#include "context.hpp"
#include "fft.hpp"
#include "facies.hpp"
#include "forward.hpp"
#include "grid.hpp"
#include "nll.hpp"
#include "posterior.hpp"
#include "simulate.hpp"
#include "swdiv.hpp"
#include "kde.hpp"
#include "localnoise.hpp"
#include "tables.hpp"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <tuple>
#include <vector>

namespace crava {
static thread_local std::string g_err;
void set_error(const std::string &msg) { g_err = msg; }
const char *get_error() { return g_err.c_str(); }

int Context::plan(int nxp, int nyp, int nzp, int prec, Fft3D **out) {
  auto key = std::make_tuple(nxp, nyp, nzp, prec);
  auto it = plans.find(key);
  if (it == plans.end()) {
    std::unique_ptr<Fft3D> f(new (std::nothrow) Fft3D(nxp, nyp, nzp, prec));
    if (!f) return CRAVA_HIP_E_OOM;
    int rc = f->init();
    if (rc) return rc;
    it = plans.emplace(key, std::move(f)).first;
  }
  *out = it->second.get();
  return CRAVA_HIP_OK;
}

int Context::workspace(size_t bytes, void **out) {
  if (bytes > work_bytes) {
    if (work) (void)hipFree(work);
    work = nullptr;
    work_bytes = 0;
    CRAVA_HIP_CHECK(hipMalloc(&work, bytes));
    work_bytes = bytes;
  }
  *out = work;
  return CRAVA_HIP_OK;
}

Context::~Context() {
  plans.clear();
  if (counters) (void)hipFree(counters);
  if (dcbuf) (void)hipFree(dcbuf);
  if (work) (void)hipFree(work);
}
}  // namespace crava

struct crava_hip_post {
  crava_hip_ctx *ctx;
  crava::PostDevice dev;
  // host copies of the per-k vectors (the likelihood rebuilds its tables)
  std::vector<unsigned char> invert;
  std::vector<float> kWD, e1, e2, e3;
  // likelihood workspace, allocated on first use
  crava::NllTab *nll_tab = nullptr;
  int *nll_klist = nullptr;
  double *nll_partial = nullptr;
  int nll_blocks = 0;
  // queue of the mixed form (count, then the list), allocated on the first
  // CRAVA_HIP_MIXED run
  crava::DeferQueue queue;
  void *queue_mem = nullptr;
};

using crava::set_error;

// ---------------------------------------------------------------------------
// Static queue of crava_hip_post_run_static: the mixed form's FP64 queue (a
// count, then up to kStaticQueueEntries {k, first point} entries) in a
// file-scope device array, so that the call allocates nothing.  Every GPU
// holds its own copy; its address is looked up once per device and kept.
namespace {
constexpr long long kStaticQueueEntries = 1LL << 20;   // 8 MiB of entries: 67 M in-band points
constexpr size_t kStaticQueueHead = 256;
__device__ __align__(256) unsigned char scratch[kStaticQueueHead + sizeof(int2) * kStaticQueueEntries];
constexpr int kMaxDevices = 64;
void *g_scratch_dev[kMaxDevices] = {nullptr};
std::mutex g_scratch_mu;

// The current device's copy of the static queue.
int static_queue(crava::DeferQueue *q) {
  int dev = -1;
  CRAVA_HIP_CHECK(hipGetDevice(&dev));
  if (dev < 0 || dev >= kMaxDevices) {
    set_error("device index out of range for the static queue");
    return CRAVA_HIP_E_INVALID;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_dev[dev]) {
    void *ptr = nullptr;
    CRAVA_HIP_CHECK(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    g_scratch_dev[dev] = ptr;
  }
  unsigned char *base = static_cast<unsigned char *>(g_scratch_dev[dev]);
  q->count = reinterpret_cast<unsigned *>(base);
  q->list = reinterpret_cast<int2 *>(base + kStaticQueueHead);
  q->cap = kStaticQueueEntries;
  return CRAVA_HIP_OK;
}

__global__ void k_zero_counts2(unsigned long long *c) {
  if (threadIdx.x < 2) c[threadIdx.x] = 0ULL;
}
}  // namespace

// transform arithmetic/storage code of a C-ABI precision
static int fft_prec(int precision) {
  return precision == CRAVA_HIP_F64 ? 2 : (precision == CRAVA_HIP_MIXED ? 1 : 0);
}

// exact (double) scale of a transform, the level scale of a deflated one
static double fft_scale_exact(int cubetype, long long N, bool inverse) {
  return crava::grid_fft_scale(cubetype == CRAVA_HIP_COVARIANCE, true, N, inverse);
}

extern "C" {

int crava_hip_version(void) { return 100; }

const char *crava_hip_last_error(void) { return crava::get_error(); }

int crava_hip_ctx_create(crava_hip_ctx **ctx, void *stream) {
  if (!ctx) { set_error("ctx is NULL"); return CRAVA_HIP_E_INVALID; }
  *ctx = nullptr;
  crava_hip_ctx *c = new (std::nothrow) crava_hip_ctx();
  if (!c) { set_error("out of host memory"); return CRAVA_HIP_E_OOM; }
  c->stream = reinterpret_cast<hipStream_t>(stream);
  hipError_t e = hipMalloc(&c->counters, 4 * sizeof(unsigned long long));
  if (e == hipSuccess) e = hipMalloc(&c->dcbuf, crava::kMaxGrids * sizeof(double));
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc(context buffers): ") + hipGetErrorString(e));
    delete c;
    return CRAVA_HIP_E_HIP;
  }
  *ctx = c;
  return CRAVA_HIP_OK;
}

int crava_hip_ctx_destroy(crava_hip_ctx *ctx) {
  if (!ctx) return CRAVA_HIP_OK;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream); else (void)hipDeviceSynchronize();
  delete ctx;
  return CRAVA_HIP_OK;
}

int crava_hip_ctx_set_stream(crava_hip_ctx *ctx, void *stream) {
  if (!ctx) { set_error("ctx is NULL"); return CRAVA_HIP_E_INVALID; }
  ctx->stream = reinterpret_cast<hipStream_t>(stream);
  return CRAVA_HIP_OK;
}

int crava_hip_synchronize(crava_hip_ctx *ctx) {
  if (!ctx) { set_error("ctx is NULL"); return CRAVA_HIP_E_INVALID; }
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return CRAVA_HIP_OK;
}

// Float transforms (deflated or not) round the factor to float as FFTGrid
// does; the mixed and double transforms keep it in double.
static double fft_scale(int cubetype, int precision, long long N, bool inverse) {
  const bool f64 = precision == CRAVA_HIP_MIXED || precision == CRAVA_HIP_F64;
  return crava::grid_fft_scale(cubetype == CRAVA_HIP_COVARIANCE, f64, N, inverse);
}

static int fft_common(crava_hip_ctx *ctx, void *const *grids, int ngrid, int nxp, int nyp, int nzp, int cubetype,
                      int precision, bool inverse) {
  if (!ctx || !grids) { set_error("NULL ctx or grid list"); return CRAVA_HIP_E_INVALID; }
  if (ngrid < 1 || ngrid > crava::kMaxGrids) {
    set_error("ngrid must be 1.." + std::to_string(crava::kMaxGrids));
    return CRAVA_HIP_E_INVALID;
  }
  if (nxp < 2 || nyp < 1 || nzp < 1) { set_error("invalid grid size"); return CRAVA_HIP_E_INVALID; }
  if (cubetype < CRAVA_HIP_DATA || cubetype > CRAVA_HIP_COVARIANCE) { set_error("invalid cubetype"); return CRAVA_HIP_E_INVALID; }
  if (precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F32_DEFLATE) {
    set_error("invalid transform precision");
    return CRAVA_HIP_E_INVALID;
  }
  for (int i = 0; i < ngrid; ++i)
    if (!grids[i]) { set_error("NULL grid"); return CRAVA_HIP_E_INVALID; }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(nxp, nyp, nzp, fft_prec(precision), &f);
  if (rc) return rc;
  double scale[crava::kMaxGrids], exact[crava::kMaxGrids];
  const long long N = (long long)nxp * nyp * nzp;
  const double sc = fft_scale(cubetype, precision, N, inverse), se = fft_scale_exact(cubetype, N, inverse);
  for (int i = 0; i < ngrid; ++i) {
    scale[i] = sc;
    exact[i] = se;
  }
  const unsigned mask = precision == CRAVA_HIP_F32_DEFLATE ? (ngrid >= 32 ? ~0u : (1u << ngrid) - 1u) : 0u;
  return inverse ? f->c2r(grids, ngrid, scale, ctx->stream, ctx->dcbuf, mask, exact)
                 : f->r2c(grids, ngrid, scale, ctx->stream, ctx->dcbuf, mask, exact);
}

int crava_hip_fft_r2c(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, int cubetype, int precision) {
  if (!grid) { set_error("NULL grid"); return CRAVA_HIP_E_INVALID; }
  return fft_common(ctx, &grid, 1, nxp, nyp, nzp, cubetype, precision, false);
}

int crava_hip_fft_c2r(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, int cubetype, int precision) {
  if (!grid) { set_error("NULL grid"); return CRAVA_HIP_E_INVALID; }
  return fft_common(ctx, &grid, 1, nxp, nyp, nzp, cubetype, precision, true);
}

int crava_hip_fft_r2c_many(crava_hip_ctx *ctx, void *const *grids, int ngrid, int nxp, int nyp, int nzp,
                           int cubetype, int precision) {
  return fft_common(ctx, grids, ngrid, nxp, nyp, nzp, cubetype, precision, false);
}

int crava_hip_fft_c2r_many(crava_hip_ctx *ctx, void *const *grids, int ngrid, int nxp, int nyp, int nzp,
                           int cubetype, int precision) {
  return fft_common(ctx, grids, ngrid, nxp, nyp, nzp, cubetype, precision, true);
}

int crava_hip_fft_pass_many(crava_hip_ctx *ctx, void *const *grids, int ngrid, int nxp, int nyp, int nzp, int axis,
                            int inverse, int precision, double scale) {
  if (!ctx || !grids || axis < 0 || axis > 2 || ngrid < 1 || ngrid > crava::kMaxGrids) {
    set_error("invalid arguments to crava_hip_fft_pass");
    return CRAVA_HIP_E_INVALID;
  }
  if (precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F64) { set_error("invalid precision"); return CRAVA_HIP_E_INVALID; }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(nxp, nyp, nzp, fft_prec(precision), &f);
  if (rc) return rc;
  double sc[crava::kMaxGrids];
  for (int i = 0; i < ngrid; ++i) sc[i] = scale;
  return f->pass(axis, grids, ngrid, inverse != 0, sc, ctx->stream);
}

int crava_hip_fft_pass(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, int axis, int inverse,
                       int precision, double scale) {
  if (!grid) { set_error("NULL grid"); return CRAVA_HIP_E_INVALID; }
  return crava_hip_fft_pass_many(ctx, &grid, 1, nxp, nyp, nzp, axis, inverse, precision, scale);
}

int crava_hip_autotune_fft(crava_hip_ctx *ctx, void *scratch, int nxp, int nyp, int nzp, int precision,
                           int *variants, float *ms) {
  if (!ctx || !scratch) { set_error("NULL argument to crava_hip_autotune_fft"); return CRAVA_HIP_E_INVALID; }
  if (precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F32_DEFLATE) {
    set_error("invalid precision");
    return CRAVA_HIP_E_INVALID;
  }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(nxp, nyp, nzp, fft_prec(precision), &f);
  if (rc) return rc;
  float t[3] = {0, 0, 0};
  if ((rc = f->autotune(&scratch, 1, ctx->stream, 5, t))) return rc;
  for (int a = 0; a < 3; ++a) {
    if (variants) variants[a] = f->var[a];
    if (ms) ms[a] = t[a];
  }
  return CRAVA_HIP_OK;
}

int crava_hip_fft_set_variants(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, int precision, const int *variants) {
  if (!ctx || !variants) { set_error("NULL argument to crava_hip_fft_set_variants"); return CRAVA_HIP_E_INVALID; }
  if (precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F32_DEFLATE) {
    set_error("invalid precision");
    return CRAVA_HIP_E_INVALID;
  }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(nxp, nyp, nzp, fft_prec(precision), &f);
  if (rc) return rc;
  for (int a = 0; a < 3; ++a)
    if ((rc = f->set_variant(a, variants[a]))) return rc;
  return CRAVA_HIP_OK;
}

int crava_hip_perk_tables(int ntheta, int nz, int nzp, double lz, float low_cut, float high_cut,
                          float min_rel_thick, const float *wavelets, const float *wnorm, unsigned char *invert,
                          float *kWD, float *e1, float *e2, float *e3) {
  if (ntheta < 1 || ntheta > crava::kMaxTheta || nzp < 2 || nz < 1 || !wavelets || !wnorm || !invert || !kWD ||
      !e1 || !e2 || !e3) {
    set_error("invalid arguments to crava_hip_perk_tables");
    return CRAVA_HIP_E_INVALID;
  }
  crava::perk_vectors(ntheta, nz, nzp, lz, low_cut, high_cut, min_rel_thick, wavelets, wnorm, invert, kWD, e1, e2, e3);
  return CRAVA_HIP_OK;
}

int crava_hip_perk_tables_nonconst(int ntheta, int nz, int nzp, double lz, float low_cut, float high_cut,
                                   float min_rel_thick, const float *wavelets, const float *wnorm,
                                   unsigned char *invert, float *kWD, float *e1, float *e2, float *e3) {
  if (ntheta < 1 || ntheta > crava::kMaxTheta || nzp < 2 || nz < 1 || !wavelets || !wnorm || !invert || !kWD ||
      !e1 || !e2 || !e3) {
    set_error("invalid arguments to crava_hip_perk_tables_nonconst");
    return CRAVA_HIP_E_INVALID;
  }
  crava::perk_vectors_nonconst(ntheta, nz, nzp, lz, low_cut, high_cut, min_rel_thick, wavelets, wnorm, invert, kWD,
                               e1, e2, e3);
  return CRAVA_HIP_OK;
}

int crava_hip_wavelet_resample(const float *wavelet, int nw, int center, float dz_wavelet, float dz, int nzp,
                               float *out, float *norm) {
  if (!wavelet || !out || nw < 1 || nzp < 2 || !(dz_wavelet > 0.0f) || !(dz > 0.0f)) {
    set_error("invalid arguments to crava_hip_wavelet_resample");
    return CRAVA_HIP_E_INVALID;
  }
  float nv = 0.0f;
  if (!crava::wavelet_resample(wavelet, nw, center, dz_wavelet, dz, nzp, out, &nv)) {
    set_error("crava_hip_wavelet_resample: kriging matrix not positive definite (sample spacing too fine for "
              "the 20 Hz Ricker correlation)");
    return CRAVA_HIP_E_INVALID;
  }
  if (norm) *norm = nv;
  return CRAVA_HIP_OK;
}

int crava_hip_wavelet_energy(int ntheta, int nzp, const float *wavelets, const float *A, const double *prior_var0,
                             const float *corr_t, const float *error_variance, const float *emp_sn,
                             const unsigned char *match, float *gain, float *model_var, float *signal_var,
                             float *theo_sn) {
  if (ntheta < 1 || ntheta > crava::kMaxTheta || nzp < 2 || !wavelets || !A || !prior_var0 || !corr_t ||
      !error_variance || !emp_sn || !gain || !model_var || !signal_var || !theo_sn) {
    set_error("invalid arguments to crava_hip_wavelet_energy");
    return CRAVA_HIP_E_INVALID;
  }
  crava::wavelet_energy(ntheta, nzp, wavelets, A, prior_var0, corr_t, error_variance, emp_sn, match, gain, model_var,
                        signal_var, theo_sn);
  return CRAVA_HIP_OK;
}

int crava_hip_padding_size(int n, double pad_fac, int *npad) {
  if (n < 1 || !npad || !(pad_fac >= 0.0) || !(n * (1.0 + pad_fac) < 1.0e8)) {
    set_error("invalid arguments to crava_hip_padding_size");
    return CRAVA_HIP_E_INVALID;
  }
  *npad = crava::padding_size(n, pad_fac);
  return CRAVA_HIP_OK;
}

int crava_hip_estimate_xy_padding(int nx, int ny, double dx, double dy, double lx, double ly, double x_pad_fac,
                                  double y_pad_fac, int estimate, float range1, float range2, float angle,
                                  int *nxpad, int *nypad, double *true_x_pad_fac, double *true_y_pad_fac) {
  if (nx < 1 || ny < 1 || !(dx > 0.0) || !(dy > 0.0) || !(lx > 0.0) || !(ly > 0.0) || !(x_pad_fac >= 0.0) ||
      !(y_pad_fac >= 0.0) || !(x_pad_fac <= 1.0e4) || !(y_pad_fac <= 1.0e4) || !std::isfinite(range1) ||
      !std::isfinite(range2) || !std::isfinite(angle) || !nxpad || !nypad || !true_x_pad_fac || !true_y_pad_fac) {
    set_error("invalid arguments to crava_hip_estimate_xy_padding");
    return CRAVA_HIP_E_INVALID;
  }
  crava::estimate_xy_padding(nx, ny, dx, dy, lx, ly, x_pad_fac, y_pad_fac, estimate != 0, range1, range2, angle, nxpad,
                             nypad, true_x_pad_fac, true_y_pad_fac);
  return CRAVA_HIP_OK;
}

int crava_hip_estimate_z_padding(int nz, double lz, double min_rel_thick, double z_pad_fac, int estimate,
                                 float default_wavelet_length, int *nzpad, double *true_z_pad_fac) {
  if (nz < 1 || !(lz > 0.0) || !(min_rel_thick > 0.0) || !(z_pad_fac >= 0.0) || !(z_pad_fac <= 1.0e4) ||
      !(default_wavelet_length >= 0.0f) || !nzpad || !true_z_pad_fac) {
    set_error("invalid arguments to crava_hip_estimate_z_padding");
    return CRAVA_HIP_E_INVALID;
  }
  crava::estimate_z_padding(nz, lz, min_rel_thick, z_pad_fac, estimate != 0, default_wavelet_length, nzpad,
                            true_z_pad_fac);
  return CRAVA_HIP_OK;
}

int crava_hip_circ_corr_t(const float *prior_corr_t, int n, int min_int_fq, int nzp, float *out) {
  if (!prior_corr_t || !out || n < 1 || nzp < 1 || nzp > (1 << 16)) {
    set_error("invalid arguments to crava_hip_circ_corr_t");
    return CRAVA_HIP_E_INVALID;
  }
  if (prior_corr_t[0] == 0.0f) {
    set_error("crava_hip_circ_corr_t: priorCorrT[0] must be nonzero");
    return CRAVA_HIP_E_INVALID;
  }
  if (!crava::circ_corr_t(prior_corr_t, n, min_int_fq, nzp, out)) {
    set_error("crava_hip_circ_corr_t: the circular temporal correlation is undefined (c(0) <= 1e-5); increase the "
              "number of layers");
    return CRAVA_HIP_E_INVALID;
  }
  return CRAVA_HIP_OK;
}

int crava_hip_prior_var0_robustify(const double *prior_var0, double *out, int *changed) {
  if (!prior_var0 || !out) {
    set_error("invalid arguments to crava_hip_prior_var0_robustify");
    return CRAVA_HIP_E_INVALID;
  }
  for (int i = 0; i < 9; ++i)
    if (!std::isfinite(prior_var0[i])) {
      set_error("crava_hip_prior_var0_robustify: non-finite prior covariance");
      return CRAVA_HIP_E_INVALID;
    }
  double tmp[9];
  const int rc = crava::prior_var0_robustify(prior_var0, tmp);
  if (rc < 0) {
    set_error("crava_hip_prior_var0_robustify: the prior covariance has an eigenvalue below 0.1 % of the largest "
              "and a zero off-diagonal entry or rows that DGEEV would balance; its LAPACK eigenvectors, which the "
              "lift uses, are not reproduced for such a matrix");
    return CRAVA_HIP_E_INVALID;
  }
  for (int i = 0; i < 9; ++i) out[i] = tmp[i];
  if (changed) *changed = rc;
  return CRAVA_HIP_OK;
}

int crava_hip_complex_noise_reference(int nxp, int nyp, int nzp, uint32_t *seed, float *host_out) {
  if (nxp < 1 || nyp < 1 || nzp < 1 || !seed || !host_out) {
    set_error("invalid arguments to crava_hip_complex_noise_reference");
    return CRAVA_HIP_E_INVALID;
  }
  crava::reference_complex_noise(nxp, nyp, nzp, seed, host_out);
  return CRAVA_HIP_OK;
}

int crava_hip_reflection_matrix(int ntheta, const float *theta, const int *seismic_type, double vsvp, float *A) {
  if (ntheta < 1 || !theta || !A) {
    set_error("invalid arguments to crava_hip_reflection_matrix");
    return CRAVA_HIP_E_INVALID;
  }
  const double vsvp2 = vsvp * vsvp;
  for (int i = 0; i < ntheta; ++i) {
    const int type = seismic_type ? seismic_type[i] : 0;
    const double angle = (double)theta[i];
    const double sint = std::sin(angle), sint2 = sint * sint;
    float *a = A + 3 * i;
    if (type == 0) {            // PP (linearised Zoeppritz, Aki-Richards form)
      const double tan2t = std::tan(angle) * std::tan(angle);
      a[0] = (float)((1.0 + tan2t) / 2.0);
      a[1] = (float)(-4 * vsvp2 * sint2);
      a[2] = (float)((1.0 - 4.0 * vsvp2 * sint2) / 2.0);
    } else if (type == 1) {     // PS
      const double cost = std::cos(angle), cosp = std::sqrt(1 - vsvp2 * sint2);
      const double fac = 0.5 * sint / cosp;
      a[0] = 0;
      a[1] = (float)(4.0 * fac * (vsvp2 * sint2 - vsvp * cost * cosp));
      a[2] = (float)(fac * (-1.0 + 2 * vsvp2 * sint2 + 2 * vsvp * cost * cosp));
    } else {
      set_error("seismic_type must be 0 (PP) or 1 (PS)");
      return CRAVA_HIP_E_INVALID;
    }
  }
  return CRAVA_HIP_OK;
}

int crava_hip_post_create(crava_hip_ctx *ctx, const crava_hip_post_params *params, const unsigned char *invert,
                          const float *kWD, const float *e1, const float *e2, const float *e3,
                          crava_hip_post **post) {
  if (!ctx || !params || !invert || !kWD || !e1 || !e2 || !e3 || !post) {
    set_error("NULL argument to crava_hip_post_create");
    return CRAVA_HIP_E_INVALID;
  }
  *post = nullptr;
  const crava_hip_post_params &p = *params;
  if (p.ntheta < 1 || p.ntheta > crava::kMaxTheta || p.nxp < 2 || p.nyp < 1 || p.nzp < 1) {
    set_error("invalid posterior problem size");
    return CRAVA_HIP_E_INVALID;
  }
  crava_hip_post *h = new (std::nothrow) crava_hip_post();
  if (!h) { set_error("out of host memory"); return CRAVA_HIP_E_OOM; }
  h->ctx = ctx;
  crava::PostProblem &pp = h->dev.pp;
  pp.nxp = p.nxp; pp.nyp = p.nyp; pp.nzp = p.nzp; pp.ntheta = p.ntheta;
  pp.wnc = p.wnc;
  std::memcpy(pp.sigma0, p.sigma0, sizeof(pp.sigma0));
  std::memcpy(pp.A, p.A, sizeof(pp.A));
  std::memcpy(pp.err_theta_cov, p.err_theta_cov, sizeof(pp.err_theta_cov));
  pp.mixed_threshold = p.mixed_threshold;
  std::vector<crava::PerK<float>> tf(p.nzp);
  std::vector<crava::PerK<double>> td(p.nzp);
  crava::build_perk(pp, invert, kWD, e1, e2, e3, tf.data(), td.data());
  h->dev.defer_need = crava::defer_capacity(pp, invert);
  const size_t nv = (size_t)p.nzp * p.ntheta * 2;
  h->invert.assign(invert, invert + p.nzp);
  h->kWD.assign(kWD, kWD + nv);
  h->e1.assign(e1, e1 + nv);
  h->e2.assign(e2, e2 + nv);
  h->e3.assign(e3, e3 + nv);
  hipError_t e = hipMalloc(&h->dev.tabf, sizeof(crava::PerK<float>) * p.nzp);
  if (e == hipSuccess) e = hipMalloc(&h->dev.tabd, sizeof(crava::PerK<double>) * p.nzp);
  if (e == hipSuccess)
    e = hipMemcpy(h->dev.tabf, tf.data(), sizeof(crava::PerK<float>) * p.nzp, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->dev.tabd, td.data(), sizeof(crava::PerK<double>) * p.nzp, hipMemcpyHostToDevice);
  // the in-band k-planes first: the solve runs over those, a light kernel
  // over the rest
  const std::vector<int> kmap = crava::posterior_kmap(p.nzp, invert, &h->dev.nk_in);
  h->dev.invert.assign(invert, invert + p.nzp);
  if (e == hipSuccess) e = hipMalloc(&h->dev.kmap, sizeof(int) * p.nzp);
  if (e == hipSuccess) e = hipMemcpy(h->dev.kmap, kmap.data(), sizeof(int) * p.nzp, hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    set_error(std::string("posterior table upload: ") + hipGetErrorString(e));
    crava_hip_post_destroy(h);
    return CRAVA_HIP_E_HIP;
  }
  *post = h;
  return CRAVA_HIP_OK;
}

int crava_hip_post_destroy(crava_hip_post *post) {
  if (!post) return CRAVA_HIP_OK;
  if (post->dev.tabf) (void)hipFree(post->dev.tabf);
  if (post->dev.tabd) (void)hipFree(post->dev.tabd);
  if (post->dev.kmap) (void)hipFree(post->dev.kmap);
  if (post->nll_tab) (void)hipFree(post->nll_tab);
  if (post->nll_klist) (void)hipFree(post->nll_klist);
  if (post->nll_partial) (void)hipFree(post->nll_partial);
  if (post->queue_mem) (void)hipFree(post->queue_mem);
  delete post;
  return CRAVA_HIP_OK;
}

}  // extern "C"

// Shared body of the posterior entry points: argument checks, the queue of
// the mixed form (the handle's, allocated on first use, or the static one),
// the launches and the optional counts.
static int post_run_impl(const char *name, crava_hip_post *post, int precision, void *const *mu, void *const *data,
                         void *const *cov, const void *corr, const void *errcorr, long long *n_fail, long long *n_fp64,
                         bool static_q) {
  if (!post || !mu || !data || !cov || !errcorr) {
    set_error(std::string("NULL argument to ") + name);
    return CRAVA_HIP_E_INVALID;
  }
  if (precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F64) { set_error("invalid precision"); return CRAVA_HIP_E_INVALID; }
  if (static_q && precision == CRAVA_HIP_MIXED && post->dev.defer_need > kStaticQueueEntries) {
    set_error("the mixed posterior needs a queue of " + std::to_string(post->dev.defer_need) +
              " entries; the static queue holds " + std::to_string(kStaticQueueEntries));
    return CRAVA_HIP_E_INVALID;
  }
  for (int c = 0; c < 3; ++c) if (!mu[c]) { set_error("NULL mean grid"); return CRAVA_HIP_E_INVALID; }
  for (int q = 0; q < 6; ++q) if (!cov[q]) { set_error("NULL covariance grid"); return CRAVA_HIP_E_INVALID; }
  for (int l = 0; l < post->dev.pp.ntheta; ++l) if (!data[l]) { set_error("NULL data grid"); return CRAVA_HIP_E_INVALID; }
  crava_hip_ctx *ctx = post->ctx;
  crava::DeferQueue sq;
  const crava::DeferQueue *q = nullptr;
  int rc;
  if (precision == CRAVA_HIP_MIXED) {
    if (static_q) {
      if ((rc = static_queue(&sq))) return rc;
      q = &sq;
    } else {
      if (!post->queue_mem) {
        const size_t n = (size_t)std::max<long long>(post->dev.defer_need, 1);
        CRAVA_HIP_CHECK(hipMalloc(&post->queue_mem, 256 + sizeof(int2) * n));
        post->queue.count = static_cast<unsigned *>(post->queue_mem);
        post->queue.list = reinterpret_cast<int2 *>(static_cast<unsigned char *>(post->queue_mem) + 256);
        post->queue.cap = (long long)n;
      }
      q = &post->queue;
    }
  }
  const bool want = n_fail || n_fp64;
  if (want) {
    hipLaunchKernelGGL(k_zero_counts2, dim3(1), dim3(64), 0, ctx->stream, ctx->counters);
    CRAVA_HIP_CHECK(hipGetLastError());
  }
  rc = crava::launch_posterior(post->dev, precision, mu, data, cov, errcorr, want ? ctx->counters : nullptr, q,
                               ctx->stream, corr);
  if (rc) return rc;
  if (want) {
    unsigned long long h[2] = {0, 0};
    CRAVA_HIP_CHECK(hipMemcpyAsync(h, ctx->counters, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
    CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (n_fail) *n_fail = (long long)h[0];
    if (n_fp64) *n_fp64 = (long long)h[1];
  }
  return CRAVA_HIP_OK;
}

extern "C" {

int crava_hip_post_run(crava_hip_post *post, int precision, void *const *mu, void *const *data, void *const *cov,
                       const void *errcorr, long long *n_fail, long long *n_fp64) {
  return post_run_impl("crava_hip_post_run", post, precision, mu, data, cov, nullptr, errcorr, n_fail, n_fp64, false);
}

int crava_hip_post_run_separable(crava_hip_post *post, int precision, void *const *mu, void *const *data,
                                 const void *corr, const void *errcorr, void *const *cov_out, long long *n_fail,
                                 long long *n_fp64) {
  if (!corr) { set_error("NULL correlation grid"); return CRAVA_HIP_E_INVALID; }
  return post_run_impl("crava_hip_post_run_separable", post, precision, mu, data, cov_out, corr, errcorr, n_fail,
                       n_fp64, false);
}

int crava_hip_post_run_static(crava_hip_post *post, int precision, void *const *mu, void *const *data,
                              void *const *cov, const void *errcorr, long long *n_fail, long long *n_fp64) {
  return post_run_impl("crava_hip_post_run_static", post, precision, mu, data, cov, nullptr, errcorr, n_fail, n_fp64,
                       true);
}

int crava_hip_nll_grad(crava_hip_post *post, int precision, void *const *mu, void *const *data, void *const *cov,
                       const void *errcorr, const double *scales, const double *noise_var, const double *chol0,
                       double *nll, double *grad) {
  if (!post || !mu || !data || !cov || !errcorr || !scales || !noise_var || !chol0 || !nll || !grad) {
    set_error("NULL argument to crava_hip_nll_grad");
    return CRAVA_HIP_E_INVALID;
  }
  if (precision != CRAVA_HIP_F32 && precision != CRAVA_HIP_F64) { set_error("grids must be float or double"); return CRAVA_HIP_E_INVALID; }
  const crava::PostProblem &pp = post->dev.pp;
  const int nt = pp.ntheta;
  for (int l = 0; l < nt; ++l)
    if (!(scales[l] > 0.0) || !(noise_var[l] > 0.0)) { set_error("scales and noise variances must be positive"); return CRAVA_HIP_E_INVALID; }
  crava_hip_ctx *ctx = post->ctx;
  // tables at the current parameters
  std::vector<crava::NllTab> tab(pp.nzp);
  std::vector<int> klist;
  const double c0 = 0.5 * (1.0 - (double)pp.wnc);
  double sig[crava::kMaxTheta * crava::kMaxTheta];
  for (int l = 0; l < nt; ++l)
    for (int m = 0; m < nt; ++m) {
      const double bl = pp.err_theta_cov[l * nt + l], bm = pp.err_theta_cov[m * nt + m];
      const double rho = (bl > 0 && bm > 0) ? pp.err_theta_cov[l * nt + m] / std::sqrt(bl * bm) : (l == m ? 1.0 : 0.0);
      sig[l * nt + m] = std::sqrt(noise_var[l] * noise_var[m]) * rho;
    }
  for (int k = 0; k < pp.nzp; ++k) {
    crava::NllTab t{};
    if (post->invert[k]) klist.push_back(k);
    crava::cx<double> w[crava::kMaxTheta], a1[crava::kMaxTheta], a2[crava::kMaxTheta];
    for (int l = 0; l < nt; ++l) {
      const size_t o = ((size_t)k * nt + l) * 2;
      w[l] = crava::mk((double)post->kWD[o] * scales[l], (double)post->kWD[o + 1] * scales[l]);
      a1[l] = crava::mk((double)post->e1[o], (double)post->e1[o + 1]);
      a2[l] = crava::mk((double)post->e2[o], (double)post->e2[o + 1]);
      t.kwd[l] = w[l];
      const double e3 = post->e3[o];
      t.wd[l] = (double)pp.wnc * e3 * e3;
      t.Ed[l] = t.wd[l] * sig[l * nt + l];
    }
    for (int l = 0; l < nt; ++l)
      for (int m = 0; m <= l; ++m) {
        const int q = l * (l + 1) / 2 + m;
        crava::cx<double> g = crava::add(crava::mulc(a1[l], a1[m]), crava::mulc(a2[l], a2[m]));
        if (l == m) g.y = 0.0;
        t.g[q] = g;
        t.Ec[q] = crava::scl(g, c0 * sig[l * nt + m]);
      }
    t.c0 = c0;
    tab[k] = t;
  }
  // Sigma0 = L0 L0^T
  const double L00 = chol0[0], L10 = chol0[1], L11 = chol0[2], L20 = chol0[3], L21 = chol0[4], L22 = chol0[5];
  const double S[3][3] = {{L00 * L00, L00 * L10, L00 * L20},
                          {L10 * L00, L10 * L10 + L11 * L11, L10 * L20 + L11 * L21},
                          {L20 * L00, L20 * L10 + L21 * L11, L20 * L20 + L21 * L21 + L22 * L22}};
  crava::NllCall c{};
  c.nt = nt;
  c.nxp = pp.nxp;
  c.nyp = pp.nyp;
  c.dbl = precision == CRAVA_HIP_F64;
  for (int q = 0; q < 3; ++q) c.mu[q] = mu[q];
  for (int l = 0; l < nt; ++l) c.data[l] = data[l];
  c.cov0 = cov[0];
  c.errcorr = errcorr;
  c.ec_plane = post->dev.ec_plane;
  c.n_inband = (int)klist.size();
  for (int i = 0; i < crava::kMaxTheta * 3; ++i) c.A[i] = pp.A[i];
  c.sig0[0] = S[0][0]; c.sig0[1] = S[1][1]; c.sig0[2] = S[2][2];
  c.sig0[3] = S[0][1]; c.sig0[4] = S[0][2]; c.sig0[5] = S[1][2];
  c.base00 = pp.sigma0[0];
  const int na = crava::nll_accumulator_count(nt);
  if (!post->nll_tab) {
    CRAVA_HIP_CHECK(hipMalloc(&post->nll_tab, sizeof(crava::NllTab) * pp.nzp));
    CRAVA_HIP_CHECK(hipMalloc(&post->nll_klist, sizeof(int) * pp.nzp));
    post->nll_blocks = 1024;
    CRAVA_HIP_CHECK(hipMalloc(&post->nll_partial, sizeof(double) * post->nll_blocks * 64));
  }
  CRAVA_HIP_CHECK(hipMemcpyAsync(post->nll_tab, tab.data(), sizeof(crava::NllTab) * pp.nzp, hipMemcpyHostToDevice, ctx->stream));
  if (!klist.empty())
    CRAVA_HIP_CHECK(hipMemcpyAsync(post->nll_klist, klist.data(), sizeof(int) * klist.size(), hipMemcpyHostToDevice, ctx->stream));
  c.tab = post->nll_tab;
  c.klist = post->nll_klist;
  c.partial = post->nll_partial;
  const long long total = (long long)c.n_inband * (pp.nxp / 2 + 1) * pp.nyp;
  c.blocks = (int)std::max<long long>(1, std::min<long long>(post->nll_blocks, (total + 255) / 256));
  std::vector<double> part((size_t)c.blocks * na, 0.0);
  if (total > 0) {
    int rc = crava::launch_nll(c, ctx->stream);
    if (rc) return rc;
    CRAVA_HIP_CHECK(hipMemcpyAsync(part.data(), post->nll_partial, sizeof(double) * part.size(), hipMemcpyDeviceToHost, ctx->stream));
  }
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  std::vector<double> acc(na, 0.0);
  for (int b = 0; b < c.blocks; ++b)
    for (int v = 0; v < na; ++v) acc[v] += part[(size_t)b * na + v];
  const int kS = 1, kN = 1 + nt, kP = 1 + nt + nt * (nt + 1) / 2;
  *nll = 0.5 * acc[0];
  for (int l = 0; l < nt; ++l) grad[l] = 0.5 * acc[kS + l] / scales[l];
  for (int l = 0; l < nt; ++l) {
    double gsum = acc[kN + l * (l + 1) / 2 + l];
    for (int m = 0; m < nt; ++m) {
      if (m == l) continue;
      const int hi = l > m ? l : m, lo = l > m ? m : l;
      gsum += acc[kN + hi * (hi + 1) / 2 + lo] * sig[l * nt + m] / (2.0 * noise_var[l]);
    }
    grad[nt + l] = 0.5 * gsum;
  }
  // dNLL/dSigma0 (symmetric P) -> dNLL/dL0 = 2 * (1/2) P L0 on the lower triangle
  const double P[3][3] = {{acc[kP + 0], acc[kP + 3], acc[kP + 4]},
                          {acc[kP + 3], acc[kP + 1], acc[kP + 5]},
                          {acc[kP + 4], acc[kP + 5], acc[kP + 2]}};
  const double Lm[3][3] = {{L00, 0, 0}, {L10, L11, 0}, {L20, L21, L22}};
  const int li[6] = {0, 1, 1, 2, 2, 2}, lj[6] = {0, 0, 1, 0, 1, 2};
  for (int q = 0; q < 6; ++q) {
    double v = 0.0;
    for (int k = 0; k < 3; ++k) v += P[li[q]][k] * Lm[k][lj[q]];
    grad[2 * nt + q] = v;
  }
  return CRAVA_HIP_OK;
}

int crava_hip_fill_complex_noise(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, uint64_t seed,
                                 uint64_t stream) {
  if (!ctx || !grid || nxp < 2 || nyp < 1 || nzp < 1) { set_error("invalid arguments to crava_hip_fill_complex_noise"); return CRAVA_HIP_E_INVALID; }
  return crava::launch_fill_noise(grid, nxp, nyp, nzp, seed, stream, ctx->stream);
}

static int sync_counter(crava_hip_ctx *ctx, long long *out) {
  unsigned long long h = 0;
  CRAVA_HIP_CHECK(hipMemcpyAsync(&h, ctx->counters, sizeof(h), hipMemcpyDeviceToHost, ctx->stream));
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  *out = (long long)h;
  return CRAVA_HIP_OK;
}

int crava_hip_simulate_points(crava_hip_ctx *ctx, long long npts, void *const *cov, void *const *seed,
                              long long *n_fail) {
  if (!ctx || !cov || !seed || npts < 0) { set_error("invalid arguments to crava_hip_simulate_points"); return CRAVA_HIP_E_INVALID; }
  for (int q = 0; q < 6; ++q) if (!cov[q]) { set_error("NULL covariance grid"); return CRAVA_HIP_E_INVALID; }
  for (int c = 0; c < 3; ++c) if (!seed[c]) { set_error("NULL seed grid"); return CRAVA_HIP_E_INVALID; }
  if (npts == 0) { if (n_fail) *n_fail = 0; return CRAVA_HIP_OK; }
  if (n_fail) CRAVA_HIP_CHECK(hipMemsetAsync(ctx->counters, 0, sizeof(unsigned long long), ctx->stream));
  int rc = crava::launch_simulate_points(npts, cov, seed, n_fail ? ctx->counters : nullptr, ctx->stream);
  if (rc) return rc;
  return n_fail ? sync_counter(ctx, n_fail) : CRAVA_HIP_OK;
}

int crava_hip_simulate(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, void *const *cov, void *const *seed,
                       void *const *mean, long long *n_fail) {
  if (!ctx || !cov || !seed || nxp < 2 || nyp < 1 || nzp < 1) { set_error("invalid arguments to crava_hip_simulate"); return CRAVA_HIP_E_INVALID; }
  const long long npts = (long long)(nxp / 2 + 1) * nyp * nzp;
  long long nf = 0;
  int rc = crava_hip_simulate_points(ctx, npts, cov, seed, n_fail ? &nf : nullptr);
  if (rc) return rc;
  if ((rc = crava_hip_fft_c2r_many(ctx, seed, 3, nxp, nyp, nzp, CRAVA_HIP_PARAMETER, CRAVA_HIP_F32))) return rc;
  for (int c = 0; c < 3; ++c) {
    if (mean && mean[c]) {
      const long long rsize = (long long)2 * (nxp / 2 + 1) * nyp * nzp;
      if ((rc = crava::launch_add_real(reinterpret_cast<float *>(seed[c]), reinterpret_cast<const float *>(mean[c]),
                                       rsize, ctx->stream)))
        return rc;
    }
  }
  if (n_fail) *n_fail = nf;
  return CRAVA_HIP_OK;
}

int crava_hip_facies(crava_hip_ctx *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp, const float *alpha,
                     const float *beta, const float *rho, int nfac, const float *tables, int tn1, int tn2, int tn3,
                     const double *volume, const int *vn, const float *prior, float p_undefined, float *const *prob,
                     int *argmax, unsigned long long *counts) {
  if (!ctx || !alpha || !beta || !rho || !tables || !volume || !vn || !prior || !prob) {
    set_error("NULL argument to crava_hip_facies");
    return CRAVA_HIP_E_INVALID;
  }
  if (nx < 1 || ny < 1 || nz < 1 || nx > nxp || ny > nyp || nz > nzp || tn1 < 1 || tn2 < 1 || tn3 < 1 ||
      vn[0] < 1 || vn[1] < 1 || vn[2] < 1 || vn[0] > tn1 || vn[1] > tn2 || vn[2] > tn3) {
    set_error("invalid sizes in crava_hip_facies");
    return CRAVA_HIP_E_INVALID;
  }
  for (int l = 0; l <= nfac && l <= crava::kMaxFacies; ++l)
    if (!prob[l]) { set_error("NULL probability grid"); return CRAVA_HIP_E_INVALID; }
  crava::FaciesCall c{nx, ny, nz, nxp, nyp, nzp, alpha, beta, rho, nfac, tables, tn1, tn2, tn3,
                      volume, vn, prior, p_undefined, prob, argmax, counts};
  if (counts && nfac >= 1 && nfac <= crava::kMaxFacies)
    CRAVA_HIP_CHECK(hipMemsetAsync(counts, 0, sizeof(unsigned long long) * nfac, ctx->stream));
  return crava::launch_facies(c, ctx->stream);
}

int crava_hip_facies_local_noise(crava_hip_ctx *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp,
                                 const float *alpha, const float *beta, const float *rho, int nfac, int nang,
                                 const float *tables, int tn1, int tn2, int tn3, const double *volumes,
                                 const int *vn, const double *const *noise_scale, const float *prior,
                                 float p_undefined, float *const *prob, int *argmax, unsigned long long *counts) {
  if (!ctx || !alpha || !beta || !rho || !tables || !volumes || !vn || !prior || !prob ||
      (nang > 0 && !noise_scale)) {
    set_error("NULL argument to crava_hip_facies_local_noise");
    return CRAVA_HIP_E_INVALID;
  }
  if (nang < 0 || nang > crava::kMaxDensAngles) { set_error("nang must be 0..8"); return CRAVA_HIP_E_INVALID; }
  const int ndens = 1 << nang;
  if (nx < 1 || ny < 1 || nz < 1 || nx > nxp || ny > nyp || nz > nzp || tn1 < 1 || tn2 < 1 || tn3 < 1 ||
      nfac < 1 || nfac > crava::kMaxFacies) {
    set_error("invalid sizes in crava_hip_facies_local_noise");
    return CRAVA_HIP_E_INVALID;
  }
  for (int d = 0; d < ndens; ++d)
    if (vn[3 * d] < 1 || vn[3 * d + 1] < 1 || vn[3 * d + 2] < 1 || vn[3 * d] > tn1 || vn[3 * d + 1] > tn2 ||
        vn[3 * d + 2] > tn3) {
      set_error("density lattice larger than its table");
      return CRAVA_HIP_E_INVALID;
    }
  for (int l = 0; l <= nfac; ++l)
    if (!prob[l]) { set_error("NULL probability grid"); return CRAVA_HIP_E_INVALID; }
  for (int g = 0; g < nang; ++g)
    if (!noise_scale[g]) { set_error("NULL noise-scale table"); return CRAVA_HIP_E_INVALID; }
  // angle weights t = (scale - min) / (max - min) per angle, 0 where the
  // table is constant (src/faciesprob.cpp:2168-2179), Grid2D::FindMin /
  // FindMax skipping RMISSING; float(t) as the loop reads them (:2221-2222)
  const size_t ntr = (size_t)nx * ny;
  std::vector<crava::FaciesVolume> vols(ndens);
  for (int d = 0; d < ndens; ++d) {
    const double *v = volumes + 6 * d;
    vols[d] = crava::FaciesVolume{v[0], v[1], v[2], v[3], v[4], v[5], vn[3 * d], vn[3 * d + 1], vn[3 * d + 2], 0};
  }
  std::vector<float> tw((size_t)nang * ntr);
  const double missing = -99999.000;
  for (int g = 0; g < nang; ++g) {
    const double *tab = noise_scale[g];
    double mn = tab[0], mx = tab[0];
    for (size_t t = 0; t < ntr; ++t) {
      if ((mn == missing || tab[t] < mn) && tab[t] != missing) mn = tab[t];
      if ((mx == missing || tab[t] > mx) && tab[t] != missing) mx = tab[t];
    }
    for (size_t t = 0; t < ntr; ++t) tw[(size_t)g * ntr + t] = mn == mx ? 0.0f : (float)((tab[t] - mn) / (mx - mn));
  }
  void *ws = nullptr;
  const size_t b_vol = sizeof(crava::FaciesVolume) * ndens, b_t = sizeof(float) * tw.size();
  int rc = ctx->workspace(b_vol + b_t + 256, &ws);
  if (rc) return rc;
  auto *dvol = static_cast<crava::FaciesVolume *>(ws);
  float *dt = reinterpret_cast<float *>(static_cast<char *>(ws) + ((b_vol + 255) / 256) * 256);
  CRAVA_HIP_CHECK(hipMemcpyAsync(dvol, vols.data(), b_vol, hipMemcpyHostToDevice, ctx->stream));
  if (b_t) CRAVA_HIP_CHECK(hipMemcpyAsync(dt, tw.data(), b_t, hipMemcpyHostToDevice, ctx->stream));
  crava::FaciesCall c{nx, ny, nz, nxp, nyp, nzp, alpha, beta, rho, nfac, tables, tn1, tn2, tn3,
                      volumes, vn, prior, p_undefined, prob, argmax, counts};
  c.ndens = ndens;
  c.nang = nang;
  c.vols = dvol;
  c.tgrid = dt;
  if (counts) CRAVA_HIP_CHECK(hipMemsetAsync(counts, 0, sizeof(unsigned long long) * nfac, ctx->stream));
  if ((rc = crava::launch_facies(c, ctx->stream))) return rc;
  // the host vectors are staged by the copies; wait so the workspace is not reused early
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return CRAVA_HIP_OK;
}

int crava_hip_kde_table(crava_hip_ctx *ctx, long long npts, const double *d1, const double *d2, const double *d3,
                        const double *sigma, int n1, int n2, int n3, const double *limits, void *table) {
  if (!ctx || !sigma || !limits || !table || npts < 0 || (npts > 0 && (!d1 || !d2 || !d3)) || n1 < 2 || n2 < 1 ||
      n3 < 1 || (long long)n1 * n2 * n3 > (1LL << 26)) {
    set_error("invalid arguments to crava_hip_kde_table");
    return CRAVA_HIP_E_INVALID;
  }
  for (int d = 0; d < 3; ++d)
    if (!std::isfinite(limits[2 * d]) || !std::isfinite(limits[2 * d + 1]) || !(limits[2 * d + 1] > limits[2 * d])) {
      set_error("crava_hip_kde_table: each dimension needs finite limits with max > min");
      return CRAVA_HIP_E_INVALID;
    }
  const size_t rsize = (size_t)(2 * (n1 / 2 + 1)) * n2 * n3;
  std::vector<float> hist(rsize), smooth(rsize);
  double s3[3][3];
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q) s3[r][q] = sigma[r * 3 + q];
  if (!crava::kde_smoother(s3, n1, n2, n3, limits, smooth.data())) {
    set_error("crava_hip_kde_table: the smoothing covariance is not positive definite");
    return CRAVA_HIP_E_INVALID;
  }
  crava::kde_histogram(npts, d1, d2, d3, n1, n2, n3, limits, hist.data());
  void *ws = nullptr;
  int rc = ctx->workspace(sizeof(float) * rsize, &ws);
  if (rc) return rc;
  CRAVA_HIP_CHECK(hipMemcpyAsync(table, hist.data(), sizeof(float) * rsize, hipMemcpyHostToDevice, ctx->stream));
  CRAVA_HIP_CHECK(hipMemcpyAsync(ws, smooth.data(), sizeof(float) * rsize, hipMemcpyHostToDevice, ctx->stream));
  // FFTGrid::fftInPlace of both (PARAMETER: 1/sqrt(N)), multiply, inverse
  // (1/sqrt(N)) and multiplyByScalar(sqrt(float(N))) (src/posteriorelasticpdf3d.cpp:80-129)
  void *both[2] = {table, ws};
  if ((rc = fft_common(ctx, both, 2, n1, n2, n3, CRAVA_HIP_PARAMETER, CRAVA_HIP_F32, false))) return rc;
  if ((rc = crava_hip_grid_op(ctx, CRAVA_HIP_OP_MULTIPLY, table, ws, n1, n2, n3, 1, 0.0f))) return rc;
  if ((rc = fft_common(ctx, &table, 1, n1, n2, n3, CRAVA_HIP_PARAMETER, CRAVA_HIP_F32, true))) return rc;
  if ((rc = crava_hip_grid_op(ctx, CRAVA_HIP_OP_MULTIPLY_SCALAR, table, nullptr, n1, n2, n3, 0,
                              std::sqrt(float(n1 * n2 * n3)))))
    return rc;
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return CRAVA_HIP_OK;
}

int crava_hip_facies_densities(crava_hip_ctx *ctx, int nsamples, const float *alpha, const float *beta,
                               const float *rho, const int *facies_log, int nfac, int ndens,
                               const double *sample_transform, const double *sigma_e, int no_vs, int n1, int n2,
                               int n3, void *const *tables, double *volumes, int *vn) {
  if (!ctx || nsamples < 2 || !alpha || !beta || !rho || !facies_log || nfac < 1 || nfac > crava::kMaxFacies ||
      ndens < 1 || ndens > (1 << crava::kMaxDensAngles) || !tables || !volumes || !vn) {
    set_error("invalid arguments to crava_hip_facies_densities");
    return CRAVA_HIP_E_INVALID;
  }
  for (int i = 0; i < ndens * nfac; ++i)
    if (!tables[i]) { set_error("NULL table grid"); return CRAVA_HIP_E_INVALID; }
  int nobs = 0;
  for (int i = 0; i < nsamples; ++i) nobs += facies_log[i] != crava::kIMissing;
  if (nobs < 2) { set_error("crava_hip_facies_densities needs two samples with a facies"); return CRAVA_HIP_E_INVALID; }
  std::vector<float> a(nsamples), b(nsamples), r(nsamples);
  std::vector<double> fa, fb, fr;
  for (int d = 0; d < ndens; ++d) {
    // the samples of density d: H (alpha, beta, rho) for d > 0
    // (src/faciesprob.cpp:967-1003), float
    for (int i = 0; i < nsamples; ++i) {
      if (d > 0 && sample_transform) {
        const double *h = sample_transform + 9 * d;
        const double h1[3] = {(double)alpha[i], (double)beta[i], (double)rho[i]};
        double h2[3];
        for (int q = 0; q < 3; ++q) h2[q] = h[q * 3 + 0] * h1[0] + h[q * 3 + 1] * h1[1] + h[q * 3 + 2] * h1[2];
        a[i] = (float)h2[0];
        b[i] = (float)h2[1];
        r[i] = (float)h2[2];
      } else {
        a[i] = alpha[i];
        b[i] = beta[i];
        r[i] = rho[i];
      }
    }
    double sig[3][3], lim[6];
    crava::facies_density_setup(nsamples, a.data(), b.data(), r.data(), facies_log,
                                sigma_e ? sigma_e + 9 * d : nullptr, no_vs != 0, n1, n2, n3, sig, lim);
    // the Simbox of the lattice (src/faciesprob.cpp:1100-1102, Simbox ctor
    // src/simbox.cpp:35-66): n = int(0.5 + l / d), base = top + lz
    const double dA = (lim[1] - lim[0]) / n1, dB = (lim[3] - lim[2]) / n2, dR = (lim[5] - lim[4]) / n3;
    double *vol = volumes + 6 * d;
    vol[0] = lim[0];
    vol[1] = lim[2];
    vol[2] = dA;
    vol[3] = dB;
    vol[4] = lim[4];
    vol[5] = lim[4] + (lim[5] - lim[4]);
    vn[3 * d] = int(0.5 + (lim[1] - lim[0]) / dA);
    vn[3 * d + 1] = int(0.5 + (lim[3] - lim[2]) / dB);
    vn[3 * d + 2] = int(0.5 + (lim[5] - lim[4]) / dR);
    double sflat[9];
    for (int q = 0; q < 9; ++q) sflat[q] = sig[q / 3][q % 3];
    for (int l = 0; l < nfac; ++l) {
      fa.clear();
      fb.clear();
      fr.clear();
      for (int i = 0; i < nsamples; ++i)
        if (facies_log[i] == l) {
          fa.push_back(a[i]);
          fb.push_back(b[i]);
          fr.push_back(r[i]);
        }
      int rc = crava_hip_kde_table(ctx, (long long)fa.size(), fa.data(), fb.data(), fr.data(), sflat, n1, n2, n3, lim,
                                   tables[d * nfac + l]);
      if (rc) return rc;
    }
  }
  return CRAVA_HIP_OK;
}

int crava_hip_local_noise_sigmanew(int ntheta, const double *G, const double *errThetaCov, const double *priorVar0,
                                   const double *postVar0, const double *scale, double *sigmanew) {
  if (ntheta < 1 || ntheta > crava::kLncMaxTheta || !G || !errThetaCov || !priorVar0 || !postVar0 || !scale ||
      !sigmanew) {
    set_error("invalid arguments to crava_hip_local_noise_sigmanew");
    return CRAVA_HIP_E_INVALID;
  }
  for (int a = 0; a < ntheta; ++a)
    if (!(scale[a] >= 0.0)) { set_error("noise scales must be >= 0"); return CRAVA_HIP_E_INVALID; }
  crava::LncConst c;
  crava::lnc_const_host(ntheta, G, errThetaCov, priorVar0, postVar0, c);
  return crava::lnc_sigmanew_host(c, ntheta, scale, sigmanew);
}

// ---------------------------------------------------------------------------
// Time-lapse merge
static int merge_common(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, crava::Merge4D &a) {
  if (!ctx || nxp < 1 || nyp < 1 || nzp < 1) { set_error("invalid arguments to the 4D merge"); return CRAVA_HIP_E_INVALID; }
  a.n = (long long)(nxp / 2 + 1) * nyp * nzp;
  return CRAVA_HIP_OK;
}

int crava_hip_merge_mean_4d(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, const void *const *mu_static,
                            const void *const *mu_dynamic, void *const *mu) {
  crava::Merge4D a{};
  int rc = merge_common(ctx, nxp, nyp, nzp, a);
  if (rc) return rc;
  if (!mu_static || !mu_dynamic || !mu) { set_error("NULL grid list"); return CRAVA_HIP_E_INVALID; }
  for (int l = 0; l < 3; ++l) {
    if (!mu_static[l] || !mu_dynamic[l] || !mu[l]) { set_error("NULL grid"); return CRAVA_HIP_E_INVALID; }
    a.ms[l] = static_cast<const crava::cx<float> *>(mu_static[l]);
    a.md[l] = static_cast<const crava::cx<float> *>(mu_dynamic[l]);
    a.mu[l] = static_cast<crava::cx<float> *>(mu[l]);
  }
  return crava::launch_merge_4d(a, false, ctx->stream);
}

int crava_hip_merge_cov_4d(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, const void *const *ss,
                           const void *const *dd, const void *const *sd, void *const *sigma) {
  crava::Merge4D a{};
  int rc = merge_common(ctx, nxp, nyp, nzp, a);
  if (rc) return rc;
  if (!ss || !dd || !sd || !sigma) { set_error("NULL grid list"); return CRAVA_HIP_E_INVALID; }
  for (int q = 0; q < 6; ++q) {
    if (!ss[q] || !dd[q] || !sigma[q]) { set_error("NULL grid"); return CRAVA_HIP_E_INVALID; }
    a.ss[q] = static_cast<const crava::cx<float> *>(ss[q]);
    a.dd[q] = static_cast<const crava::cx<float> *>(dd[q]);
    a.sig[q] = static_cast<crava::cx<float> *>(sigma[q]);
  }
  for (int q = 0; q < 9; ++q) {
    if (!sd[q]) { set_error("NULL grid"); return CRAVA_HIP_E_INVALID; }
    a.sd[q] = static_cast<const crava::cx<float> *>(sd[q]);
  }
  // each output point depends on the same point of the inputs only, so the
  // outputs may alias ss or dd
  return crava::launch_merge_4d(a, true, ctx->stream);
}

// ---------------------------------------------------------------------------
// Forward modelling
int crava_hip_synthetic_seismic(crava_hip_ctx *ctx, const crava_hip_synt_params *p, const void *alpha,
                                const void *beta, const void *rho, void *const *out) {
  if (!ctx || !p || !alpha || !beta || !rho || !out || !p->A || !p->wavelets) {
    set_error("NULL argument to crava_hip_synthetic_seismic");
    return CRAVA_HIP_E_INVALID;
  }
  const int nt = p->ntheta;
  if (nt < 1 || nt > crava::kFwdMaxTheta) { set_error("ntheta must be 1..8"); return CRAVA_HIP_E_INVALID; }
  if (p->nx < 1 || p->ny < 1 || p->nz < 1 || p->nx > p->nxp || p->ny > p->nyp || p->nz > p->nzp || p->nzp < 2) {
    set_error("invalid volume / padded grid sizes");
    return CRAVA_HIP_E_INVALID;
  }
  for (int l = 0; l < nt; ++l)
    if (!out[l]) { set_error("NULL output grid"); return CRAVA_HIP_E_INVALID; }
  bool any_shift = false;
  for (int l = 0; l < nt; ++l) any_shift = any_shift || (p->shift && p->shift[l]);
  if (p->stretch)
    for (int l = 0; l < nt; ++l)
      if (p->stretch[l])
        for (size_t t = 0; t < (size_t)p->nx * p->ny; ++t)
          if (!(p->stretch[l][t] > 0.0f) || !std::isfinite(p->stretch[l][t])) {
            set_error("stretch factors must be positive and finite");
            return CRAVA_HIP_E_INVALID;
          }
  if (any_shift && !(p->dz > 0.0f)) { set_error("dz must be positive when time shifts are given"); return CRAVA_HIP_E_INVALID; }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(p->nxp, p->nyp, p->nzp, 0, &f);
  if (rc) return rc;
  // tables: wavelet half spectra, then the per-trace arrays that are given
  const int nk = p->nzp / 2 + 1;
  const size_t ntr = (size_t)p->nx * p->ny;
  std::vector<float> host((size_t)nt * nk * 2);
  for (int l = 0; l < nt; ++l)
    crava::wavelet_half_spectrum(p->wavelets + (size_t)l * p->nzp, p->nzp, host.data() + (size_t)l * nk * 2);
  const float *const *tabs[3] = {p->stretch, p->shift, p->gain};
  size_t ntab = 0;
  for (int t = 0; t < 3; ++t)
    for (int l = 0; l < nt; ++l) ntab += (tabs[t] && tabs[t][l]) ? 1 : 0;
  const size_t spec_bytes = ((host.size() * sizeof(float) + 255) / 256) * 256;
  const size_t tab_bytes = ((ntr * sizeof(float) + 255) / 256) * 256;
  void *ws = nullptr;
  if ((rc = ctx->workspace(spec_bytes + ntab * tab_bytes, &ws))) return rc;
  char *wp = static_cast<char *>(ws);
  CRAVA_HIP_CHECK(hipMemcpyAsync(wp, host.data(), host.size() * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  crava::FwdArgs a{};
  a.alpha = static_cast<const float *>(alpha);
  a.beta = static_cast<const float *>(beta);
  a.rho = static_cast<const float *>(rho);
  a.W = reinterpret_cast<const crava::cx<float> *>(wp);
  size_t off = spec_bytes;
  const float **dst[3] = {a.sf, a.shift, a.gain};
  for (int t = 0; t < 3; ++t)
    for (int l = 0; l < nt; ++l) {
      dst[t][l] = nullptr;
      if (tabs[t] && tabs[t][l]) {
        CRAVA_HIP_CHECK(hipMemcpyAsync(wp + off, tabs[t][l], ntr * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
        dst[t][l] = reinterpret_cast<const float *>(wp + off);
        off += tab_bytes;
      }
    }
  // the host tables above go out of scope: finish the uploads first
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  for (int l = 0; l < nt; ++l) {
    a.out[l] = static_cast<float *>(out[l]);
    for (int c = 0; c < 3; ++c) a.A[l][c] = p->A[l * 3 + c];
  }
  a.dz = p->dz;
  a.inv_n = (float)(1.0 / (double)p->nzp);
  a.nt = nt;
  a.nx = p->nx;
  a.ny = p->ny;
  a.nz = p->nz;
  a.nyp = p->nyp;
  a.nzp = p->nzp;
  a.rnxp = 2 * (p->nxp / 2 + 1);
  if (crava::fm_fused_available(p->nzp) && f->stw[2]) return crava::launch_fm_fused(a, f->stw[2], ctx->stream);
  if ((rc = crava::launch_fm_reflect(a, ctx->stream))) return rc;
  if ((rc = f->pass(2, out, nt, false, nullptr, ctx->stream))) return rc;
  if ((rc = crava::launch_fm_filter(a, ctx->stream))) return rc;
  if ((rc = f->pass(2, out, nt, true, nullptr, ctx->stream))) return rc;
  return crava::launch_fm_padding(a, ctx->stream);
}


// ---------------------------------------------------------------------------
// Non-constant-thickness data preparation

int crava_hip_divide_by_scale_wavelet(crava_hip_ctx *ctx, const crava_hip_swdiv_params *p, void *const *grids) {
  if (!ctx || !p || !grids || !p->wavelets || !p->wavelet_norm || !p->A || !p->prior_var0 || !p->corr_t ||
      !p->err_var || !p->rel_thick) {
    set_error("NULL argument to crava_hip_divide_by_scale_wavelet");
    return CRAVA_HIP_E_INVALID;
  }
  const int nt = p->ntheta;
  if (nt < 1 || nt > crava::kSwdMaxTheta) { set_error("ntheta must be 1..8"); return CRAVA_HIP_E_INVALID; }
  if (p->nx < 1 || p->ny < 1 || p->nz < 1 || p->nx > p->nxp || p->ny > p->nyp || p->nz > p->nzp || p->nzp < 2) {
    set_error("invalid volume / padded grid sizes");
    return CRAVA_HIP_E_INVALID;
  }
  if (!(p->dz > 0.0f) || !(p->lz > 0.0)) { set_error("dz and lz must be positive"); return CRAVA_HIP_E_INVALID; }
  for (int l = 0; l < nt; ++l)
    if (!grids[l]) { set_error("NULL grid"); return CRAVA_HIP_E_INVALID; }
  const size_t ntr_pad = (size_t)p->nxp * p->nyp;
  for (size_t t = 0; t < ntr_pad; ++t)
    if (!(p->rel_thick[t] > 0.0) || !std::isfinite(p->rel_thick[t])) {
      set_error("relative thicknesses must be positive and finite");
      return CRAVA_HIP_E_INVALID;
    }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(p->nxp, p->nyp, p->nzp, 0, &f);
  if (rc) return rc;
  // per-angle tables [nt][cn]: W {re, im}, t3, tau2, wg2
  const int cn = p->nzp / 2 + 1;
  std::vector<float> tw((size_t)nt * cn * 2), t3((size_t)nt * cn), tau2((size_t)nt * cn), wg2((size_t)nt * cn);
  crava::SwdArgs a{};
  for (int l = 0; l < nt; ++l) {
    const size_t o = (size_t)l * cn;
    crava::swd_angle_tables(p->nz, p->nzp, p->wavelets + (size_t)l * p->nzp, p->wavelet_norm[l], p->A + 3 * l,
                            p->prior_var0, p->corr_t, p->dz, p->lz, p->low_cut, p->high_cut, tw.data() + 2 * o,
                            t3.data() + o, tau2.data() + o, wg2.data() + o, &a.mod_w[l]);
    a.ev[l] = p->err_var[l];
    a.iwn2[l] = 1.0f / (p->wavelet_norm[l] * p->wavelet_norm[l]);
  }
  const auto al = [](size_t b) { return ((b + 255) / 256) * 256; };
  const size_t ntr = (size_t)p->nx * p->ny;
  size_t ntab = 0;
  for (int l = 0; l < nt; ++l) ntab += ((p->shift && p->shift[l]) ? 1 : 0) + ((p->gain && p->gain[l]) ? 1 : 0);
  const size_t b_w = al(tw.size() * sizeof(float)), b_t = al(t3.size() * sizeof(float));
  const size_t b_rel = al(ntr_pad * sizeof(double)), b_tab = al(ntr * sizeof(float));
  void *ws = nullptr;
  if ((rc = ctx->workspace(b_w + 3 * b_t + b_rel + ntab * b_tab, &ws))) return rc;
  char *wp = static_cast<char *>(ws);
  size_t off = 0;
  const auto up = [&](const void *src, size_t bytes, size_t slot) -> const void * {
    void *dst = wp + off;
    off += slot;
    return hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream) == hipSuccess ? dst : nullptr;
  };
  a.W = static_cast<const crava::cx<float> *>(up(tw.data(), tw.size() * sizeof(float), b_w));
  a.t3 = static_cast<const float *>(up(t3.data(), t3.size() * sizeof(float), b_t));
  a.tau2 = static_cast<const float *>(up(tau2.data(), tau2.size() * sizeof(float), b_t));
  a.wg2 = static_cast<const float *>(up(wg2.data(), wg2.size() * sizeof(float), b_t));
  a.rel = static_cast<const double *>(up(p->rel_thick, ntr_pad * sizeof(double), b_rel));
  bool ok = a.W && a.t3 && a.tau2 && a.wg2 && a.rel;
  for (int l = 0; l < nt; ++l) {
    a.shift[l] = (p->shift && p->shift[l]) ? static_cast<const float *>(up(p->shift[l], ntr * sizeof(float), b_tab))
                                           : nullptr;
    a.gain[l] = (p->gain && p->gain[l]) ? static_cast<const float *>(up(p->gain[l], ntr * sizeof(float), b_tab))
                                        : nullptr;
    ok = ok && (!(p->shift && p->shift[l]) || a.shift[l]) && (!(p->gain && p->gain[l]) || a.gain[l]);
    a.g[l] = static_cast<float *>(grids[l]);
  }
  if (!ok) { set_error("table upload failed"); return CRAVA_HIP_E_HIP; }
  // the host tables above go out of scope: finish the uploads first
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  a.dz = p->dz;
  a.wnc = p->wnc;
  a.low_cut = p->low_cut;
  a.high_cut = p->high_cut;
  a.sq = std::sqrt((float)p->nzp);
  a.inv_n = (float)(1.0 / (double)p->nzp);
  a.lz = p->lz;
  a.nt = nt;
  a.nx = p->nx;
  a.ny = p->ny;
  a.nz = p->nz;
  a.nxp = p->nxp;
  a.nyp = p->nyp;
  a.nzp = p->nzp;
  a.rnxp = 2 * (p->nxp / 2 + 1);
  if (crava::swd_fused_available(p->nzp) && f->stw[2]) return crava::launch_swd_fused(a, f->stw[2], ctx->stream);
  if ((rc = crava::launch_swd_prepare(a, ctx->stream))) return rc;
  if ((rc = f->pass(2, grids, nt, false, nullptr, ctx->stream))) return rc;
  if ((rc = crava::launch_swd_filter(a, ctx->stream))) return rc;
  if ((rc = f->pass(2, grids, nt, true, nullptr, ctx->stream))) return rc;
  return crava::launch_swd_finalize(a, ctx->stream);
}

int crava_hip_interpolate_seismic(crava_hip_ctx *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp,
                                  float energy_threshold, void *grid, int *n_flagged, int *n_zero) {
  if (!ctx || !grid) { set_error("NULL argument to crava_hip_interpolate_seismic"); return CRAVA_HIP_E_INVALID; }
  if (nx < 1 || ny < 1 || nz < 1 || nx > nxp || ny > nyp || nz > nzp) {
    set_error("invalid volume / padded grid sizes");
    return CRAVA_HIP_E_INVALID;
  }
  if ((long long)nx * ny > (1LL << 30)) { set_error("too many traces"); return CRAVA_HIP_E_INVALID; }
  const int rnxp = 2 * (nxp / 2 + 1);
  float *g = static_cast<float *>(grid);
  const size_t ntr = (size_t)nx * ny;
  void *ws = nullptr;
  int rc;
  if ((rc = ctx->workspace(ntr * sizeof(float), &ws))) return rc;
  if ((rc = crava::launch_isq_energy(g, nx, ny, nz, nyp, rnxp, static_cast<float *>(ws), ctx->stream))) return rc;
  std::vector<float> energy(ntr);
  CRAVA_HIP_CHECK(hipMemcpyAsync(energy.data(), ws, ntr * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  crava::IsqSchedule sch;
  crava::isq_schedule(energy.data(), nx, ny, energy_threshold, &sch);
  if (n_flagged) *n_flagged = sch.n_flagged;
  if (n_zero) *n_zero = sch.n_zero;
  if (sch.imin > sch.imax) {
    set_error("no trace has energy above the threshold");
    return CRAVA_HIP_E_INVALID;
  }
  if (!sch.ops.empty()) {
    if ((rc = ctx->workspace(sch.ops.size() * sizeof(crava::IsqOp), &ws))) return rc;
    const crava::IsqOp *dops = static_cast<const crava::IsqOp *>(ws);
    CRAVA_HIP_CHECK(hipMemcpyAsync(ws, sch.ops.data(), sch.ops.size() * sizeof(crava::IsqOp), hipMemcpyHostToDevice,
                                   ctx->stream));
    for (size_t v = 0; v + 1 < sch.level_start.size(); ++v) {
      const int b = sch.level_start[v], e = sch.level_start[v + 1];
      if ((rc = crava::launch_isq_interp(g, dops + b, e - b, nx, nyp, nzp, rnxp, ctx->stream))) return rc;
    }
  }
  if ((rc = crava::launch_isq_extrapolate(g, nx, ny, nz, nxp, nyp, nzp, rnxp, sch.imin, sch.imax, sch.jmin, sch.jmax,
                                          ctx->stream)))
    return rc;
  // the schedule's host copy goes out of scope
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return CRAVA_HIP_OK;
}

// ---------------------------------------------------------------------------
// Local-noise correction
int crava_hip_local_noise_g(int ntheta, const double *errThetaCov, const double *priorVar0, const double *postVar0,
                            double *G) {
  if (!errThetaCov || !priorVar0 || !postVar0 || !G) { set_error("NULL argument to crava_hip_local_noise_g"); return CRAVA_HIP_E_INVALID; }
  if (ntheta < 1 || ntheta > crava::kLncMaxTheta) { set_error("ntheta must be in 1..8"); return CRAVA_HIP_E_INVALID; }
  const int rc = crava::lnc_compute_g(ntheta, errThetaCov, priorVar0, postVar0, G);
  if (rc) set_error("crava_hip_local_noise_g failed");
  return rc;
}

int crava_hip_local_noise_correct(crava_hip_ctx *ctx, const crava_hip_local_noise *p, void *const *post,
                                  const void *const *mean2, double *traceCov) {
  if (!ctx || !p || !post || !mean2 || !p->G || !p->errThetaCov || !p->priorVar0 || !p->postVar0 || !p->noiseScale) {
    set_error("NULL argument to crava_hip_local_noise_correct");
    return CRAVA_HIP_E_INVALID;
  }
  if (p->ntheta < 1 || p->ntheta > crava::kLncMaxTheta) { set_error("ntheta must be in 1..8"); return CRAVA_HIP_E_INVALID; }
  if (p->nx < 1 || p->ny < 1 || p->nz < 1 || p->nx > p->nxp || p->ny > p->nyp || p->nz > p->nzp) {
    set_error("invalid volume / padded grid sizes");
    return CRAVA_HIP_E_INVALID;
  }
  if ((long long)p->nx * p->ny > (1LL << 28)) { set_error("too many traces"); return CRAVA_HIP_E_INVALID; }
  for (int c = 0; c < 3; ++c)
    if (!post[c] || !mean2[c]) { set_error("NULL grid"); return CRAVA_HIP_E_INVALID; }
  const int nt = p->ntheta;
  for (int a = 0; a < nt; ++a)
    if (!p->noiseScale[a]) { set_error("NULL noise-scale table"); return CRAVA_HIP_E_INVALID; }
  crava::LncConst c{};
  for (int r = 0; r < nt; ++r) {
    for (int q = 0; q < 3; ++q) c.g[r][q] = p->G[r * 3 + q];
    for (int q = 0; q < nt; ++q) c.e[r][q] = p->errThetaCov[r * nt + q];
  }
  double sd[3][3], smd[3][3];
  for (int r = 0; r < 3; ++r)
    for (int q = 0; q < 3; ++q) {
      c.sm[r][q] = p->priorVar0[r * 3 + q];
      smd[r][q] = p->postVar0[r * 3 + q];
      sd[r][q] = c.sm[r][q] - smd[r][q];
    }
  crava::lnc_inv_sqrt_host(sd, 0.0000001, c.sdi);  // src/crava.cpp:2294-2307
  crava::lnc_inv_sqrt_host(smd, 0.0, c.old);       // :2376-2385
  const size_t ntr = (size_t)p->nx * p->ny;
  // Grid2D::FindMin(RMISSING), libs/nrlib/grid/grid2d.hpp:182-191: the
  // smallest value that is not RMISSING, or RMISSING when every value is.
  // Written branch-free (a select and a min per value) so the host compiler
  // vectorises it.
  const double missing = -99999.000;
  for (int a = 0; a < nt; ++a) {
    const double *tab = p->noiseScale[a];
    double mn = HUGE_VAL;
    int found = 0;
    for (size_t t = 0; t < ntr; ++t) {
      const bool ok = tab[t] != missing;
      const double v = ok ? tab[t] : HUGE_VAL;
      mn = v < mn ? v : mn;
      found |= ok;
    }
    c.min_scale[a] = found ? mn : missing;
  }
  void *ws = nullptr;
  int rc;
  const size_t b_scale = sizeof(double) * nt * ntr, b_sn = sizeof(double) * 9 * ntr;
  if ((rc = ctx->workspace(b_scale + b_sn, &ws))) return rc;
  double *dscale = static_cast<double *>(ws);
  double *dsn = reinterpret_cast<double *>(static_cast<char *>(ws) + b_scale);
  for (int a = 0; a < nt; ++a)
    CRAVA_HIP_CHECK(hipMemcpyAsync(dscale + (size_t)a * ntr, p->noiseScale[a], sizeof(double) * ntr,
                                   hipMemcpyHostToDevice, ctx->stream));
  if ((rc = crava::launch_lnc_trace(c, nt, dscale, p->nx, p->ny, dsn, traceCov, ctx->stream))) return rc;
  const int rnxp = 2 * (p->nxp / 2 + 1);
  float *pg[3];
  const float *mg[3];
  for (int q = 0; q < 3; ++q) {
    pg[q] = static_cast<float *>(post[q]);
    mg[q] = static_cast<const float *>(mean2[q]);
  }
  if ((rc = crava::launch_lnc_apply(pg, mg, dsn, p->nx, p->ny, p->nz, p->nyp, rnxp, ctx->stream))) return rc;
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return CRAVA_HIP_OK;
}

int crava_hip_local_noise_realisation(crava_hip_ctx *ctx, int nx, int ny, int nz, int nxp, int nyp, int nzp,
                                      const double *traceCov, void *const *seed) {
  if (!ctx || !traceCov || !seed || !seed[0] || !seed[1] || !seed[2]) {
    set_error("NULL argument to crava_hip_local_noise_realisation");
    return CRAVA_HIP_E_INVALID;
  }
  if (nx < 1 || ny < 1 || nz < 1 || nx > nxp || ny > nyp || nz > nzp) {
    set_error("invalid volume / padded grid sizes");
    return CRAVA_HIP_E_INVALID;
  }
  if ((long long)nx * ny > (1LL << 28)) { set_error("too many traces"); return CRAVA_HIP_E_INVALID; }
  float *g[3];
  for (int q = 0; q < 3; ++q) g[q] = static_cast<float *>(seed[q]);
  return crava::launch_lnc_realisation(g, traceCov, nx, ny, nz, nyp, 2 * (nxp / 2 + 1), ctx->stream);
}

}  // extern "C"

struct crava_hip_batch {
  crava_hip_ctx *ctx;
  crava::BatchExec exec;
  crava_hip_batch(crava_hip_ctx *c, int precision) : ctx(c), exec(c, precision) {}
};

int crava_hip_batch_create(crava_hip_ctx *ctx, int precision, int nvol, const crava_hip_batch_volume *vols,
                           crava_hip_batch **batch) {
  if (!ctx || !vols || !batch || nvol < 1) { set_error("invalid arguments to crava_hip_batch_create"); return CRAVA_HIP_E_INVALID; }
  *batch = nullptr;
  std::vector<crava::BatchVolume> bv((size_t)nvol);
  for (int v = 0; v < nvol; ++v) {
    if (!vols[v].post) { set_error("volume " + std::to_string(v) + " has no problem"); return CRAVA_HIP_E_INVALID; }
    bv[v].pd = &vols[v].post->dev;
    for (int c = 0; c < 3; ++c) bv[v].mu[c] = vols[v].mu[c];
    for (int l = 0; l < crava::kMaxTheta; ++l) bv[v].data[l] = vols[v].data[l];
    for (int q = 0; q < 6; ++q) bv[v].cov[q] = vols[v].cov[q];
    bv[v].errcorr = vols[v].errcorr;
    bv[v].corr = vols[v].corr;
  }
  crava_hip_batch *b = new (std::nothrow) crava_hip_batch(ctx, precision);
  if (!b) { set_error("out of host memory"); return CRAVA_HIP_E_OOM; }
  const int rc = b->exec.init(bv.data(), nvol);
  if (rc) {
    delete b;
    return rc;
  }
  *batch = b;
  return CRAVA_HIP_OK;
}

int crava_hip_batch_autotune(crava_hip_batch *batch, int *points_per_block, int *blocks, int *form,
                             float *post_ms) {
  if (!batch) { set_error("batch is NULL"); return CRAVA_HIP_E_INVALID; }
  return batch->exec.autotune(batch->ctx->stream, points_per_block, blocks, form, post_ms);
}

int crava_hip_batch_set_launch(crava_hip_batch *batch, int points_per_block, int blocks, int form) {
  if (!batch) { set_error("batch is NULL"); return CRAVA_HIP_E_INVALID; }
  return batch->exec.set_launch(points_per_block, blocks, form);
}

int crava_hip_batch_run(crava_hip_batch *batch, int use_graph) {
  if (!batch) { set_error("batch is NULL"); return CRAVA_HIP_E_INVALID; }
  return batch->exec.run(batch->ctx->stream, use_graph != 0);
}

int crava_hip_batch_phase_times(crava_hip_batch *batch, float *ms) {
  if (!batch || !ms) { set_error("NULL argument to crava_hip_batch_phase_times"); return CRAVA_HIP_E_INVALID; }
  return batch->exec.phase_times(batch->ctx->stream, ms);
}

int crava_hip_batch_counts(crava_hip_batch *batch, long long *n_fail, long long *n_fp64) {
  if (!batch) { set_error("batch is NULL"); return CRAVA_HIP_E_INVALID; }
  return batch->exec.counts(batch->ctx->stream, n_fail, n_fp64);
}

int crava_hip_batch_destroy(crava_hip_batch *batch) {
  delete batch;
  return CRAVA_HIP_OK;
}

// ---------------------------------------------------------------------------
// Grid fills, element-wise operators, data variance, facies geomodel.
static bool grid_dims_ok(int nxp, int nyp, int nzp) { return nxp >= 1 && nyp >= 1 && nzp >= 1; }

int crava_hip_fill_param_corr(crava_hip_ctx *ctx, void *const *grids, int ngrid, const float *scale, int nxp, int nyp,
                              int nzp, const float *corr_xy, const float *circ_corr_t, float grad_i, float grad_j) {
  if (!ctx || !grids || !scale || !corr_xy || !circ_corr_t || ngrid < 1 || ngrid > crava::kMaxGrids ||
      !grid_dims_ok(nxp, nyp, nzp)) {
    set_error("invalid arguments to crava_hip_fill_param_corr");
    return CRAVA_HIP_E_INVALID;
  }
  crava::GridSet gs{};
  gs.n = ngrid;
  for (int g = 0; g < ngrid; ++g) {
    if (!grids[g]) { set_error("NULL grid"); return CRAVA_HIP_E_INVALID; }
    gs.dst[g] = grids[g];
    gs.src[g] = grids[g];
    gs.scale[g] = (double)scale[g];
  }
  return crava::launch_fill_param_corr(gs, nxp, nyp, nzp, corr_xy, circ_corr_t, grad_i, grad_j, ctx->stream);
}

int crava_hip_fill_err_corr(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, const float *corr_xy,
                            float grad_i, float grad_j) {
  if (!ctx || !grid || !corr_xy || !grid_dims_ok(nxp, nyp, nzp)) {
    set_error("invalid arguments to crava_hip_fill_err_corr");
    return CRAVA_HIP_E_INVALID;
  }
  return crava::launch_fill_err_corr(static_cast<float *>(grid), nxp, nyp, nzp, corr_xy, grad_i, grad_j, ctx->stream);
}

int crava_hip_err_corr_plane(crava_hip_ctx *ctx, void *plane, int nxp, int nyp, int nzp, const float *corr_xy,
                             int precision) {
  if (!ctx || !plane || !corr_xy || !grid_dims_ok(nxp, nyp, nzp) ||
      (precision != CRAVA_HIP_F32 && precision != CRAVA_HIP_MIXED && precision != CRAVA_HIP_F64)) {
    set_error("invalid arguments to crava_hip_err_corr_plane");
    return CRAVA_HIP_E_INVALID;
  }
  int rc = crava::launch_fill_err_corr_plane(plane, precision == CRAVA_HIP_F64, nxp, nyp, nzp, corr_xy, ctx->stream);
  if (rc) return rc;
  // the unscaled 2-D transform: a one-plane grid's z axis is the identity
  return fft_common(ctx, &plane, 1, nxp, nyp, 1, CRAVA_HIP_COVARIANCE, precision, false);
}

int crava_hip_corr_spectrum_separable(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp,
                                      const float *corr_xy, const float *circ_t, int precision) {
  if (!ctx || !grid || !corr_xy || !circ_t || !grid_dims_ok(nxp, nyp, nzp) ||
      (precision != CRAVA_HIP_F32 && precision != CRAVA_HIP_MIXED && precision != CRAVA_HIP_F64)) {
    set_error("invalid arguments to crava_hip_corr_spectrum_separable");
    return CRAVA_HIP_E_INVALID;
  }
  const bool f64 = precision == CRAVA_HIP_F64;
  // plane 0 of the grid: corrXY, then its unscaled 2-D transform in place
  int rc = crava::launch_fill_err_corr_plane(grid, f64, nxp, nyp, nzp, corr_xy, ctx->stream);
  if (rc) return rc;
  if ((rc = fft_common(ctx, &grid, 1, nxp, nyp, 1, CRAVA_HIP_COVARIANCE, precision, false))) return rc;
  return crava::launch_corr_outer(grid, f64, nxp, nyp, nzp, circ_t, ctx->stream);
}

int crava_hip_post_set_errcorr_plane(crava_hip_post *post, int plane) {
  if (!post) {
    set_error("NULL posterior");
    return CRAVA_HIP_E_INVALID;
  }
  post->dev.ec_plane = plane != 0;
  return CRAVA_HIP_OK;
}

int crava_hip_grid_op(crava_hip_ctx *ctx, int op, void *grid, const void *other, int nxp, int nyp, int nzp,
                      int transformed, float scalar) {
  if (!ctx || !grid || !grid_dims_ok(nxp, nyp, nzp) || op < CRAVA_HIP_OP_ADD || op > CRAVA_HIP_OP_COPY) {
    set_error("invalid arguments to crava_hip_grid_op");
    return CRAVA_HIP_E_INVALID;
  }
  const bool binary = op == CRAVA_HIP_OP_ADD || op == CRAVA_HIP_OP_SUBTRACT || op == CRAVA_HIP_OP_MULTIPLY ||
                      op == CRAVA_HIP_OP_COPY;
  if (binary && !other) { set_error("this operator needs a second grid"); return CRAVA_HIP_E_INVALID; }
  // the reference asserts the domain of these operators (src/fftgrid.cpp:1903,1976,1984)
  if (transformed && (op == CRAVA_HIP_OP_MULTIPLY_SCALAR || op == CRAVA_HIP_OP_ADD_SCALAR ||
                      op == CRAVA_HIP_OP_FILL_CONSTANT)) {
    set_error("scalar operators apply in the real domain only");
    return CRAVA_HIP_E_STATE;
  }
  if (!transformed && op == CRAVA_HIP_OP_CONJUGATE) {
    set_error("conjugate applies to transformed grids only");
    return CRAVA_HIP_E_STATE;
  }
  const long long csize = (long long)(nxp / 2 + 1) * nyp * nzp;
  const long long n = transformed ? csize : 2 * csize;
  return crava::launch_grid_op(op, static_cast<float *>(grid), static_cast<const float *>(other), n, transformed != 0,
                               scalar, ctx->stream);
}

int crava_hip_data_variance(crava_hip_ctx *ctx, const void *grid, int nx, int ny, int nz, int nxp, int nyp, int nzp,
                            float *variance, long long *ndata) {
  if (!ctx || !grid || !variance || !grid_dims_ok(nxp, nyp, nzp) || nx < 0 || ny < 0 || nz < 0 || nx > nxp ||
      ny > nyp || nz > nzp) {
    set_error("invalid arguments to crava_hip_data_variance");
    return CRAVA_HIP_E_INVALID;
  }
  void *work = nullptr;
  int rc = ctx->workspace(sizeof(double) * crava::data_variance_work_doubles(nyp, nzp), &work);
  if (rc) return rc;
  double out[2] = {0.0, 0.0};
  if ((rc = crava::data_variance(static_cast<const float *>(grid), nx, ny, nz, nxp, nyp, nzp,
                                 static_cast<double *>(work), out, ctx->stream)))
    return rc;
  const long long cnt = (long long)out[1];
  *variance = cnt == 0 ? 0.0f : static_cast<float>(out[0] / static_cast<double>(cnt));
  if (ndata) *ndata = cnt;
  return CRAVA_HIP_OK;
}

int crava_hip_facies_geomodel(crava_hip_ctx *ctx, int nx, int ny, int nz, int nfac, const float *prior,
                              float *const *prob) {
  if (!ctx || !prior || !prob || nfac < 1 || nfac > crava::kMaxGrids - 1 || nx < 1 || ny < 1 || nz < 1) {
    set_error("invalid arguments to crava_hip_facies_geomodel");
    return CRAVA_HIP_E_INVALID;
  }
  crava::GridSet gs{};
  gs.n = nfac;
  for (int l = 0; l <= nfac; ++l)
    if (!prob[l]) { set_error("NULL probability grid"); return CRAVA_HIP_E_INVALID; }
  for (int l = 0; l < nfac; ++l) {
    gs.dst[l] = prob[l];
    gs.src[l] = prob[l];
    gs.scale[l] = (double)prior[l];
  }
  return crava::launch_facies_geomodel(gs, prob[nfac], nfac, nx, ny, nz, ctx->stream);
}

// ---------------------------------------------------------------------------
// Device grid storage for host code that holds no HIP types.
int crava_hip_grid_alloc(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, int precision, void **grid) {
  if (!ctx || !grid || nxp < 1 || nyp < 1 || nzp < 1 || precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F64) {
    set_error("invalid arguments to crava_hip_grid_alloc");
    return CRAVA_HIP_E_INVALID;
  }
  *grid = nullptr;
  const size_t bytes = (size_t)2 * (nxp / 2 + 1) * nyp * nzp * (precision == CRAVA_HIP_F64 ? 8 : 4);
  const hipError_t e = hipMalloc(grid, bytes);
  if (e != hipSuccess) {
    *grid = nullptr;
    set_error(std::string("hipMalloc: ") + hipGetErrorString(e));
    return e == hipErrorOutOfMemory ? CRAVA_HIP_E_OOM : CRAVA_HIP_E_HIP;
  }
  return CRAVA_HIP_OK;
}

int crava_hip_grid_free(void *grid) {
  if (grid) CRAVA_HIP_CHECK(hipFree(grid));
  return CRAVA_HIP_OK;
}

int crava_hip_grid_upload(crava_hip_ctx *ctx, void *grid, const void *host, size_t bytes) {
  if (!ctx || !grid || !host) { set_error("invalid arguments to crava_hip_grid_upload"); return CRAVA_HIP_E_INVALID; }
  CRAVA_HIP_CHECK(hipMemcpyAsync(grid, host, bytes, hipMemcpyHostToDevice, ctx->stream));
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return CRAVA_HIP_OK;
}

int crava_hip_grid_read_trace(crava_hip_ctx *ctx, const void *grid, int nxp, int nyp, int nzp, int i, int j, int k0,
                              int nk, float *out) {
  if (!ctx || !grid || !out || nxp < 1 || nyp < 1 || nzp < 1) {
    set_error("invalid arguments to crava_hip_grid_read_trace");
    return CRAVA_HIP_E_INVALID;
  }
  const int rnxp = 2 * (nxp / 2 + 1);
  if (i < 0 || i >= rnxp || j < 0 || j >= nyp || k0 < 0 || nk < 0 || k0 + nk > nzp) {
    set_error("trace outside the grid in crava_hip_grid_read_trace");
    return CRAVA_HIP_E_INVALID;
  }
  if (nk == 0) return CRAVA_HIP_OK;
  const size_t plane = (size_t)rnxp * nyp * sizeof(float);
  const char *src = static_cast<const char *>(grid) + (size_t)k0 * plane + ((size_t)j * rnxp + i) * sizeof(float);
  CRAVA_HIP_CHECK(hipMemcpy2DAsync(out, sizeof(float), src, plane, sizeof(float), (size_t)nk, hipMemcpyDeviceToHost,
                                   ctx->stream));
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return CRAVA_HIP_OK;
}

int crava_hip_grid_download(crava_hip_ctx *ctx, void *host, const void *grid, size_t bytes) {
  if (!ctx || !grid || !host) { set_error("invalid arguments to crava_hip_grid_download"); return CRAVA_HIP_E_INVALID; }
  CRAVA_HIP_CHECK(hipMemcpyAsync(host, grid, bytes, hipMemcpyDeviceToHost, ctx->stream));
  CRAVA_HIP_CHECK(hipStreamSynchronize(ctx->stream));
  return CRAVA_HIP_OK;
}
