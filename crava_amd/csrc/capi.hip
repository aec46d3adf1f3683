// C ABI of libcrava_hip.so (declared in include/crava_hip.h).
#include "../../include/crava_hip.h"
#include "common.hpp"
#include "context.hpp"
#include "fft.hpp"
#include "facies.hpp"
#include "posterior.hpp"
#include "simulate.hpp"
#include "tables.hpp"

#include <cmath>
#include <cstring>
#include <map>
#include <memory>
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

Context::~Context() {
  plans.clear();
  if (counters) (void)hipFree(counters);
}
}  // namespace crava

struct crava_hip_post {
  crava_hip_ctx *ctx;
  crava::PostDevice dev;
};

using crava::set_error;

// transform arithmetic/storage code of a C-ABI precision
static int fft_prec(int precision) {
  return precision == CRAVA_HIP_F64 ? 2 : (precision == CRAVA_HIP_MIXED ? 1 : 0);
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
  if (e != hipSuccess) {
    set_error(std::string("hipMalloc(counters): ") + hipGetErrorString(e));
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

static int fft_common(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, int cubetype, int precision,
                      bool inverse) {
  if (!ctx || !grid) { set_error("NULL ctx or grid"); return CRAVA_HIP_E_INVALID; }
  if (nxp < 2 || nyp < 1 || nzp < 1) { set_error("invalid grid size"); return CRAVA_HIP_E_INVALID; }
  if (cubetype < CRAVA_HIP_DATA || cubetype > CRAVA_HIP_COVARIANCE) { set_error("invalid cubetype"); return CRAVA_HIP_E_INVALID; }
  if (precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F64) {
    set_error("invalid transform precision");
    return CRAVA_HIP_E_INVALID;
  }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(nxp, nyp, nzp, fft_prec(precision), &f);
  if (rc) return rc;
  const long long N = (long long)nxp * nyp * nzp;
  double scale;
  if (!inverse) {
    // FFTGrid::fftInPlace multiplies by 1.0f/sqrt(float(N)) before the transform
    scale = cubetype == CRAVA_HIP_COVARIANCE ? 1.0
            : (precision == CRAVA_HIP_F64 ? 1.0 / std::sqrt((double)N) : (double)(1.0f / std::sqrt((float)N)));
    return f->r2c(grid, ctx->stream, scale);
  }
  if (cubetype == CRAVA_HIP_COVARIANCE)
    scale = precision == CRAVA_HIP_F64 ? 1.0 / (double)N : (double)(float)(1.0 / (double)N);
  else
    scale = precision == CRAVA_HIP_F64 ? 1.0 / std::sqrt((double)N) : (double)(float)(1.0 / std::sqrt((float)N));
  return f->c2r(grid, ctx->stream, scale);
}

int crava_hip_fft_r2c(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, int cubetype, int precision) {
  return fft_common(ctx, grid, nxp, nyp, nzp, cubetype, precision, false);
}

int crava_hip_fft_c2r(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, int cubetype, int precision) {
  return fft_common(ctx, grid, nxp, nyp, nzp, cubetype, precision, true);
}

int crava_hip_fft_pass(crava_hip_ctx *ctx, void *grid, int nxp, int nyp, int nzp, int axis, int inverse,
                       int precision, double scale) {
  if (!ctx || !grid || axis < 0 || axis > 2) { set_error("invalid arguments to crava_hip_fft_pass"); return CRAVA_HIP_E_INVALID; }
  if (precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F64) { set_error("invalid precision"); return CRAVA_HIP_E_INVALID; }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(nxp, nyp, nzp, fft_prec(precision), &f);
  if (rc) return rc;
  if (axis == 0) return f->pass_x(grid, ctx->stream, inverse != 0, scale);
  if (axis == 1) return f->pass_y(grid, ctx->stream, inverse != 0);
  return f->pass_z(grid, ctx->stream, inverse != 0);
}

int crava_hip_autotune_fft(crava_hip_ctx *ctx, void *scratch, int nxp, int nyp, int nzp, int precision,
                           int *variants, float *ms) {
  if (!ctx || !scratch) { set_error("NULL argument to crava_hip_autotune_fft"); return CRAVA_HIP_E_INVALID; }
  if (precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F64) { set_error("invalid precision"); return CRAVA_HIP_E_INVALID; }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(nxp, nyp, nzp, fft_prec(precision), &f);
  if (rc) return rc;
  float t[3] = {0, 0, 0};
  if ((rc = f->autotune(scratch, ctx->stream, 5, t))) return rc;
  for (int a = 0; a < 3; ++a) {
    if (variants) variants[a] = f->var[a];
    if (ms) ms[a] = t[a];
  }
  return CRAVA_HIP_OK;
}

int crava_hip_fft_set_variants(crava_hip_ctx *ctx, int nxp, int nyp, int nzp, int precision, const int *variants) {
  if (!ctx || !variants) { set_error("NULL argument to crava_hip_fft_set_variants"); return CRAVA_HIP_E_INVALID; }
  if (precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F64) { set_error("invalid precision"); return CRAVA_HIP_E_INVALID; }
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
  hipError_t e = hipMalloc(&h->dev.tabf, sizeof(crava::PerK<float>) * p.nzp);
  if (e == hipSuccess) e = hipMalloc(&h->dev.tabd, sizeof(crava::PerK<double>) * p.nzp);
  if (e == hipSuccess)
    e = hipMemcpy(h->dev.tabf, tf.data(), sizeof(crava::PerK<float>) * p.nzp, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(h->dev.tabd, td.data(), sizeof(crava::PerK<double>) * p.nzp, hipMemcpyHostToDevice);
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
  delete post;
  return CRAVA_HIP_OK;
}

int crava_hip_post_run(crava_hip_post *post, int precision, void *const *mu, void *const *data, void *const *cov,
                       const void *errcorr, long long *n_fail, long long *n_fp64) {
  if (!post || !mu || !data || !cov || !errcorr) { set_error("NULL argument to crava_hip_post_run"); return CRAVA_HIP_E_INVALID; }
  if (precision < CRAVA_HIP_F32 || precision > CRAVA_HIP_F64) { set_error("invalid precision"); return CRAVA_HIP_E_INVALID; }
  for (int c = 0; c < 3; ++c) if (!mu[c]) { set_error("NULL mean grid"); return CRAVA_HIP_E_INVALID; }
  for (int q = 0; q < 6; ++q) if (!cov[q]) { set_error("NULL covariance grid"); return CRAVA_HIP_E_INVALID; }
  for (int l = 0; l < post->dev.pp.ntheta; ++l) if (!data[l]) { set_error("NULL data grid"); return CRAVA_HIP_E_INVALID; }
  crava_hip_ctx *ctx = post->ctx;
  const bool want = n_fail || n_fp64;
  if (want) CRAVA_HIP_CHECK(hipMemsetAsync(ctx->counters, 0, 2 * sizeof(unsigned long long), ctx->stream));
  int rc = crava::launch_posterior(post->dev, precision, mu, data, cov, errcorr, want ? ctx->counters : nullptr,
                                   ctx->stream);
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
  for (int c = 0; c < 3; ++c) {
    if ((rc = crava_hip_fft_c2r(ctx, seed[c], nxp, nyp, nzp, CRAVA_HIP_PARAMETER, CRAVA_HIP_F32))) return rc;
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

}  // extern "C"
