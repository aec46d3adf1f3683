// C ABI of libcrava_hip.so (declared in include/crava_hip.h).
#include "../../include/crava_hip.h"
#include "common.hpp"
#include "context.hpp"
#include "fft.hpp"
#include "posterior.hpp"
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

int Context::plan(int nxp, int nyp, int nzp, bool dbl, Fft3D **out) {
  auto key = std::make_tuple(nxp, nyp, nzp, dbl);
  auto it = plans.find(key);
  if (it == plans.end()) {
    std::unique_ptr<Fft3D> f(new (std::nothrow) Fft3D(nxp, nyp, nzp, dbl));
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
  if (precision != CRAVA_HIP_F32 && precision != CRAVA_HIP_F64) {
    set_error("transform precision must be CRAVA_HIP_F32 or CRAVA_HIP_F64");
    return CRAVA_HIP_E_INVALID;
  }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(nxp, nyp, nzp, precision == CRAVA_HIP_F64, &f);
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
  if (precision != CRAVA_HIP_F32 && precision != CRAVA_HIP_F64) { set_error("invalid precision"); return CRAVA_HIP_E_INVALID; }
  crava::Fft3D *f = nullptr;
  int rc = ctx->plan(nxp, nyp, nzp, precision == CRAVA_HIP_F64, &f);
  if (rc) return rc;
  if (axis == 0) return f->pass_x(grid, ctx->stream, inverse != 0, scale);
  if (axis == 1) return f->pass_y(grid, ctx->stream, inverse != 0);
  return f->pass_z(grid, ctx->stream, inverse != 0);
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

}  // extern "C"
