// Launch-shape autotuner for the fused Model C step.
//
// The step's results are bitwise independent of its launch shape (kernel choice,
// strips per block, rows loaded ahead), so the shape is a pure performance choice
// that depends on the block shape, the number of blocks and whether noise is
// drawn.  lgd_autotune_modelc_step times every candidate on the caller's state
// with HIP events on the context's stream and keeps the fastest for that key;
// lgd_modelc_step and the ensemble step then use it whenever no knob is set
// explicitly (lgd_ctx_set_tuning).
#include <algorithm>
#include <vector>

#include "lgd_internal.h"

using namespace lgd;

lgd_launch_shape lgd_modelc_shape(const lgd_ctx *ctx, int64_t nx, int64_t ny, int64_t nblocks,
                                  bool noise, const lgd_launch_shape &dflt) {
  lgd_launch_shape s = dflt;
  const bool explicit_knobs =
      ctx->modelc_kernel != 0 || ctx->modelc_strips != 0 || ctx->modelc_prefetch != 0;
  if (!explicit_knobs) {
    auto it = ctx->tuned.find(std::make_tuple(nx, ny, nblocks, noise ? 1 : 0));
    if (it != ctx->tuned.end()) return it->second;
    return s;
  }
  if (ctx->modelc_kernel != 0) s.kernel = (int32_t)ctx->modelc_kernel;
  if (ctx->modelc_strips > 0) s.strips = (int32_t)ctx->modelc_strips;
  if (ctx->modelc_prefetch > 0) s.prefetch = (int32_t)ctx->modelc_prefetch;
  return s;
}

namespace {

// Mean duration of one lgd_modelc_step launch in shape `s` (ms).
int time_shape(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
               const lgd_buf *phi, const lgd_buf *u, lgd_buf *pn, lgd_buf *un, double dt,
               int32_t wiener, int reps, const lgd_launch_shape &s, hipEvent_t e0, hipEvent_t e1,
               double *ms) {
  const int64_t k0 = ctx->modelc_kernel, s0 = ctx->modelc_strips, p0 = ctx->modelc_prefetch;
  ctx->modelc_kernel = s.kernel;
  ctx->modelc_strips = s.strips;
  ctx->modelc_prefetch = s.prefetch;
  int rc = LGD_OK;
  for (int w = 0; w < 2 && !rc; ++w)
    rc = lgd_modelc_step(ctx, grid, params, phi, u, pn, un, dt, 0, 0, wiener);
  hipError_t e = hipSuccess;
  if (!rc) e = hipEventRecord(e0, ctx->stream);
  for (int r = 0; r < reps && !rc && e == hipSuccess; ++r)
    rc = lgd_modelc_step(ctx, grid, params, phi, u, pn, un, dt, 0, 0, wiener);
  if (!rc && e == hipSuccess) e = hipEventRecord(e1, ctx->stream);
  if (!rc && e == hipSuccess) e = hipEventSynchronize(e1);
  float t = 0.0f;
  if (!rc && e == hipSuccess) e = hipEventElapsedTime(&t, e0, e1);
  ctx->modelc_kernel = k0;
  ctx->modelc_strips = s0;
  ctx->modelc_prefetch = p0;
  if (rc) return rc;
  if (e != hipSuccess) return hip_fail(ctx, e, "autotune timing");
  *ms = (double)t / reps;
  return LGD_OK;
}

}  // namespace

extern "C" {

int lgd_autotune_modelc_step(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                             const lgd_buf *phi, const lgd_buf *u, double dt, int32_t wiener,
                             int32_t reps, lgd_tuning_result *out) {
  if (!ctx || !grid || !params || !out) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  const lgd_cart &c = grid->cart;
  if (c.dim != 2) return fail(ctx, LGD_ERR_UNSUPPORTED, "Model C is 2-D");
  int rc;
  if ((rc = check_field(ctx, grid, phi, 1, "phi"))) return rc;
  if ((rc = check_field(ctx, grid, u, 1, "u"))) return rc;
  if (reps < 1) reps = 5;
  const bool noise = wiener && params->noise_amplitude != 0.0;
  std::vector<lgd_launch_shape> cand;
  cand.push_back(lgd_launch_shape{1, 0, 0});
  const bool pairable = lgd_stream_pairable(c.n[0], c.nb[0]);
  if (c.n[0] + 2 <= 128)
    for (int32_t strips : {1, 2, 4, 8, 16, 32})
      if (strips <= c.n[1])
        for (int32_t pf : {1, 2}) {
          cand.push_back(lgd_launch_shape{2, strips, pf});
          if (pairable) cand.push_back(lgd_launch_shape{3, strips, pf});
        }
  lgd_buf *pn = nullptr, *un = nullptr;
  const size_t len = phi->len;
  if ((rc = lgd_buf_alloc(ctx, len, &pn))) return rc;
  if ((rc = lgd_buf_alloc(ctx, len, &un))) {
    lgd_buf_free(pn);
    return rc;
  }
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) rc = hip_fail(ctx, e, "autotune events");
  const lgd_launch_shape dflt = lgd_modelc_default_shape(noise);
  double best = 1e300, dflt_ms = -1.0;
  lgd_launch_shape pick = dflt;
  for (const lgd_launch_shape &s : cand) {
    if (rc) break;
    double ms = 0.0;
    rc = time_shape(ctx, grid, params, phi, u, pn, un, dt, wiener, reps, s, e0, e1, &ms);
    if (rc) break;
    if (s.kernel == dflt.kernel && s.strips == dflt.strips && s.prefetch == dflt.prefetch)
      dflt_ms = ms;
    if (ms < best) {
      best = ms;
      pick = s;
    }
  }
  if (e0) (void)hipEventDestroy(e0);
  if (e1) (void)hipEventDestroy(e1);
  lgd_buf_free(pn);
  lgd_buf_free(un);
  if (rc) return rc;
  ctx->tuned[std::make_tuple(c.n[0], c.n[1], c.nblocks, noise ? 1 : 0)] = pick;
  ++ctx->tuning_epoch;
  out->kernel = pick.kernel;
  out->strips = pick.strips;
  out->prefetch = pick.prefetch;
  out->ms = best;
  out->default_ms = dflt_ms;
  out->candidates = (int32_t)cand.size();
  return LGD_OK;
}

}  // extern "C"
