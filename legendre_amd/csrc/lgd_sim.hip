// Host-side orchestration: the GpuScheduler-driven Simulation.
//
// Mirrors Simulation::step (src/core/simulation.rs:255-290) with
// EulerMaruyama::step (src/integrators/euler_maruyama.rs:34-75) or
// ForwardEuler::step (src/integrators/euler.rs:104-130): the host owns the
// control flow and the f64 time accumulation (t += dt, the noise salt), the
// device owns every per-block sweep.  Initial conditions are host work, as in
// the reference (ModelC::initialize_grains, src/physics/phasefield.rs:227-314).
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "lgd_regrid.h"

struct lgd_sim {
  lgd_ctx *ctx = nullptr;
  const lgd_grid *grid = nullptr;
  enum Kind { kModelC = 0, kHeat = 1 } kind = kModelC;
  int32_t mode = LGD_MODE_FUSED;
  lgd_modelc_params params{};
  int32_t wiener = 0;
  uint64_t seed = 0;
  double kappa = 0.0;
  bool has_bcs = false;
  lgd_face_bc bcs[6]{};
  std::vector<lgd_buf *> state;  // current state, one buffer per field
  std::vector<lgd_buf *> next;   // fused mode: ping-pong partner
  std::vector<lgd_buf *> k;      // per-intent mode: Tendency(Time) buffer
  lgd_buf *amp = nullptr;        // per-intent mode: Tendency(Wiener(0)) buffer (phi only)
  // AMR hierarchy (owned) — set for simulations on an AmrGrid.
  lgd_amr *amr = nullptr;
  int32_t subcycle = 0;
  bool adaptive = false;
  lgd_regrid_policy policy{};
  std::vector<lgd_buf *> snaps;  // subcycling: per (level, field) time-interpolation snapshot
  double t = 0.0;
  int64_t step_index = 0;
};

using namespace lgd;

namespace {

void free_stages(lgd_sim *s) {
  for (auto *b : s->next) lgd_buf_free(b);
  for (auto *b : s->k) lgd_buf_free(b);
  for (auto *b : s->snaps) lgd_buf_free(b);
  lgd_buf_free(s->amp);
  s->next.clear();
  s->k.clear();
  s->snaps.clear();
  s->amp = nullptr;
}

void free_all(lgd_sim *s) {
  for (auto *b : s->state) lgd_buf_free(b);
  s->state.clear();
  free_stages(s);
}

int alloc_fields(lgd_sim *s, int nfields) {
  size_t len = (size_t)(s->grid->cart.nblocks * slab_len(s->grid->cart, 1));
  for (int f = 0; f < nfields; ++f) {
    lgd_buf *b = nullptr;
    int rc = lgd_buf_alloc(s->ctx, len, &b);
    if (rc) return rc;
    s->state.push_back(b);
    if (s->mode == LGD_MODE_FUSED) {
      if ((rc = lgd_buf_alloc(s->ctx, len, &b))) return rc;
      s->next.push_back(b);
    } else {
      if ((rc = lgd_buf_alloc(s->ctx, len, &b))) return rc;
      s->k.push_back(b);
    }
  }
  if (s->mode == LGD_MODE_PER_INTENT && s->kind == lgd_sim::kModelC && s->wiener) {
    int rc = lgd_buf_alloc(s->ctx, len, &s->amp);
    if (rc) return rc;
  }
  return LGD_OK;
}

bool all_mirror(const lgd_sim *s) {
  if (!s->has_bcs) return true;
  for (int d = 0; d < s->grid->cart.dim; ++d)
    for (int e = 0; e < 2; ++e)
      if (s->bcs[d * 2 + e].kind != LGD_BC_MIRROR) return false;
  return true;
}

int step_modelc(lgd_sim *s, double dt, uint64_t salt) {
  lgd_ctx *ctx = s->ctx;
  int rc;
  if (s->mode == LGD_MODE_FUSED) {
    rc = lgd_modelc_step(ctx, s->grid, &s->params, s->state[0], s->state[1], s->next[0],
                         s->next[1], dt, s->seed, salt, s->wiener);
    if (rc) return rc;
    std::swap(s->state, s->next);
    return LGD_OK;
  }
  // eval_drift: ModelC::fill_ghosts then the per-block Time vector field.
  if ((rc = lgd_fill_ghosts(ctx, s->grid, s->state[0], 1, nullptr))) return rc;
  if ((rc = lgd_fill_ghosts(ctx, s->grid, s->state[1], 1, nullptr))) return rc;
  if ((rc = lgd_modelc_rhs(ctx, s->grid, &s->params, s->state[0], s->state[1], s->k[0], s->k[1])))
    return rc;
  if (s->wiener && (rc = lgd_modelc_amplitude(ctx, s->grid, &s->params, s->amp))) return rc;
  // apply_step_with: Time on phi and u, then Wiener(0) on phi.
  if ((rc = lgd_axpy(ctx, s->state[0], s->k[0], dt))) return rc;
  if ((rc = lgd_axpy(ctx, s->state[1], s->k[1], dt))) return rc;
  if (s->wiener) {
    const uint64_t stream = (1ull << 32) | 0ull;
    if ((rc = lgd_apply_wiener(ctx, s->grid, s->state[0], s->amp, 1, dt, s->seed, salt, stream)))
      return rc;
  }
  return LGD_OK;
}

int step_heat(lgd_sim *s, double dt) {
  int rc;
  if (s->mode == LGD_MODE_FUSED) {
    if ((rc = lgd_heat_step(s->ctx, s->grid, s->kappa, s->state[0], s->next[0], dt))) return rc;
    std::swap(s->state, s->next);
    return LGD_OK;
  }
  if ((rc = lgd_fill_ghosts(s->ctx, s->grid, s->state[0], 1, s->has_bcs ? s->bcs : nullptr)))
    return rc;
  if ((rc = lgd_heat_rhs(s->ctx, s->grid, s->kappa, s->state[0], s->k[0]))) return rc;
  return lgd_axpy(s->ctx, s->state[0], s->k[0], dt);
}

// cell_center (src/geometry/cartesian.rs:589-597) for a 2-D grid.
inline void cell_center2(const lgd_cart &c, int64_t b, int64_t i, int64_t j, double out[2]) {
  int64_t bx = b % c.nb[0], by = b / c.nb[0];
  out[0] = std::fma((double)(bx * c.n[0]) + (double)i + 0.5, c.h[0], c.origin[0]);
  out[1] = std::fma((double)(by * c.n[1]) + (double)j + 0.5, c.h[1], c.origin[1]);
}


// ------------------------------------------------------------------ AMR path
int nfields(const lgd_sim *s) { return s->kind == lgd_sim::kModelC ? 2 : 1; }

// Stage buffers of a simulation on an AmrGrid: Tendency(Time) per field, the
// Wiener(0) amplitude (Model C with noise), one snapshot per (level, field)
// when subcycling (src/integrators/subcycling.rs:63-73).  After a regrid the
// buffers are kept when they still hold the new field length (kernels address
// them by block offsets, so a longer buffer is fine) and grown by a quarter
// otherwise; tendencies are cleared (their ghost entries stay zero, as the
// reference's zero-initialised stages), snapshots need no clearing (a level's
// range is copied in before it is read).
int alloc_amr_stages(lgd_sim *s) {
  const size_t len = (size_t)s->amr->off.back();
  const int nf = nfields(s);
  const bool noise = s->kind == lgd_sim::kModelC && s->wiener;
  const size_t nsnap = s->subcycle ? (size_t)(s->amr->max_levels() * nf) : 0;
  bool fits = s->k.size() == (size_t)nf && s->snaps.size() == nsnap && (!noise || s->amp);
  for (auto *b : s->k) fits = fits && b->len >= len;
  for (auto *b : s->snaps) fits = fits && b->len >= len;
  if (noise && s->amp) fits = fits && s->amp->len >= len;
  int rc;
  if (fits) {
    for (auto *b : s->k)
      if ((rc = lgd_buf_fill_zero(s->ctx, b))) return rc;
    if (noise && (rc = lgd_buf_fill_zero(s->ctx, s->amp))) return rc;
    return LGD_OK;
  }
  const size_t cap = s->k.empty() ? len : len + len / 4;  // grow with headroom after a regrid
  free_stages(s);
  for (int f = 0; f < nf; ++f) {
    lgd_buf *b = nullptr;
    if ((rc = lgd_buf_alloc(s->ctx, cap, &b))) return rc;
    s->k.push_back(b);
  }
  if (noise && (rc = lgd_buf_alloc(s->ctx, cap, &s->amp))) return rc;
  for (size_t i = 0; i < nsnap; ++i) {
    lgd_buf *b = nullptr;
    if ((rc = lgd_buf_alloc(s->ctx, cap, &b))) return rc;
    s->snaps.push_back(b);
  }
  return LGD_OK;
}

double amr_stable_dt(const lgd_sim *s, int level) {
  double h = std::fmin(s->amr->spacing[(size_t)level][0], s->amr->spacing[(size_t)level][1]);
  return s->kind == lgd_sim::kModelC ? 0.2 * std::pow(h, 2) / s->params.d_thermal
                                     : 0.25 * std::pow(h, 2) / s->kappa;
}

// Vector field + stage update on one level's patches (eval_drift_level,
// eval_tendency_level, apply_step_level: src/integrators/mod.rs:205-240,
// src/core/state.rs:706-753).
int amr_level_update(lgd_sim *s, int level, double dt, uint64_t salt) {
  lgd_ctx *ctx = s->ctx;
  const lgd_amr &g = *s->amr;
  int rc;
  if (s->kind == lgd_sim::kModelC)
    rc = lgd_amr_modelc_rhs(ctx, g, s->params, level, s->state[0]->data, s->state[1]->data,
                            s->k[0]->data, s->k[1]->data);
  else
    rc = lgd_amr_heat_rhs(ctx, g, s->kappa, level, s->state[0]->data, s->k[0]->data);
  return rc;
}

int amr_level_apply(lgd_sim *s, int level, double dt, uint64_t salt) {
  int rc;
  for (int f = 0; f < nfields(s); ++f)
    if ((rc = lgd_amr_apply(s->ctx, *s->amr, level, s->state[f]->data, s->k[f]->data, dt))) return rc;
  if (s->kind == lgd_sim::kModelC && s->wiener && s->params.noise_amplitude != 0.0) {
    double scale = std::sqrt(dt) * s->params.noise_amplitude;
    if ((rc = lgd_amr_wiener(s->ctx, *s->amr, level, s->state[0]->data, scale, s->seed, salt)))
      return rc;
  }
  return LGD_OK;
}

// EulerMaruyama on the AmrGrid with one global dt: ModelC/Heat fill_ghosts
// (restrict every field finest→coarsest, then fill every level), every level's
// vector field, then every level's update.
int amr_global_step(lgd_sim *s, double dt, uint64_t salt) {
  const lgd_amr &g = *s->amr;
  int rc, nf = nfields(s);
  for (int f = 0; f < nf; ++f)
    for (int l = g.num_levels() - 1; l >= 1; --l)
      if ((rc = lgd_amr_restrict(s->ctx, g, s->state[f]->data, l))) return rc;
  for (int f = 0; f < nf; ++f)
    for (int l = 0; l < g.num_levels(); ++l)
      if ((rc = lgd_amr_fill(s->ctx, g, s->state[f]->data, nullptr, 0.0, l))) return rc;
  for (int l = 0; l < g.num_levels(); ++l)
    if ((rc = amr_level_update(s, l, dt, salt))) return rc;
  for (int l = 0; l < g.num_levels(); ++l)
    if ((rc = amr_level_apply(s, l, dt, salt))) return rc;
  return LGD_OK;
}

// Subcycle::advance (src/integrators/subcycling.rs:195-267), flattened into a
// launch sequence: snapshot the level, fill its ghosts (coarse source blended in
// time from the parent's snapshot), step it, recurse n = round(dt_l/dt_{l+1})
// times into the finer level at t + s·dt_child (fma), then restrict it back.
int amr_advance(lgd_sim *s, int level, double t, double dt, double alpha) {
  const lgd_amr &g = *s->amr;
  const int nf = nfields(s);
  int rc;
  // Only this level's cells of a snapshot are ever read (as prolongation
  // donors of the finer level), and a level's patches are contiguous.
  int64_t lo = g.off[g.level_start[level]], hi = g.off[g.level_start[level + 1]];
  for (int f = 0; f < nf; ++f) {
    hipError_t e = hipMemcpyAsync(s->snaps[(size_t)(level * nf + f)]->data + lo,
                                  s->state[f]->data + lo, (size_t)(hi - lo) * sizeof(double),
                                  hipMemcpyDeviceToDevice, s->ctx->stream);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "snapshot");
  }
  for (int f = 0; f < nf; ++f) {
    const double *old = level > 0 ? s->snaps[(size_t)((level - 1) * nf + f)]->data : nullptr;
    if ((rc = lgd_amr_fill(s->ctx, g, s->state[f]->data, old, alpha, level))) return rc;
  }
  uint64_t salt;
  std::memcpy(&salt, &t, sizeof salt);
  if ((rc = amr_level_update(s, level, dt, salt))) return rc;
  if ((rc = amr_level_apply(s, level, dt, salt))) return rc;
  const int child = level + 1;
  if (child < g.num_levels()) {
    double n = std::fmax(std::round(amr_stable_dt(s, level) / amr_stable_dt(s, child)), 1.0);
    uint64_t nn = (uint64_t)n;
    double dt_child = dt / (double)nn;
    for (uint64_t k = 0; k < nn; ++k)
      if ((rc = amr_advance(s, child, std::fma((double)k, dt_child, t), dt_child,
                            (double)k / (double)nn)))
        return rc;
    for (int f = 0; f < nf; ++f)
      if ((rc = lgd_amr_restrict(s->ctx, g, s->state[f]->data, child))) return rc;
  }
  return LGD_OK;
}

// BergerOliger::regrid at the top of a step (src/core/simulation.rs:255-270).
int amr_regrid(lgd_sim *s) {
  lgd_amr &g = *s->amr;
  const int tf = s->policy.tag_field;
  int rc;
  // GradientTagger::prepare: restrict + fill the tagged field (adapt.rs:72-76).
  for (int l = g.num_levels() - 1; l >= 1; --l)
    if ((rc = lgd_amr_restrict(s->ctx, g, s->state[(size_t)tf]->data, l))) return rc;
  for (int l = 0; l < g.num_levels(); ++l)
    if ((rc = lgd_amr_fill(s->ctx, g, s->state[(size_t)tf]->data, nullptr, 0.0, l))) return rc;
  std::vector<double> tag((size_t)g.off.back());
  if ((rc = lgd_buf_download(s->ctx, s->state[(size_t)tf], 0, tag.data(), tag.size()))) return rc;
  std::vector<std::vector<lgd_box>> levels;
  if (lgd_regrid_propose(g, tag, s->policy, levels) == 0) return LGD_OK;
  lgd_amr *ng = new (std::nothrow) lgd_amr();
  if (!ng) return fail(s->ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  if ((rc = lgd_amr_build(s->ctx, g.base, g.ratios, levels, *ng))) {
    ng->free_device();
    delete ng;
    return rc;
  }
  std::vector<lgd_buf *> ns;
  std::vector<double> old_f, new_f;
  const int64_t fine0 = ng->off[ng->level_start[1]];  // level-0 elements (same in both)
  for (size_t f = 0; f < s->state.size(); ++f) {
    if ((int)f == tf) {
      old_f.swap(tag);  // already downloaded
    } else {
      old_f.resize((size_t)g.off.back());
      if ((rc = lgd_buf_download(s->ctx, s->state[f], 0, old_f.data(), old_f.size()))) return rc;
    }
    lgd_regrid_migrate(g, old_f, *ng, new_f);
    lgd_buf *b = nullptr;
    if ((rc = lgd_buf_alloc(s->ctx, (size_t)ng->off.back(), &b))) return rc;
    hipError_t e = hipMemcpyAsync(b->data, s->state[f]->data, (size_t)fine0 * sizeof(double),
                                  hipMemcpyDeviceToDevice, s->ctx->stream);
    if (e != hipSuccess) return hip_fail(s->ctx, e, "migrate level 0");
    if (!new_f.empty() && (rc = lgd_buf_upload(s->ctx, b, (size_t)fine0, new_f.data(), new_f.size())))
      return rc;
    ns.push_back(b);
  }
  for (auto *b : s->state) lgd_buf_free(b);
  s->state = ns;
  (void)hipStreamSynchronize(s->ctx->stream);
  s->amr->free_device();
  delete s->amr;
  s->amr = ng;
  return alloc_amr_stages(s);
}

int step_amr(lgd_sim *s, double dt, uint64_t salt) {
  int rc;
  if (s->adaptive && s->step_index % s->policy.every == 0 && (rc = amr_regrid(s))) return rc;
  if (s->subcycle) return amr_advance(s, 0, s->t, dt, 0.0);
  return amr_global_step(s, dt, salt);
}

}  // namespace

extern "C" {

int lgd_sim_create_modelc(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                          int32_t wiener, uint64_t seed, int32_t mode, lgd_sim **out) {
  if (!ctx || !grid || !params || !out) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (grid->cart.dim != 2) return fail(ctx, LGD_ERR_UNSUPPORTED, "Model C runs on a 2-D grid");
  if (mode != LGD_MODE_FUSED && mode != LGD_MODE_PER_INTENT)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "unknown mode");
  int rc = check_ghost(ctx, grid, 1);
  if (rc) return rc;
  lgd_sim *s = new (std::nothrow) lgd_sim();
  if (!s) return fail(ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->ctx = ctx;
  s->grid = grid;
  s->kind = lgd_sim::kModelC;
  s->mode = mode;
  s->params = *params;
  s->wiener = wiener ? 1 : 0;
  s->seed = seed;
  if ((rc = alloc_fields(s, 2))) {
    free_all(s);
    delete s;
    return rc;
  }
  *out = s;
  return LGD_OK;
}

int lgd_sim_create_heat(lgd_ctx *ctx, const lgd_grid *grid, double kappa,
                        const lgd_face_bc *bcs, int32_t mode, lgd_sim **out) {
  if (!ctx || !grid || !out) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (mode != LGD_MODE_FUSED && mode != LGD_MODE_PER_INTENT)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "unknown mode");
  int rc = check_ghost(ctx, grid, 1);
  if (rc) return rc;
  lgd_sim *s = new (std::nothrow) lgd_sim();
  if (!s) return fail(ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->ctx = ctx;
  s->grid = grid;
  s->kind = lgd_sim::kHeat;
  s->mode = mode;
  s->kappa = kappa;
  if (bcs) {
    s->has_bcs = true;
    for (int i = 0; i < 2 * grid->cart.dim; ++i) s->bcs[i] = bcs[i];
  }
  if (mode == LGD_MODE_FUSED && !all_mirror(s)) {
    delete s;
    return fail(ctx, LGD_ERR_UNSUPPORTED,
                "the fused heat step implements mirror and periodic faces; use "
                "LGD_MODE_PER_INTENT for Dirichlet/Flux/Extrapolate faces");
  }
  if ((rc = alloc_fields(s, 1))) {
    free_all(s);
    delete s;
    return rc;
  }
  *out = s;
  return LGD_OK;
}

int lgd_sim_create_amr(lgd_ctx *ctx, lgd_amr *amr, int32_t model, const lgd_modelc_params *params,
                       double kappa, int32_t wiener, uint64_t seed, int32_t subcycle,
                       const lgd_regrid_policy *policy, lgd_sim **out) {
  if (!ctx || !amr || !out) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (model != 0 && model != 1) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "model is 0 (heat) or 1 (Model C)");
  if (model == 1 && !params) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "Model C needs parameters");
  if (policy && (policy->every < 1 || policy->tag_field < 0 || policy->tag_field >= (model == 1 ? 2 : 1) ||
                 policy->min_width < 1 || !(policy->efficiency > 0.0 && policy->efficiency <= 1.0)))
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "invalid regrid policy");
  lgd_sim *s = new (std::nothrow) lgd_sim();
  if (!s) return fail(ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  s->ctx = ctx;
  s->grid = nullptr;
  s->kind = model == 1 ? lgd_sim::kModelC : lgd_sim::kHeat;
  s->mode = LGD_MODE_PER_INTENT;
  if (params) s->params = *params;
  s->kappa = kappa;
  s->wiener = wiener ? 1 : 0;
  s->seed = seed;
  s->amr = amr;
  s->subcycle = subcycle ? 1 : 0;
  if (policy) {
    s->adaptive = true;
    s->policy = *policy;
  }
  int rc = LGD_OK;
  size_t len = (size_t)amr->off.back();
  for (int f = 0; f < nfields(s) && !rc; ++f) {
    lgd_buf *b = nullptr;
    rc = lgd_buf_alloc(ctx, len, &b);
    if (!rc) s->state.push_back(b);
  }
  if (!rc) rc = alloc_amr_stages(s);
  if (rc) {
    free_all(s);
    s->amr = nullptr;  // the caller keeps the hierarchy on failure
    delete s;
    return rc;
  }
  *out = s;
  return LGD_OK;
}

const lgd_amr *lgd_sim_amr(const lgd_sim *sim) { return sim ? sim->amr : nullptr; }

int lgd_sim_destroy(lgd_sim *sim) {
  if (!sim) return LGD_OK;
  free_all(sim);
  if (sim->amr) lgd_amr_destroy(sim->amr);
  delete sim;
  return LGD_OK;
}

int lgd_sim_field(lgd_sim *sim, int32_t index, lgd_buf **out) {
  if (!sim || !out) return LGD_ERR_INVALID_ARGUMENT;
  if (index < 0 || index >= (int32_t)sim->state.size())
    return fail(sim->ctx, LGD_ERR_INVALID_ARGUMENT, "field index out of range");
  *out = sim->state[index];
  return LGD_OK;
}

int lgd_sim_step(lgd_sim *sim, double dt, int64_t nsteps) {
  if (!sim) return LGD_ERR_INVALID_ARGUMENT;
  for (int64_t n = 0; n < nsteps; ++n) {
    uint64_t salt;
    std::memcpy(&salt, &sim->t, sizeof salt);  // t.to_bits()
    int rc = sim->amr ? step_amr(sim, dt, salt)
             : sim->kind == lgd_sim::kModelC ? step_modelc(sim, dt, salt)
                                             : step_heat(sim, dt);
    if (rc) return rc;
    sim->t = sim->t + dt;
    sim->step_index += 1;
  }
  return LGD_OK;
}

double lgd_sim_time(const lgd_sim *sim) { return sim ? sim->t : 0.0; }

int64_t lgd_sim_step_index(const lgd_sim *sim) { return sim ? sim->step_index : -1; }

int lgd_sim_initialize_grains(lgd_sim *sim, const double *grains, size_t ngrains, double u0) {
  if (!sim || (!grains && ngrains)) return LGD_ERR_INVALID_ARGUMENT;
  if (sim->kind != lgd_sim::kModelC)
    return fail(sim->ctx, LGD_ERR_INVALID_ARGUMENT, "initialize_grains is a Model C initial condition");
  const lgd_cart &c = sim->grid->cart;
  const double cutoff = 27.0;
  const double sqrt2 = std::sqrt(2.0);
  int64_t nx = c.n[0], ny = c.n[1], sx = nx + 2;
  int64_t slab = sx * (ny + 2);
  std::vector<double> phi((size_t)(c.nblocks * slab), 0.0), u((size_t)(c.nblocks * slab), 0.0);
  std::vector<size_t> near;
  for (int64_t b = 0; b < c.nblocks; ++b) {
    double hx = c.h[0], hy = c.h[1], c0[2], c1[2];
    cell_center2(c, b, 0, 0, c0);
    cell_center2(c, b, nx - 1, ny - 1, c1);
    double lo[2] = {std::fma(0.5, -hx, c0[0]), std::fma(0.5, -hy, c0[1])};
    double hi[2] = {std::fma(0.5, hx, c1[0]), std::fma(0.5, hy, c1[1])};
    near.clear();
    for (size_t g = 0; g < ngrains; ++g) {
      double cx = grains[3 * g], cy = grains[3 * g + 1];
      double dx = std::fmax(std::fmax(lo[0] - cx, cx - hi[0]), 0.0);
      double dy = std::fmax(std::fmax(lo[1] - cy, cy - hi[1]), 0.0);
      if (std::hypot(dx, dy) <= grains[3 * g + 2] + cutoff) near.push_back(g);
    }
    double *pb = phi.data() + b * slab, *ub = u.data() + b * slab;
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        double v = -1.0;
        if (!near.empty()) {
          double xy[2];
          cell_center2(c, b, i, j, xy);
          for (size_t g : near) {
            double r = std::hypot(xy[0] - grains[3 * g], xy[1] - grains[3 * g + 1]);
            v = std::fmax(v, -std::tanh((r - grains[3 * g + 2]) / sqrt2));
          }
        }
        pb[(j + 1) * sx + (i + 1)] = v;
        ub[(j + 1) * sx + (i + 1)] = -u0;
      }
  }
  int rc = lgd_buf_upload(sim->ctx, sim->state[0], 0, phi.data(), phi.size());
  if (rc) return rc;
  return lgd_buf_upload(sim->ctx, sim->state[1], 0, u.data(), u.size());
}

}  // extern "C"
