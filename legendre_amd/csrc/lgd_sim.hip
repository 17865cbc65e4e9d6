// Host-side orchestration: the GpuScheduler-driven Simulation.
//
// Mirrors Simulation::step (src/core/simulation.rs:255-290) with
// EulerMaruyama::step (src/integrators/euler_maruyama.rs:34-75) or
// ForwardEuler::step (src/integrators/euler.rs:104-130): the host owns the
// control flow and the f64 time accumulation (t += dt, the noise salt), the
// device owns every per-block sweep.  Initial conditions are host work, as in
// the reference (ModelC::initialize_grains, src/physics/phasefield.rs:227-314).
//
// Whole-step graphs (lgd_sim_set_graph_steps; no reference counterpart): a run of
// steps — every launch of each step, the subcycled level recursion and the time
// advance — is captured once into a HIP graph and replayed.  The noise salt, t's
// bits, is the one value that changes from step to step; during capture the noise
// kernels read it from a device copy of the times (d_t[0] = t, d_t[l] = the current
// substep time of AMR level l), which small kernels in the graph advance with the
// host's own arithmetic (t + dt, fma(k, dt_l, t)), so replays are bitwise equal to
// eager steps.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <limits>
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
  std::vector<int8_t> cur;       // AMR: per level, whether next[] holds its current values
  // Integrator (LGD_INTEGRATOR_*); RK4 stages: k1..k4 then y_tmp, one buffer per field each.
  int32_t integrator = LGD_INTEGRATOR_EULER;
  std::vector<lgd_buf *> rk;
  // Multi-grain Model C: θ₀ per cell (ghost-0 slabs) as cos θ₀ and sin θ₀, and θ₀
  // itself for observers.
  lgd_buf *cos0 = nullptr, *sin0 = nullptr, *theta0 = nullptr;
  std::vector<lgd_observer *> observers;  // attached, not owned
  double t = 0.0;
  int64_t step_index = 0;
  // Graphs.
  struct graph {
    lgd_buf *state0;   // state[0] at capture (the fused step's ping-pong parity)
    uint64_t cur0;     // AMR: the levels' buffer parities at capture, and after
    uint64_t cur1;
    int64_t epoch;     // AMR hierarchy epoch
    int64_t tuning;    // context tuning epoch
    uint64_t dt_bits;
    int64_t nsteps;
    hipGraphExec_t exec;
  };
  int64_t graph_steps = 0;    // 0: eager steps
  std::vector<graph> graphs;  // at most the two parities of one key
  int64_t amr_epoch = 0;      // bumped by every regrid
  double *d_t = nullptr;      // device times: t, then one substep time per level
  bool capturing = false;     // noise launches read their salt from d_t
  bool d_t_known = false;     // d_t[0] holds d_t_value
  double d_t_value = 0.0;
  // Regrid wall time by phase (seconds): tag field to the host, proposal (tagging
  // and clustering), new hierarchy (plans and uploads), state migration; and count.
  double regrid_seconds[4] = {0.0, 0.0, 0.0, 0.0};
  int64_t regrids = 0;
  lgd_regrid_scratch rscratch;  // device tagging scratch
  // A caller's TagCells::tag_level in place of the GradientTagger (NULL: gradient).
  lgd_tag_level_fn tagger = nullptr;
  void *tagger_user = nullptr;
};

// The flags a TagCells::tag_level call appends to (n × [x, y]).
struct lgd_tag_sink {
  std::vector<int64_t> cells;
};

using namespace lgd;

namespace {

// Simulation::step's time accumulation t += dt, and Subcycle's substep times
// fma(k, dt_child, t) (src/integrators/subcycling.rs:245-250), on the device.
__global__ void time_advance_kernel(double *t, double dt) { t[0] = t[0] + dt; }
__global__ void time_set_kernel(double *t, double v) { t[0] = v; }
__global__ void time_substep_kernel(double *t, int dst, int src, double k, double dtc) {
  t[dst] = fma(k, dtc, t[src]);
}
__global__ void copy_kernel(int64_t n, const double *src, double *dst) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) dst[i] = src[i];
}

// The device time the noise of a (sub)step at `level` reads while capturing.
const double *time_slot(const lgd_sim *s, int level) {
  return s->capturing ? s->d_t + level : nullptr;
}

// Graphs are destroyed only once the stream has drained (a launch may be in flight).
void drop_graphs(lgd_sim *s) {
  if (s->graphs.empty()) return;
  (void)hipStreamSynchronize(s->ctx->stream);
  for (auto &g : s->graphs) (void)hipGraphExecDestroy(g.exec);
  s->graphs.clear();
}

void free_stages(lgd_sim *s) {
  for (auto *b : s->next) lgd_buf_free(b);
  for (auto *b : s->k) lgd_buf_free(b);
  lgd_buf_free(s->amp);
  s->next.clear();
  s->k.clear();
  s->amp = nullptr;
}

void free_all(lgd_sim *s) {
  for (auto *b : s->state) lgd_buf_free(b);
  s->state.clear();
  for (auto *b : s->rk) lgd_buf_free(b);
  s->rk.clear();
  lgd_buf_free(s->cos0);
  lgd_buf_free(s->sin0);
  lgd_buf_free(s->theta0);
  s->cos0 = s->sin0 = s->theta0 = nullptr;
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

// eval_drift of Model C: ModelC::fill_ghosts, then the per-block Time vector field
// (axis-aligned or multi-grain) of (phi, u) into (kphi, ku).
int modelc_eval(lgd_sim *s, lgd_buf *phi, lgd_buf *u, lgd_buf *kphi, lgd_buf *ku) {
  lgd_ctx *ctx = s->ctx;
  int rc;
  if ((rc = lgd_fill_ghosts(ctx, s->grid, phi, 1, nullptr))) return rc;
  if ((rc = lgd_fill_ghosts(ctx, s->grid, u, 1, nullptr))) return rc;
  if (s->cos0) return lgd_modelc_rhs_oriented(ctx, s->grid, &s->params, phi, u, s->cos0, s->sin0, kphi, ku);
  return lgd_modelc_rhs(ctx, s->grid, &s->params, phi, u, kphi, ku);
}

int heat_eval(lgd_sim *s, lgd_buf *u, lgd_buf *k) {
  int rc;
  if ((rc = lgd_fill_ghosts(s->ctx, s->grid, u, 1, s->has_bcs ? s->bcs : nullptr))) return rc;
  return lgd_heat_rhs(s->ctx, s->grid, s->kappa, u, k);
}

int copy_buf(lgd_ctx *ctx, const lgd_buf *src, lgd_buf *dst) {
  copy_kernel<<<blocks_for((int64_t)src->len, 256), 256, 0, ctx->stream>>>((int64_t)src->len,
                                                                         src->data, dst->data);
  return launch_check(ctx, "copy");
}

// RungeKutta4::step (src/integrators/rk4.rs:30-80): k1 = f(y); y_tmp = y + dt/2·k1,
// k2 = f(y_tmp); y_tmp = y + dt/2·k2, k3; y_tmp = y + dt·k3, k4; then
// y += dt/6·k1, += dt/3·k2, += dt/3·k3, += dt/6·k4, each a whole-slab sweep.
int step_rk4(lgd_sim *s, double dt) {
  const int nf = (int)s->state.size();
  if (s->kind == lgd_sim::kModelC && s->mode == LGD_MODE_FUSED) {
    // Fused stages: the same operations in the same order, half the field passes.
    lgd_buf *kb[2] = {s->rk[0], s->rk[1]};
    lgd_buf *tb[4] = {s->rk[2], s->rk[3], s->rk[4], s->rk[5]};
    return lgd_modelc_rk4_step(s->ctx, s->grid, &s->params, s->state[0], s->state[1], kb, tb, dt,
                               s->cos0 ? s->cos0->data : nullptr, s->sin0 ? s->sin0->data : nullptr);
  }
  auto k = [&](int stage, int f) { return s->rk[(size_t)(stage * nf + f)]; };
  auto y = [&](int f) { return s->rk[(size_t)(4 * nf + f)]; };
  auto eval = [&](bool from_state, int stage) {
    std::vector<lgd_buf *> in(nf);
    for (int f = 0; f < nf; ++f) in[f] = from_state ? s->state[f] : y(f);
    return s->kind == lgd_sim::kModelC ? modelc_eval(s, in[0], in[1], k(stage, 0), k(stage, 1))
                                       : heat_eval(s, in[0], k(stage, 0));
  };
  int rc;
  if ((rc = eval(true, 0))) return rc;
  const double half_dt = 0.5 * dt;
  const double c[3] = {half_dt, half_dt, dt};
  for (int st = 1; st < 4; ++st) {
    for (int f = 0; f < nf; ++f) {
      if ((rc = copy_buf(s->ctx, s->state[f], y(f)))) return rc;
      if ((rc = lgd_axpy(s->ctx, y(f), k(st - 1, f), c[st - 1]))) return rc;
    }
    if ((rc = eval(false, st))) return rc;
  }
  const double sixth = dt / 6.0, third = dt / 3.0;
  const double w[4] = {sixth, third, third, sixth};
  for (int st = 0; st < 4; ++st)
    for (int f = 0; f < nf; ++f)
      if ((rc = lgd_axpy(s->ctx, s->state[f], k(st, f), w[st]))) return rc;
  return LGD_OK;
}

int step_modelc(lgd_sim *s, double dt, uint64_t salt) {
  lgd_ctx *ctx = s->ctx;
  int rc;
  if (s->integrator == LGD_INTEGRATOR_RK4) return step_rk4(s, dt);
  if (s->mode == LGD_MODE_FUSED) {
    ctx->salt_src = time_slot(s, 0);
    rc = lgd_modelc_step_impl(ctx, s->grid, &s->params, s->state[0], s->state[1], s->next[0],
                              s->next[1], dt, s->seed, salt, s->wiener,
                              s->cos0 ? s->cos0->data : nullptr, s->sin0 ? s->sin0->data : nullptr);
    ctx->salt_src = nullptr;
    return rc;  // step_done swaps the ping-pong pair
  }
  if ((rc = modelc_eval(s, s->state[0], s->state[1], s->k[0], s->k[1]))) return rc;
  if (s->wiener && (rc = lgd_modelc_amplitude(ctx, s->grid, &s->params, s->amp))) return rc;
  // apply_step_with: Time on phi and u, then Wiener(0) on phi.
  if ((rc = lgd_axpy(ctx, s->state[0], s->k[0], dt))) return rc;
  if ((rc = lgd_axpy(ctx, s->state[1], s->k[1], dt))) return rc;
  if (s->wiener) {
    const uint64_t stream = (1ull << 32) | 0ull;
    ctx->salt_src = time_slot(s, 0);
    rc = lgd_apply_wiener(ctx, s->grid, s->state[0], s->amp, 1, dt, s->seed, salt, stream);
    ctx->salt_src = nullptr;
    if (rc) return rc;
  }
  return LGD_OK;
}

int step_heat(lgd_sim *s, double dt) {
  int rc;
  if (s->integrator == LGD_INTEGRATOR_RK4) return step_rk4(s, dt);
  if (s->mode == LGD_MODE_FUSED) {
    return lgd_heat_step_bc(s->ctx, s->grid, s->kappa, s->has_bcs ? s->bcs : nullptr, s->state[0],
                            s->next[0], dt);
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

// Ping-pong state: state[f] and next[f] hold every field twice and cur[l] says
// which one holds level l's current values (0: state, 1: next).  A substep of
// level l reads its current buffer and writes the other, so the values before the
// substep stay in place — the coarse source the finer level blends with in time
// (subcycling.rs:63-73) — without a snapshot copy.  Ghost entries are scratch:
// every substep fills the ghosts it reads first.
lgd_buf *amr_cur(const lgd_sim *s, int f, int level) {
  return s->cur[(size_t)level] ? s->next[(size_t)f] : s->state[(size_t)f];
}
lgd_buf *amr_prev(const lgd_sim *s, int f, int level) {
  return s->cur[(size_t)level] ? s->state[(size_t)f] : s->next[(size_t)f];
}

// Bring every level's current values into state[] (copying the levels that live
// in next[]; the larger part is moved by swapping the pair when level 0 is there).
int amr_consolidate(lgd_sim *s) {
  const lgd_amr &g = *s->amr;
  if (s->cur[0]) {
    std::swap(s->state, s->next);
    for (auto &c : s->cur) c ^= 1;
  }
  int rc;
  for (int l = 1; l < g.num_levels(); ++l) {
    if (!s->cur[(size_t)l]) continue;
    for (int f = 0; f < nfields(s); ++f)
      if ((rc = lgd_amr_copy_level(s->ctx, g, l, s->next[(size_t)f]->data, s->state[(size_t)f]->data)))
        return rc;
    s->cur[(size_t)l] = 0;
  }
  return LGD_OK;
}

// The second buffer of every field, at least the state's length; kept while it is
// large enough, else reallocated with a quarter of headroom.
int alloc_amr_stages(lgd_sim *s) {
  const size_t len = (size_t)s->amr->off.back();
  const int nf = nfields(s);
  int rc;
  if (s->next.size() != (size_t)nf) {
    for (auto *b : s->next) lgd_buf_free(b);
    s->next.assign((size_t)nf, nullptr);
  }
  for (int f = 0; f < nf; ++f) {
    lgd_buf *&b = s->next[(size_t)f];
    if (b && b->cap >= len) {
      b->len = len;
      continue;
    }
    lgd_buf_free(b);
    b = nullptr;
    if ((rc = lgd_buf_alloc(s->ctx, s->regrids ? len + len / 4 : len, &b))) return rc;
    b->len = len;
  }
  s->cur.assign((size_t)s->amr->max_levels(), 0);
  return LGD_OK;
}

double amr_stable_dt(const lgd_sim *s, int level) {
  double h = std::fmin(s->amr->spacing[(size_t)level][0], s->amr->spacing[(size_t)level][1]);
  return s->kind == lgd_sim::kModelC ? 0.2 * std::pow(h, 2) / s->params.d_thermal
                                     : 0.25 * std::pow(h, 2) / s->kappa;
}

// Fill the ghosts of `level` in its current buffers: same-level copies and
// prolongation from the parent (blended between the parent's previous and current
// values when `blend`).  Level 0 on the streaming kernel needs none.  At α = 0 the
// blend old·1 + new·0 is the corner sum of the old values bit for bit (that sum
// starts from +0, so it is never −0) whenever the new ones are finite: the fill
// reads the previous buffer alone (half the corner loads).
int amr_fill_level(lgd_sim *s, int level, bool blend, double alpha) {
  const lgd_amr &g = *s->amr;
  if (level == 0 && s->kind == lgd_sim::kModelC && lgd_amr_level0_streams(s->ctx, g)) return LGD_OK;
  const int nf = nfields(s);
  const bool at_old = blend && alpha == 0.0;
  double *lvl[2] = {nullptr, nullptr};
  const double *par[2] = {nullptr, nullptr}, *old[2] = {nullptr, nullptr};
  for (int f = 0; f < nf; ++f) {
    lvl[f] = amr_cur(s, f, level)->data;
    if (level > 0) {
      par[f] = at_old ? amr_prev(s, f, level - 1)->data : amr_cur(s, f, level - 1)->data;
      old[f] = amr_prev(s, f, level - 1)->data;
    }
  }
  return lgd_amr_fill_pp(s->ctx, g, level, nf, lvl, level > 0 ? par : nullptr,
                         level > 0 && blend && !at_old ? old : nullptr, alpha);
}

// Restrict level + 1 into `level` and fill `level`'s ghosts (as amr_fill_level with
// blending) in one launch.
int amr_restrict_fill_level(lgd_sim *s, int level, double alpha) {
  const int nf = nfields(s);
  double *lvl[2] = {nullptr, nullptr};
  const double *par[2] = {nullptr, nullptr}, *old[2] = {nullptr, nullptr}, *fine[2] = {nullptr, nullptr};
  for (int f = 0; f < nf; ++f) {
    lvl[f] = amr_cur(s, f, level)->data;
    par[f] = amr_cur(s, f, level - 1)->data;
    old[f] = amr_prev(s, f, level - 1)->data;
    fine[f] = amr_cur(s, f, level + 1)->data;
  }
  return lgd_amr_restrict_fill_pp(s->ctx, *s->amr, level, nf, lvl, par, old, alpha, fine);
}

// Restrict `fine_level` into its parent (both current).
int amr_restrict_level(lgd_sim *s, int fine_level) {
  const int nf = nfields(s);
  const double *fine[2] = {nullptr, nullptr};
  double *coarse[2] = {nullptr, nullptr};
  for (int f = 0; f < nf; ++f) {
    fine[f] = amr_cur(s, f, fine_level)->data;
    coarse[f] = amr_cur(s, f, fine_level - 1)->data;
  }
  return lgd_amr_restrict_pp(s->ctx, *s->amr, fine_level, nf, fine, coarse);
}

// One fused stage of `level` (eval_drift_level + apply_step_level,
// src/integrators/mod.rs:205-240, src/core/state.rs:706-753): current → other
// buffer, which becomes current.  `tslot`: the device time slot this substep's
// noise reads while a graph is captured.
// With next_alpha >= 0 the same launch also fills the ghosts of the level's next
// substep (a level without children; the parent's buffers blended at next_alpha);
// with child_fill the ghosts of level + 1 for its first substep (blended at α = 0);
// with parent_step 1 the restriction of the level's new values into its parent, 2
// also the parent's ghosts for its next substep (the grandparent blended at
// parent_alpha).
int amr_level_step(lgd_sim *s, int level, double dt, uint64_t salt, int tslot,
                   double next_alpha = -1.0, bool child_fill = false, int parent_step = 0,
                   double parent_alpha = 0.0) {
  const lgd_amr &g = *s->amr;
  int rc;
  lgd_amr_child_fill cf{};
  if (child_fill)
    for (int f = 0; f < nfields(s); ++f) cf.lvl[f] = amr_cur(s, f, level + 1)->data;
  lgd_amr_parent_step ps{};
  if (parent_step) {
    ps.fill = parent_step == 2;
    ps.alpha = parent_alpha;
    for (int f = 0; f < nfields(s); ++f) {
      ps.coarse[f] = amr_cur(s, f, level - 1)->data;
      if (ps.fill) {
        ps.par[f] = amr_cur(s, f, level - 2)->data;
        ps.old[f] = amr_prev(s, f, level - 2)->data;
      }
    }
  }
  lgd_amr_next_fill nf{};
  if (next_alpha >= 0.0) {
    for (int f = 0; f < nfields(s); ++f) {
      nf.par[f] = amr_cur(s, f, level - 1)->data;
      nf.old[f] = amr_prev(s, f, level - 1)->data;
    }
    nf.alpha = next_alpha;
  }
  const lgd_amr_next_fill *nfp = next_alpha >= 0.0 ? &nf : nullptr;
  if (s->kind == lgd_sim::kModelC) {
    s->ctx->salt_src = time_slot(s, tslot);
    rc = lgd_amr_modelc_step(s->ctx, g, s->params, level, dt, amr_cur(s, 0, level)->data,
                             amr_cur(s, 1, level)->data, amr_prev(s, 0, level)->data,
                             amr_prev(s, 1, level)->data, s->seed, salt, s->wiener != 0, nfp,
                             child_fill ? &cf : nullptr, parent_step ? &ps : nullptr);
    s->ctx->salt_src = nullptr;
  } else if (child_fill || parent_step) {
    return fail(s->ctx, LGD_ERR_UNSUPPORTED, "fused fills and restrictions share only a Model C stage");
  } else {
    rc = lgd_amr_heat_step(s->ctx, g, s->kappa, level, dt, amr_cur(s, 0, level)->data,
                           amr_prev(s, 0, level)->data, nfp);
  }
  if (rc) return rc;
  s->cur[(size_t)level] ^= 1;
  return LGD_OK;
}

// EulerMaruyama on the AmrGrid with one global dt: ModelC/Heat fill_ghosts
// (restrict every field finest→coarsest, then fill every level), then every
// level's stage.
int amr_global_step(lgd_sim *s, double dt, uint64_t salt) {
  const lgd_amr &g = *s->amr;
  int rc;
  for (int l = g.num_levels() - 1; l >= 1; --l)
    if ((rc = amr_restrict_level(s, l))) return rc;
  for (int l = 0; l < g.num_levels(); ++l)
    if ((rc = amr_fill_level(s, l, false, 0.0))) return rc;
  for (int l = 0; l < g.num_levels(); ++l)
    if ((rc = amr_level_step(s, l, dt, salt, 0))) return rc;
  return LGD_OK;
}

// Subcycle::advance (src/integrators/subcycling.rs:195-267) as a launch sequence:
// fill the level's ghosts (coarse source blended in time between the parent's
// previous and current values), step it, recurse n = round(dt_l/dt_{l+1}) times
// into the finer level at t + s·dt_child (fma), then restrict it back.
// Between two substeps of a level that has a finer level, the restriction that
// ends one substep (defer_out) and the fill that starts the next (fuse_in) share a
// launch; the last substep restricts on its own.  The child's first fill shares the
// level's stage launch where that streams (filled_in: the caller's launch wrote
// this level's ghosts).  A finest child whose restriction shares its last stage
// launch restricts there, and with defer_out also fills this level's ghosts for
// its next substep (at next_alpha): *carry tells the caller which (1 restricted,
// 2 restricted and filled; 0 neither).
int amr_advance(lgd_sim *s, int level, double t, double dt, double alpha, bool fuse_in = false,
                bool defer_out = false, bool filled_in = false, double next_alpha = -1.0,
                int *carry = nullptr) {
  if (carry) *carry = 0;
  const lgd_amr &g = *s->amr;
  int rc;
  if (fuse_in) {
    if ((rc = amr_restrict_fill_level(s, level, alpha))) return rc;
  } else if (!filled_in && (rc = amr_fill_level(s, level, true, alpha))) {
    return rc;
  }
  uint64_t salt;
  std::memcpy(&salt, &t, sizeof salt);
  const bool child_fill = s->kind == lgd_sim::kModelC && lgd_amr_child_fill_fuses(s->ctx, g, level);
  if ((rc = amr_level_step(s, level, dt, salt, level, -1.0, child_fill))) return rc;
  const int child = level + 1;
  if (child < g.num_levels()) {
    double n = std::fmax(std::round(amr_stable_dt(s, level) / amr_stable_dt(s, child)), 1.0);
    uint64_t nn = (uint64_t)n;
    double dt_child = dt / (double)nn;
    // The finest level's substeps after the first take their ghosts from the stage
    // launch before them (one fill and nn stage launches instead of nn of each).
    const bool fused = child + 1 == g.num_levels() && !s->ctx->amr_per_block &&
                       !s->ctx->amr_split_fill && nn > 1;
    // A child with a finer level: its restriction and its next fill share launches.
    const bool chain = child + 1 < g.num_levels() && !s->ctx->amr_per_block &&
                       !s->ctx->amr_split_fill && nn > 1;
    // The finest child's last substep restricts into this level in its stage launch.
    const bool fuse_restrict =
        fused && s->kind == lgd_sim::kModelC && lgd_amr_restrict_fuses(s->ctx, g, child);
    int from_child = 0;  // what the child's last substep did for the next one
    for (uint64_t k = 0; k < nn; ++k) {
      if (s->capturing) {
        time_substep_kernel<<<1, 1, 0, s->ctx->stream>>>(s->d_t, child, level, (double)k, dt_child);
        if ((rc = launch_check(s->ctx, "substep time"))) return rc;
      }
      const double tk = std::fma((double)k, dt_child, t), ak = (double)k / (double)nn;
      if (!fused) {
        const double na = k + 1 < nn ? (double)(k + 1) / (double)nn : -1.0;
        const int prev = from_child;
        if ((rc = amr_advance(s, child, tk, dt_child, ak, chain && k > 0 && prev == 0,
                              chain && k + 1 < nn, (child_fill && k == 0) || prev == 2, na,
                              &from_child)))
          return rc;
        continue;
      }
      // Subcycle::advance of the finest level (no recursion below it).
      if (k == 0 && !child_fill && (rc = amr_fill_level(s, child, true, ak))) return rc;
      uint64_t salt_k;
      std::memcpy(&salt_k, &tk, sizeof salt_k);
      const double next = k + 1 < nn ? (double)(k + 1) / (double)nn : -1.0;
      if (fuse_restrict && k + 1 == nn) {
        const int ps = defer_out && next_alpha >= 0.0 && level >= 1 ? 2 : 1;
        if ((rc = amr_level_step(s, child, dt_child, salt_k, child, -1.0, false, ps, next_alpha)))
          return rc;
        if (carry) *carry = ps;
        continue;
      }
      if ((rc = amr_level_step(s, child, dt_child, salt_k, child, next))) return rc;
    }
    if (!defer_out && !fuse_restrict && (rc = amr_restrict_level(s, child))) return rc;
  }
  return LGD_OK;
}

double seconds_since(std::chrono::steady_clock::time_point &t0) {
  const auto t1 = std::chrono::steady_clock::now();
  const double s = std::chrono::duration<double>(t1 - t0).count();
  t0 = t1;
  return s;
}

// GradientTagger::prepare: restrict + fill the tagged field (adapt.rs:72-76).
int gradient_prepare(lgd_sim *s) {
  const lgd_amr &g = *s->amr;
  double *f = s->state[(size_t)s->policy.tag_field]->data;
  int rc;
  for (int l = g.num_levels() - 1; l >= 1; --l)
    if ((rc = lgd_amr_restrict(s->ctx, g, f, l))) return rc;
  for (int l = 0; l < g.num_levels(); ++l)
    if ((rc = lgd_amr_fill(s->ctx, g, f, nullptr, 0.0, l))) return rc;
  return LGD_OK;
}

// BergerOliger::regrid at the top of a step (src/core/simulation.rs:255-270).
int amr_regrid(lgd_sim *s) {
  int rc;
  if ((rc = amr_consolidate(s))) return rc;
  lgd_amr &g = *s->amr;
  const int tf = s->policy.tag_field;
  auto clock = std::chrono::steady_clock::now();
  ++s->regrids;
  // The tags come from a caller's TagCells, or the GradientTagger tags and dilates
  // on the device (bitmasks come to the host), or the field comes to the host and
  // is tagged there (the "regrid_host" knob, which also migrates on the host).
  const bool dev = !s->ctx->regrid_host && lgd_regrid_device_supported(g);
  std::vector<double> tag;
  std::vector<std::vector<lgd_box>> levels;
  int changed;
  if (s->tagger) {
    // A caller's TagCells: the trait's default prepare (nothing), then tag_level
    // per existing source level, finest first (adapt.rs:165-175).
    (void)hipStreamSynchronize(s->ctx->stream);
    const int capacity = g.max_levels() - 1;
    std::vector<std::vector<uint64_t>> bits((size_t)g.num_levels());
    for (int src = capacity - 1; src >= 0; --src) {
      if (src >= g.num_levels()) continue;
      lgd_tag_sink sink;
      if (s->tagger(s->tagger_user, &g, src, &sink) != 0)
        return fail(s->ctx, LGD_ERR_INVALID_ARGUMENT, "the tagger failed");
      lgd_regrid_bits_from_cells(g, src, sink.cells, s->policy.buffer, bits[(size_t)src]);
    }
    s->regrid_seconds[0] += seconds_since(clock);
    std::vector<uint64_t *> ptr(bits.size(), nullptr);
    for (size_t l = 0; l < bits.size(); ++l)
      if (!bits[l].empty()) ptr[l] = bits[l].data();
    changed = lgd_regrid_propose_bits(g, ptr, s->policy, levels);
  } else if (dev) {
    if ((rc = gradient_prepare(s))) return rc;
    std::vector<lgd_flag_words> flags;
    if ((rc = lgd_regrid_flag_words(s->ctx, g, s->state[(size_t)tf]->data, s->policy, s->rscratch,
                                    flags)))
      return rc;
    s->regrid_seconds[0] += seconds_since(clock);
    changed = lgd_regrid_propose_words(g, flags, s->policy, levels);
  } else {
    if ((rc = gradient_prepare(s))) return rc;
    tag.resize((size_t)g.off.back());
    if ((rc = lgd_buf_download(s->ctx, s->state[(size_t)tf], 0, tag.data(), tag.size()))) return rc;
    s->regrid_seconds[0] += seconds_since(clock);
    changed = lgd_regrid_propose(g, tag, s->policy, levels);
  }
  s->regrid_seconds[1] += seconds_since(clock);
  if (changed == 0) return LGD_OK;
  lgd_amr *ng = new (std::nothrow) lgd_amr();
  if (!ng) return fail(s->ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  if ((rc = lgd_amr_build(s->ctx, g.base, g.ratios, levels, *ng, &g))) {
    ng->free_device();
    delete ng;
    return rc;
  }
  s->regrid_seconds[2] += seconds_since(clock);
  // The new state goes into the spare buffers (next[]) when they hold the new
  // length, else into fresh ones.
  const size_t len = (size_t)ng->off.back();
  std::vector<lgd_buf *> ns, fresh;
  std::vector<double> old_f, new_f;
  const int64_t fine0 = ng->off[ng->level_start[1]];  // level-0 elements (same in both)
  // A failure from here on leaves the old hierarchy and state untouched.
  auto abandon = [&](int code) {
    (void)hipStreamSynchronize(s->ctx->stream);
    for (auto *b : fresh) lgd_buf_free(b);
    ng->free_device();
    delete ng;
    return code;
  };
  for (size_t f = 0; f < s->state.size(); ++f) {
    lgd_buf *b = s->next.size() > f ? s->next[f] : nullptr;
    if (!b || b->cap < len) {
      b = nullptr;
      if ((rc = lgd_buf_alloc(s->ctx, len + len / 4, &b))) return abandon(rc);
      fresh.push_back(b);
    }
    ns.push_back(b);
    hipError_t e = hipMemcpyAsync(b->data, s->state[f]->data, (size_t)fine0 * sizeof(double),
                                  hipMemcpyDeviceToDevice, s->ctx->stream);
    if (e != hipSuccess) return abandon(hip_fail(s->ctx, e, "migrate level 0"));
    if (dev) continue;
    if ((int)f == tf && !tag.empty()) {
      old_f.swap(tag);  // already downloaded
    } else {
      old_f.resize((size_t)g.off.back());
      if ((rc = lgd_buf_download(s->ctx, s->state[f], 0, old_f.data(), old_f.size())))
        return abandon(rc);
    }
    lgd_regrid_migrate(g, old_f, *ng, new_f);
    if (!new_f.empty()) {
      e = hipMemcpy(b->data + fine0, new_f.data(), new_f.size() * sizeof(double),
                    hipMemcpyHostToDevice);
      if (e != hipSuccess) return abandon(hip_fail(s->ctx, e, "migrate upload"));
    }
  }
  if (dev) {
    std::vector<const double *> of;
    std::vector<double *> nfp;
    for (size_t f = 0; f < s->state.size(); ++f) {
      of.push_back(s->state[f]->data);
      nfp.push_back(ns[f]->data);
    }
    if ((rc = lgd_regrid_migrate_device(s->ctx, g, of.data(), *ng, nfp.data(), (int)of.size())))
      return abandon(rc);
  }
  int32_t bad = 0;
  if ((rc = lgd_amr_plan_error(s->ctx, *ng, &bad))) return abandon(rc);
  {
    hipError_t e = hipStreamSynchronize(s->ctx->stream);
    if (e != hipSuccess) return abandon(hip_fail(s->ctx, e, "regrid"));
  }
  if (bad)
    return abandon(fail(s->ctx, LGD_ERR_INVALID_ARGUMENT,
                        "regrid proposal has a ghost cell without a donor"));
  drop_graphs(s);  // they hold the old hierarchy's plans and buffers
  ++s->amr_epoch;
  // The old state buffers become the spares; spares that went unused are freed.
  std::vector<lgd_buf *> spare = s->state;
  for (size_t f = 0; f < s->next.size(); ++f)
    if (s->next[f] != ns[f]) lgd_buf_free(s->next[f]);
  for (auto *b : ns) b->len = len;
  s->state = ns;
  s->next = spare;
  lgd_amr_take_level0_plan(*ng, *s->amr);
  s->amr->free_device();
  delete s->amr;
  s->amr = ng;
  rc = alloc_amr_stages(s);
  s->regrid_seconds[3] += seconds_since(clock);
  return rc;
}

// One step after any regrid due at its top.
int step_body(lgd_sim *s, double dt, uint64_t salt) {
  if (s->amr) {
    if (s->subcycle) return amr_advance(s, 0, s->t, dt, 0.0);
    return amr_global_step(s, dt, salt);
  }
  return s->kind == lgd_sim::kModelC ? step_modelc(s, dt, salt) : step_heat(s, dt);
}

bool regrid_due(const lgd_sim *s) {
  return s->amr && s->adaptive && s->step_index % s->policy.every == 0;
}

// Host-side effects of one step: the fused ping-pong swap and the time accumulation
// (Simulation::step, src/core/simulation.rs:281-289).
void step_done(lgd_sim *s, double dt) {
  if (!s->amr && s->mode == LGD_MODE_FUSED && s->integrator == LGD_INTEGRATOR_EULER)
    std::swap(s->state, s->next);
  s->t = s->t + dt;
  s->step_index += 1;
}

// Capture `n` steps from the current buffers into *out.  The host bookkeeping the
// capture pass performs (buffer swaps) is undone: the graph has not run yet.
uint64_t parity_bits(const lgd_sim *s) {
  uint64_t m = 0;
  for (size_t l = 0; l < s->cur.size() && l < 64; ++l)
    if (s->cur[l]) m |= 1ull << l;
  return m;
}

void set_parity_bits(lgd_sim *s, uint64_t m) {
  for (size_t l = 0; l < s->cur.size() && l < 64; ++l) s->cur[l] = (int8_t)((m >> l) & 1u);
}

int capture(lgd_sim *s, double dt, int64_t n, hipGraphExec_t *out, uint64_t *cur_after) {
  lgd_ctx *ctx = s->ctx;
  const auto state = s->state, next = s->next;
  const auto cur = s->cur;
  const double t = s->t;
  const int64_t step_index = s->step_index;
  hipError_t e = hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal);
  if (e != hipSuccess) return hip_fail(ctx, e, "graph capture (needs a non-default stream)");
  s->capturing = true;
  int rc = LGD_OK;
  for (int64_t k = 0; k < n && !rc; ++k) {
    uint64_t salt;
    std::memcpy(&salt, &s->t, sizeof salt);
    rc = step_body(s, dt, salt);
    if (!rc) {
      time_advance_kernel<<<1, 1, 0, ctx->stream>>>(s->d_t, dt);
      rc = launch_check(ctx, "time advance");
    }
    if (!rc) step_done(s, dt);
  }
  s->capturing = false;
  ctx->salt_src = nullptr;
  hipGraph_t graph = nullptr;
  e = hipStreamEndCapture(ctx->stream, &graph);
  *cur_after = parity_bits(s);
  s->cur = cur;
  s->state = state;
  s->next = next;
  s->t = t;
  s->step_index = step_index;
  if (rc) {
    if (graph) (void)hipGraphDestroy(graph);
    return rc;
  }
  if (e != hipSuccess) return hip_fail(ctx, e, "graph capture");
  e = hipGraphInstantiate(out, graph, nullptr, nullptr, 0);
  (void)hipGraphDestroy(graph);
  if (e != hipSuccess) return hip_fail(ctx, e, "graph instantiate");
  return LGD_OK;
}

// Run `n` steps (no regrid due among steps 2..n) as one graph launch.
int graph_steps(lgd_sim *s, double dt, int64_t n) {
  lgd_ctx *ctx = s->ctx;
  uint64_t dt_bits;
  std::memcpy(&dt_bits, &dt, sizeof dt_bits);
  lgd_sim::graph *hit = nullptr;
  const uint64_t cur0 = parity_bits(s);
  for (auto &g : s->graphs)
    if (g.state0 == s->state[0] && g.cur0 == cur0 && g.epoch == s->amr_epoch &&
        g.tuning == ctx->tuning_epoch && g.dt_bits == dt_bits && g.nsteps == n)
      hit = &g;
  int rc;
  if (!hit) {
    // Keep only graphs of this key (the other ping-pong parity).
    std::vector<lgd_sim::graph> keep;
    if (!s->graphs.empty()) (void)hipStreamSynchronize(ctx->stream);
    for (auto &g : s->graphs) {
      if (g.epoch == s->amr_epoch && g.tuning == ctx->tuning_epoch && g.dt_bits == dt_bits &&
          g.nsteps == n)
        keep.push_back(g);
      else
        (void)hipGraphExecDestroy(g.exec);
    }
    s->graphs = keep;
    lgd_sim::graph g{s->state[0], cur0, 0, s->amr_epoch, ctx->tuning_epoch, dt_bits, n, nullptr};
    if ((rc = capture(s, dt, n, &g.exec, &g.cur1))) return rc;
    s->graphs.push_back(g);
    hit = &s->graphs.back();
  }
  uint64_t tb, kb;
  std::memcpy(&tb, &s->t, sizeof tb);
  std::memcpy(&kb, &s->d_t_value, sizeof kb);
  if (!s->d_t_known || tb != kb) {
    time_set_kernel<<<1, 1, 0, ctx->stream>>>(s->d_t, s->t);
    if ((rc = launch_check(ctx, "time set"))) return rc;
  }
  hipError_t e = hipGraphLaunch(hit->exec, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "graph launch");
  set_parity_bits(s, hit->cur1);
  for (int64_t k = 0; k < n; ++k) step_done(s, dt);
  s->d_t_known = true;
  s->d_t_value = s->t;
  return LGD_OK;
}

// Simulation::step's observer loop (simulation.rs:286-289): every attached observer
// sees the completed step; only snapshot steps touch the device.
int notify_observers(lgd_sim *s) {
  int rc;
  for (auto *o : s->observers) {
    if (!observer_due(o, (uint64_t)s->step_index)) continue;
    std::vector<const lgd_buf *> f(observer_fields(o).size());
    if (s->amr) {
      if ((rc = amr_consolidate(s))) return rc;
      for (size_t c = 0; c < f.size(); ++c) f[c] = s->state[(size_t)observer_fields(o)[c]];
      rc = lgd_observer_observe_amr(o, s->amr, (uint64_t)s->step_index, s->t,
                                    (uint64_t)s->amr_epoch, f.data());
    } else {
      std::vector<uint32_t> fg(f.size(), 1u);
      for (size_t c = 0; c < f.size(); ++c) f[c] = s->state[(size_t)observer_fields(o)[c]];
      const size_t ns = observer_statics(o).size();
      if (ns && !s->theta0)
        return fail(s->ctx, LGD_ERR_INVALID_ARGUMENT, "theta0 is observed but no orientation is set");
      std::vector<const lgd_buf *> st(ns, s->theta0);
      std::vector<uint32_t> sg(ns, 0u);
      rc = lgd_observer_observe(o, s->grid, (uint64_t)s->step_index, s->t, 0, f.data(), fg.data(),
                                st.data(), sg.data());
    }
    if (rc) return rc;
  }
  return LGD_OK;
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
  if (bcs) {
    for (int d = 0; d < grid->cart.dim; ++d)
      for (int e = 0; e < 2; ++e) {
        const int k = bcs[2 * d + e].kind;
        if (k < LGD_BC_MIRROR || k > LGD_BC_EXTRAPOLATE ||
            (k == LGD_BC_EXTRAPOLATE && grid->cart.n[d] < 2)) {
          delete s;
          return fail(ctx, LGD_ERR_INVALID_ARGUMENT,
                      "unknown FaceBc kind, or extrapolation across fewer than two cells");
        }
      }
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
  if (sim->ctx && sim->ctx->stream) (void)hipStreamSynchronize(sim->ctx->stream);
  drop_graphs(sim);
  if (sim->d_t) (void)hipFree(sim->d_t);
  lgd_regrid_scratch_free(sim->rscratch);
  free_all(sim);
  if (sim->amr) lgd_amr_destroy(sim->amr);
  delete sim;
  return LGD_OK;
}

int lgd_sim_field(lgd_sim *sim, int32_t index, lgd_buf **out) {
  if (!sim || !out) return LGD_ERR_INVALID_ARGUMENT;
  if (index < 0 || index >= (int32_t)sim->state.size())
    return fail(sim->ctx, LGD_ERR_INVALID_ARGUMENT, "field index out of range");
  int rc;
  if (sim->amr && (rc = amr_consolidate(sim))) return rc;
  *out = sim->state[index];
  return LGD_OK;
}

int lgd_sim_step(lgd_sim *sim, double dt, int64_t nsteps) {
  if (!sim) return LGD_ERR_INVALID_ARGUMENT;
  int rc;
  for (int64_t n = 0; n < nsteps;) {
    // BergerOliger::regrid at the top of a step (src/core/simulation.rs:255-270).
    if (regrid_due(sim) && (rc = amr_regrid(sim))) return rc;
    if (sim->graph_steps > 0) {
      int64_t k = std::min(sim->graph_steps, nsteps - n);
      if (sim->amr && sim->adaptive)  // no regrid inside a graph
        k = std::min(k, sim->policy.every - sim->step_index % sim->policy.every);
      for (auto *o : sim->observers)  // a graph ends at each snapshot step
        k = std::min(k, observer_steps_to_due(o, (uint64_t)sim->step_index));
      if ((rc = graph_steps(sim, dt, k))) return rc;
      n += k;
      if ((rc = notify_observers(sim))) return rc;
      continue;
    }
    uint64_t salt;
    std::memcpy(&salt, &sim->t, sizeof salt);  // t.to_bits()
    if ((rc = step_body(sim, dt, salt))) return rc;
    step_done(sim, dt);
    ++n;
    if ((rc = notify_observers(sim))) return rc;
  }
  return LGD_OK;
}

int lgd_sim_detach_observer(lgd_sim *sim, lgd_observer *obs) {
  if (!sim || !obs) return fail(sim ? sim->ctx : nullptr, LGD_ERR_INVALID_ARGUMENT, "null argument");
  auto &v = sim->observers;
  for (size_t i = 0; i < v.size(); ++i)
    if (v[i] == obs) {
      v.erase(v.begin() + (long)i);
      return LGD_OK;
    }
  return fail(sim->ctx, LGD_ERR_INVALID_ARGUMENT, "the observer is not attached to this simulation");
}

int lgd_sim_attach_observer(lgd_sim *sim, lgd_observer *obs) {
  if (!sim || !obs) return fail(sim ? sim->ctx : nullptr, LGD_ERR_INVALID_ARGUMENT, "null argument");
  for (int32_t f : observer_fields(obs))
    if (f < 0 || f >= (int32_t)sim->state.size())
      return fail(sim->ctx, LGD_ERR_INVALID_ARGUMENT, "observed field index out of range");
  for (int32_t f : observer_statics(obs)) {
    if (f != 0 || sim->kind != lgd_sim::kModelC || sim->amr)
      return fail(sim->ctx, LGD_ERR_INVALID_ARGUMENT,
                  "the only static field is theta0 (field 0) of a Cartesian Model C");
  }
  // Size the observer's buffers for this grid now: snapshots then allocate nothing.
  int rc = observer_prepare(obs, sim->amr ? nullptr : sim->grid, sim->amr,
                            sim->amr ? (uint64_t)sim->amr_epoch : 0);
  if (rc) return rc;
  for (auto *o : sim->observers)
    if (o == obs) return LGD_OK;  // attached once
  sim->observers.push_back(obs);
  return LGD_OK;
}

int lgd_sim_set_graph_steps(lgd_sim *sim, int64_t steps) {
  if (!sim) return LGD_ERR_INVALID_ARGUMENT;
  if (steps < 0) return fail(sim->ctx, LGD_ERR_INVALID_ARGUMENT, "graph steps are non-negative");
  drop_graphs(sim);
  sim->graph_steps = steps;
  if (steps > 0 && !sim->d_t) {
    const size_t n = (size_t)(sim->amr ? sim->amr->max_levels() + 1 : 1);
    hipError_t e = hipMalloc((void **)&sim->d_t, n * sizeof(double));
    if (e != hipSuccess) {
      sim->d_t = nullptr;
      sim->graph_steps = 0;
      return hip_fail(sim->ctx, e, "graph time slots");
    }
    sim->d_t_known = false;
  }
  return LGD_OK;
}

double lgd_sim_time(const lgd_sim *sim) { return sim ? sim->t : 0.0; }

int lgd_sim_regrid_stats(const lgd_sim *sim, int64_t *regrids, double *seconds) {
  if (!sim) return LGD_ERR_INVALID_ARGUMENT;
  if (regrids) *regrids = sim->regrids;
  if (seconds)
    for (int i = 0; i < 4; ++i) seconds[i] = sim->regrid_seconds[i];
  return LGD_OK;
}

int64_t lgd_sim_step_index(const lgd_sim *sim) { return sim ? sim->step_index : -1; }

int lgd_sim_autotune_amr(lgd_sim *sim, double dt, int32_t reps, lgd_amr_tuning_result *out) {
  if (!sim || !out || reps < 1) return LGD_ERR_INVALID_ARGUMENT;
  if (!sim->amr) return fail(sim->ctx, LGD_ERR_INVALID_ARGUMENT, "not an AMR simulation");
  lgd_ctx *ctx = sim->ctx;
  const lgd_amr &g = *sim->amr;
  const int L = g.num_levels(), nf = nfields(sim);
  const bool modelc = sim->kind == lgd_sim::kModelC;
  int rc = LGD_OK;
  // Substeps of each level per coarse step (Subcycle::advance's n = round(dt_l/dt_{l+1})),
  // one each with a global dt.
  std::vector<double> N((size_t)L, 1.0);
  for (int l = 1; l < L; ++l)
    N[(size_t)l] = N[(size_t)l - 1] *
                   (sim->subcycle ? std::fmax(std::round(amr_stable_dt(sim, l - 1) / amr_stable_dt(sim, l)), 1.0)
                                  : 1.0);
  hipEvent_t e0 = nullptr, e1 = nullptr;
  hipError_t e = hipEventCreate(&e0);
  if (e == hipSuccess) e = hipEventCreate(&e1);
  if (e != hipSuccess) {
    if (e0) (void)hipEventDestroy(e0);
    return hip_fail(ctx, e, "autotune events");
  }
  // Mean ms of `reps` launches of f after one untimed launch.
  auto timed = [&](auto &&f, double *ms) -> int {
    int r = f();
    if (r) return r;
    hipError_t er = hipEventRecord(e0, ctx->stream);
    for (int i = 0; i < reps && !r; ++i) r = f();
    if (r) return r;
    if (er == hipSuccess) er = hipEventRecord(e1, ctx->stream);
    if (er == hipSuccess) er = hipEventSynchronize(e1);
    float t = 0.0f;
    if (er == hipSuccess) er = hipEventElapsedTime(&t, e0, e1);
    if (er != hipSuccess) return hip_fail(ctx, er, "autotune timing");
    *ms = (double)t / reps;
    return LGD_OK;
  };
  uint64_t salt;
  std::memcpy(&salt, &sim->t, sizeof salt);
  // Every timed launch writes only what the next coarse step overwrites before it
  // reads it: a level's other buffer (its stage output), the ghost entries of the
  // current one (fills), the parent's other buffer (restrictions).
  // A level's stage as a coarse step issues it: with the next substep's ghosts
  // (next_fill, the finest level) or with the finer level's first fill (child_fill,
  // which writes only ghosts of that level's current buffers).
  auto stage = [&](int l, bool next_fill = false, bool child_fill = false) {
    lgd_amr_next_fill nfa{};
    lgd_amr_child_fill cfa{};
    for (int f = 0; f < nf; ++f) {
      if (next_fill) {
        nfa.par[f] = amr_cur(sim, f, l - 1)->data;
        nfa.old[f] = amr_prev(sim, f, l - 1)->data;
      }
      if (child_fill) cfa.lvl[f] = amr_cur(sim, f, l + 1)->data;
    }
    nfa.alpha = 0.5;
    if (modelc)
      return lgd_amr_modelc_step(ctx, g, sim->params, l, dt, amr_cur(sim, 0, l)->data,
                                 amr_cur(sim, 1, l)->data, amr_prev(sim, 0, l)->data,
                                 amr_prev(sim, 1, l)->data, sim->seed, salt, sim->wiener != 0,
                                 next_fill ? &nfa : nullptr, child_fill ? &cfa : nullptr);
    return lgd_amr_heat_step(ctx, g, sim->kappa, l, dt, amr_cur(sim, 0, l)->data,
                             amr_prev(sim, 0, l)->data, next_fill ? &nfa : nullptr);
  };
  auto fill = [&](int l) { return amr_fill_level(sim, l, l > 0, 0.5); };
  auto restrict_prev = [&](int l) {
    const double *fine[2] = {nullptr, nullptr};
    double *coarse[2] = {nullptr, nullptr};
    for (int f = 0; f < nf; ++f) {
      fine[f] = amr_cur(sim, f, l)->data;
      coarse[f] = amr_prev(sim, f, l - 1)->data;
    }
    return lgd_amr_restrict_pp(ctx, g, l, nf, fine, coarse);
  };
  const bool streams0 = modelc && lgd_amr_level0_streams(ctx, g);
  const int64_t k0 = ctx->modelc_kernel, s0 = ctx->modelc_strips, p0 = ctx->modelc_prefetch;
  const int64_t t0 = ctx->amr_tile, w0 = ctx->amr_wg, ak0 = ctx->amr_kernel, r0 = ctx->amr_rows;
  int32_t candidates = 0;
  // Level stages (the levels not on the Cartesian streaming kernel): the streaming
  // stage at several strip heights (Model C), then the per-cell stage's tile shapes.
  struct StageShape {
    int kernel, rows, tile;
  };
  std::vector<StageShape> stage_cand;
  if (modelc)
    for (int rows : {0, 2, 3, 4, 6, 8, 12, 16}) stage_cand.push_back({2, rows, 0});
  for (int sh = 0; sh <= 6; ++sh) stage_cand.push_back({1, 0, sh});
  double best_tiles = 1e300, dflt_tiles = 0.0;
  StageShape best_stage = stage_cand.front();
  for (size_t c = 0; c < stage_cand.size() && !rc; ++c) {
    const StageShape &sc = stage_cand[c];
    ctx->amr_kernel = sc.kernel;
    ctx->amr_rows = sc.rows;
    ctx->amr_tile = sc.tile;
    double tot = 0.0;
    for (int l = streams0 ? 1 : 0; l < L && !rc; ++l) {
      double ms = 0.0;
      const bool batched = !ctx->amr_per_block && !ctx->amr_split_fill;
      if (l + 1 < L) {
        // A level with a finer one: its stage, and the finer level's first fill in
        // the same launch or in one of its own.
        const bool cf = modelc && lgd_amr_child_fill_fuses(ctx, g, l);
        rc = timed([&] { return stage(l, false, cf); }, &ms);
        tot += N[(size_t)l] * ms;
        if (!cf && !rc) {
          rc = timed([&] { return fill(l + 1); }, &ms);
          tot += N[(size_t)l] * ms;
        }
      } else if (l > 0 && batched && N[(size_t)l] > N[(size_t)l - 1]) {
        // The finest level: all but the last substep of each parent substep also
        // write the next substep's ghosts; the last one restricts into the parent
        // (and fills the parent's next ghosts) in its own launch or before the
        // parent's next fill.  Into the parent's other buffers, which its next stage
        // overwrites.
        rc = timed([&] { return stage(l, true); }, &ms);
        tot += (N[(size_t)l] - N[(size_t)l - 1]) * ms;
        if (!rc && modelc && lgd_amr_restrict_fuses(ctx, g, l)) {
          lgd_amr_parent_step ps{};
          ps.fill = l >= 2;
          ps.alpha = 0.5;
          for (int f = 0; f < nf; ++f) {
            ps.coarse[f] = amr_prev(sim, f, l - 1)->data;
            if (ps.fill) {
              ps.par[f] = amr_cur(sim, f, l - 2)->data;
              ps.old[f] = amr_prev(sim, f, l - 2)->data;
            }
          }
          rc = timed([&] {
            return lgd_amr_modelc_step(ctx, g, sim->params, l, dt, amr_cur(sim, 0, l)->data,
                                       amr_cur(sim, 1, l)->data, amr_prev(sim, 0, l)->data,
                                       amr_prev(sim, 1, l)->data, sim->seed, salt, sim->wiener != 0,
                                       nullptr, nullptr, &ps);
          }, &ms);
          tot += N[(size_t)l - 1] * ms;
        } else if (!rc) {
          rc = timed([&] { return stage(l); }, &ms);
          tot += N[(size_t)l - 1] * ms;
          if (!rc && l >= 2) {
            double *lv[2] = {nullptr, nullptr};
            const double *par[2] = {nullptr, nullptr}, *old[2] = {nullptr, nullptr}, *fine[2] = {nullptr, nullptr};
            for (int f = 0; f < nf; ++f) {
              lv[f] = amr_prev(sim, f, l - 1)->data;
              par[f] = amr_cur(sim, f, l - 2)->data;
              old[f] = amr_prev(sim, f, l - 2)->data;
              fine[f] = amr_cur(sim, f, l)->data;
            }
            rc = timed([&] { return lgd_amr_restrict_fill_pp(ctx, g, l - 1, nf, lv, par, old, 0.5, fine); },
                       &ms);
            tot += N[(size_t)l - 1] * ms;
          } else if (!rc) {
            rc = timed([&] { return restrict_prev(l); }, &ms);
            tot += N[(size_t)l - 1] * ms;
          }
        }
      } else {
        rc = timed([&] { return stage(l); }, &ms);
        tot += N[(size_t)l] * ms;
      }
    }
    ++candidates;
    if (c == 0) dflt_tiles = tot;
    if (tot < best_tiles) {
      best_tiles = tot;
      best_stage = sc;
    }
  }
  ctx->amr_kernel = ak0;
  ctx->amr_rows = r0;
  ctx->amr_tile = t0;
  // Fill and restriction widths.
  double best_wg_ms = 1e300, dflt_wg = 0.0;
  int best_wg = 256;
  const int widths[] = {64, 128, 256, 512, 1024};
  for (int w : widths) {
    if (rc) break;
    ctx->amr_wg = w;
    double tot = 0.0;
    for (int l = 0; l < L && !rc; ++l) {
      double ms = 0.0;
      if (!(l == 0 && streams0)) {
        rc = timed([&] { return fill(l); }, &ms);
        tot += N[(size_t)l] * ms;
      }
      if (l >= 1 && !rc) {
        rc = timed([&] { return restrict_prev(l); }, &ms);
        tot += N[(size_t)l - 1] * ms;
      }
    }
    ++candidates;
    if (w == 256) dflt_wg = tot;
    if (tot < best_wg_ms) {
      best_wg_ms = tot;
      best_wg = w;
    }
  }
  // Level 0 on the streaming kernel: one block or two per wavefront, strips and rows
  // loaded ahead.
  const bool noise0 = modelc && sim->wiener != 0 && sim->params.noise_amplitude != 0.0;
  lgd_launch_shape best0 = lgd_amr_level0_default_shape(g, noise0);
  double best0_ms = 0.0, dflt0_ms = 0.0;
  if (streams0 && !rc) {
    const lgd_launch_shape d = lgd_amr_level0_default_shape(g, noise0);
    const bool pairable = lgd_stream_pairable(g.base.n[0], g.base.nb[0]);
    best0_ms = 1e300;
    const int strips[] = {1, 2, 3, 4, 6, 8};
    for (int kern = 2; kern <= (pairable ? 3 : 2); ++kern)
      for (int st : strips)
        for (int pf = 1; pf <= 2 && !rc; ++pf) {
          ctx->modelc_kernel = kern;
          ctx->modelc_strips = st;
          ctx->modelc_prefetch = pf;
          double ms = 0.0;
          rc = timed([&] { return stage(0); }, &ms);
          ++candidates;
          if (kern == d.kernel && st == d.strips && pf == d.prefetch) dflt0_ms = ms;
          if (ms < best0_ms) {
            best0_ms = ms;
            best0 = lgd_launch_shape{kern, st, pf};
          }
        }
  }
  ctx->modelc_kernel = k0;
  ctx->modelc_strips = s0;
  ctx->modelc_prefetch = p0;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (rc) {
    ctx->amr_tile = t0;
    ctx->amr_wg = w0;
    ctx->amr_kernel = ak0;
    ctx->amr_rows = r0;
    return rc;
  }
  ctx->amr_kernel = best_stage.kernel == 1 ? 1 : 0;
  ctx->amr_rows = best_stage.kernel == 1 ? 0 : best_stage.rows;
  ctx->amr_tile = best_stage.kernel == 1 ? best_stage.tile : 0;
  ctx->amr_wg = best_wg;
  if (streams0)
    ctx->tuned[std::make_tuple(g.base.n[0], g.base.n[1], (int64_t)g.level_start[1], noise0 ? 1 : 0)] =
        best0;
  ++ctx->tuning_epoch;  // captured graphs hold the old shapes
  out->stage_rows = 0;
  if (best_stage.kernel == 1) {
    out->tile_x = kAmrTileShapes[best_stage.tile][0];
    out->tile_y = kAmrTileShapes[best_stage.tile][1];
  } else {
    out->tile_x = out->tile_y = 0;
    // The strip height the finest streamed level runs at.
    out->stage_rows = best_stage.rows > 0 ? best_stage.rows : (L > 1 ? g.plans[(size_t)L - 1].auto_rows : 0);
  }
  out->level0_kernel = streams0 ? best0.kernel : 0;
  out->wg = best_wg;
  out->level0_strips = streams0 ? best0.strips : 0;
  out->level0_prefetch = streams0 ? best0.prefetch : 0;
  out->candidates = candidates;
  out->ms = best_tiles + best_wg_ms + best0_ms;
  out->default_ms = dflt_tiles + dflt_wg + dflt0_ms;
  return LGD_OK;
}

int lgd_sim_set_tagger(lgd_sim *sim, lgd_tag_level_fn fn, void *user) {
  if (!sim) return LGD_ERR_INVALID_ARGUMENT;
  if (!sim->amr || !sim->adaptive)
    return fail(sim->ctx, LGD_ERR_INVALID_ARGUMENT, "a tagger needs an adaptive AMR simulation");
  sim->tagger = fn;
  sim->tagger_user = fn ? user : nullptr;
  return LGD_OK;
}

int lgd_tag_sink_push(lgd_tag_sink *sink, const int64_t *cells, size_t ncells) {
  if (!sink || (!cells && ncells)) return LGD_ERR_INVALID_ARGUMENT;
  sink->cells.insert(sink->cells.end(), cells, cells + 2 * ncells);
  return LGD_OK;
}

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

int lgd_sim_set_integrator(lgd_sim *sim, int32_t integrator) {
  if (!sim) return LGD_ERR_INVALID_ARGUMENT;
  if (integrator != LGD_INTEGRATOR_EULER && integrator != LGD_INTEGRATOR_RK4)
    return fail(sim->ctx, LGD_ERR_INVALID_ARGUMENT, "unknown integrator");
  if (integrator == sim->integrator) return LGD_OK;
  if (integrator == LGD_INTEGRATOR_RK4) {
    // RungeKutta4 implements Integrator<G, D, NoNoise> only (rk4.rs:23).
    if (sim->amr) return fail(sim->ctx, LGD_ERR_UNSUPPORTED, "RK4 runs on Cartesian simulations");
    if (sim->kind == lgd_sim::kModelC && sim->wiener)
      return fail(sim->ctx, LGD_ERR_UNSUPPORTED, "RK4 pairs with deterministic models (no Wiener driver)");
    const size_t len = sim->state[0]->len;
    // Fused Model C stages: the running sum and two stage inputs (φ, u each); the
    // reference's sweeps: k1..k4 and y + c·k per field.
    const size_t need = (sim->kind == lgd_sim::kModelC && sim->mode == LGD_MODE_FUSED)
                            ? 6
                            : 5 * sim->state.size();
    while (sim->rk.size() < need) {
      lgd_buf *b = nullptr;
      int rc = lgd_buf_alloc(sim->ctx, len, &b);
      if (rc) return rc;
      sim->rk.push_back(b);
    }
    // Fused ping-pong state is not used by RK4: keep the current values in state[].
  }
  drop_graphs(sim);
  sim->integrator = integrator;
  return LGD_OK;
}

int lgd_sim_set_orientation(lgd_sim *sim, const double *theta0, size_t n) {
  if (!sim) return LGD_ERR_INVALID_ARGUMENT;
  if (sim->kind != lgd_sim::kModelC || sim->amr)
    return fail(sim->ctx, LGD_ERR_UNSUPPORTED, "orientations are a Cartesian Model C feature");
  const lgd_cart &c = sim->grid->cart;
  const size_t len = (size_t)(c.nblocks * c.n[0] * c.n[1]);
  if (theta0 && n != len) return fail(sim->ctx, LGD_ERR_SIZE_MISMATCH, "theta0 needs nblocks·nx·ny values");
  int rc;
  if (!sim->cos0) {
    if ((rc = lgd_buf_alloc(sim->ctx, len, &sim->cos0))) return rc;
    if ((rc = lgd_buf_alloc(sim->ctx, len, &sim->sin0))) return rc;
    if ((rc = lgd_buf_alloc(sim->ctx, len, &sim->theta0))) return rc;
  }
  if (theta0) {
    if ((rc = lgd_buf_upload(sim->ctx, sim->theta0, 0, theta0, len))) return rc;
  } else if ((rc = lgd_buf_fill_zero(sim->ctx, sim->theta0))) {
    return rc;
  }
  // (sin θ₀, cos θ₀) with libm, as the reference's per-cell sin_cos (phasefield.rs:382).
  std::vector<double> cs(len), sn(len);
  for (size_t i = 0; i < len; ++i) {
    const double th = theta0 ? theta0[i] : 0.0;
    sn[i] = std::sin(th);
    cs[i] = std::cos(th);
  }
  if ((rc = lgd_buf_upload(sim->ctx, sim->cos0, 0, cs.data(), len))) return rc;
  if ((rc = lgd_buf_upload(sim->ctx, sim->sin0, 0, sn.data(), len))) return rc;
  drop_graphs(sim);
  return LGD_OK;
}

int lgd_sim_initialize_grains_oriented(lgd_sim *sim, const double *grains, size_t ngrains,
                                       double u0) {
  if (!sim || (!grains && ngrains)) return LGD_ERR_INVALID_ARGUMENT;
  std::vector<double> g3(3 * ngrains);
  for (size_t g = 0; g < ngrains; ++g)
    for (int k = 0; k < 3; ++k) g3[3 * g + k] = grains[4 * g + k];
  int rc = lgd_sim_initialize_grains(sim, g3.data(), ngrains, u0);
  if (rc) return rc;
  // θ₀ by nearest-seed Voronoi (phasefield.rs:286-314): per block, the grains whose
  // closest approach beats every grain's farthest corner; the first nearest wins.
  const lgd_cart &c = sim->grid->cart;
  const int64_t nx = c.n[0], ny = c.n[1];
  std::vector<double> theta((size_t)(c.nblocks * nx * ny), 0.0);
  std::vector<size_t> cand;
  for (int64_t b = 0; b < c.nblocks; ++b) {
    double c0[2], c1[2];
    cell_center2(c, b, 0, 0, c0);
    cell_center2(c, b, nx - 1, ny - 1, c1);
    const double lo[2] = {std::fma(0.5, -c.h[0], c0[0]), std::fma(0.5, -c.h[1], c0[1])};
    const double hi[2] = {std::fma(0.5, c.h[0], c1[0]), std::fma(0.5, c.h[1], c1[1])};
    double threshold = std::numeric_limits<double>::max();
    for (size_t g = 0; g < ngrains; ++g) {
      const double cx = grains[4 * g], cy = grains[4 * g + 1];
      const double fx = std::fmax(std::fabs(cx - lo[0]), std::fabs(cx - hi[0]));
      const double fy = std::fmax(std::fabs(cy - lo[1]), std::fabs(cy - hi[1]));
      threshold = std::fmin(threshold, std::hypot(fx, fy));
    }
    cand.clear();
    for (size_t g = 0; g < ngrains; ++g) {
      const double cx = grains[4 * g], cy = grains[4 * g + 1];
      const double dx = std::fmax(std::fmax(lo[0] - cx, cx - hi[0]), 0.0);
      const double dy = std::fmax(std::fmax(lo[1] - cy, cy - hi[1]), 0.0);
      if (std::hypot(dx, dy) <= threshold) cand.push_back(g);
    }
    double *tb = theta.data() + b * nx * ny;
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        double xy[2];
        cell_center2(c, b, i, j, xy);
        double best = 0.0, bestd = 0.0;
        bool have = false;
        for (size_t g : cand) {
          const double dx = xy[0] - grains[4 * g], dy = xy[1] - grains[4 * g + 1];
          const double d = std::fma(dy, dy, dx * dx);
          if (!have || d < bestd) {
            have = true;
            bestd = d;
            best = grains[4 * g + 3];
          }
        }
        tb[j * nx + i] = best;
      }
  }
  return lgd_sim_set_orientation(sim, theta.data(), theta.size());
}

}  // extern "C"
