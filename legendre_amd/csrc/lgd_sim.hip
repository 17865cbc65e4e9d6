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

#include "lgd_internal.h"

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
  double t = 0.0;
  int64_t step_index = 0;
};

using namespace lgd;

namespace {

void free_all(lgd_sim *s) {
  for (auto *b : s->state) lgd_buf_free(b);
  for (auto *b : s->next) lgd_buf_free(b);
  for (auto *b : s->k) lgd_buf_free(b);
  lgd_buf_free(s->amp);
  s->state.clear();
  s->next.clear();
  s->k.clear();
  s->amp = nullptr;
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

int lgd_sim_destroy(lgd_sim *sim) {
  if (!sim) return LGD_OK;
  free_all(sim);
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
    int rc = sim->kind == lgd_sim::kModelC ? step_modelc(sim, dt, salt) : step_heat(sim, dt);
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
