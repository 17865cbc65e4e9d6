// The explicit update of every driver in one launch: State::apply_step_with
// (src/core/state.rs:646-699) with Driver::apply_slab's three arms
// (src/core/driver.rs:115-182).
//
// One thread per slab entry of one ghost width carries every field of that width
// through the drivers in order (Time, then each stochastic driver), so an entry is
// read and written once per step whatever the number of drivers. The per-cell draw
// of a stochastic driver (a Wiener deviate, a jump fire) is made once and shared by
// every field the driver moves, as the reference's multi-field kernel shares it.
// Workgroups tile one slab at a time, so the block id, and with it each driver's
// block key mix_key(seed, [salt, stream, block]), is uniform across a workgroup and
// is computed once per wave on the scalar unit.
#include <cmath>

#include "lgd_internal.h"

namespace {

constexpr int kMaxFields = 8;
constexpr int kMaxDrivers = 8;
constexpr int kWG = 256;

struct DriverArgs {
  int64_t e[3];   // slab extents of this ghost width
  int64_t g;      // ghost width
  int64_t slab;   // entries per slab
  int64_t ncell;  // interior cells per block (the intensity slab)
  uint32_t tiles; // workgroups per slab
  int32_t dim;
  int32_t nf, nd;
  double *x[kMaxFields];
  int32_t kind[kMaxDrivers];
  uint64_t stream[kMaxDrivers];
  const double *amp[kMaxDrivers][kMaxFields];  // nullptr: the driver does not move the field
  const double *rate[kMaxDrivers];
  double dt, sqrt_dt;
  uint64_t seed, salt;
  const double *salt_src;
};

// Five waves per SIMD fit in 96 VGPRs without spills; the kernel is latency-bound
// between its loads and its transcendental chains, so the extra wave pays.
__global__ __launch_bounds__(kWG) __attribute__((amdgpu_waves_per_eu(5))) void drivers_kernel(
    DriverArgs a) {
  const uint32_t b = blockIdx.x / a.tiles;
  int64_t r = (int64_t)(blockIdx.x - b * a.tiles) * kWG + threadIdx.x;
  if (r >= a.slab) return;
  const int64_t t = (int64_t)b * a.slab + r;
  // Grid::cell_key (src/geometry/cartesian.rs:599-619): the interior linear index,
  // or -1 for a ghost entry.
  int64_t cell = 0, stride = 1;
  if (a.g == 0) {
    cell = r;
  } else {
    for (int d = 0; d < a.dim; ++d) {
      int64_t cc = r % a.e[d];
      r /= a.e[d];
      int64_t n = a.e[d] - 2 * a.g;
      if (cc < a.g || cc >= n + a.g) {
        cell = -1;
        break;
      }
      cell += (cc - a.g) * stride;
      stride *= n;
    }
  }
  double x[kMaxFields];
#pragma unroll
  for (int f = 0; f < kMaxFields; ++f)
    if (f < a.nf) x[f] = a.x[f][t];
  const uint64_t salt = lgd_salt(a.salt, a.salt_src);
  for (int d = 0; d < a.nd; ++d) {
    const int32_t kind = a.kind[d];
    if (kind == LGD_DRIVER_TIME) {
      // x += dt·v over the whole slab, ghosts included (driver.rs:131-138).
#pragma unroll
      for (int f = 0; f < kMaxFields; ++f)
        if (f < a.nf && a.amp[d][f]) x[f] = x[f] + a.dt * a.amp[d][f][t];
      continue;
    }
    if (cell < 0) continue;  // stochastic drivers never touch ghosts
    const uint64_t block_key = lgd_mix_key3(a.seed, salt, a.stream[d], (uint64_t)b);
    const uint64_t key = lgd_splitmix64(block_key ^ (uint64_t)cell);
    if (kind == LGD_DRIVER_WIENER) {
      // x += (√dt·v)·ξ where v ≠ 0 (driver.rs:139-158); ξ drawn at most once.
      bool drawn = false;
      double xi = 0.0;
#pragma unroll
      for (int f = 0; f < kMaxFields; ++f) {
        if (f >= a.nf || !a.amp[d][f]) continue;
        const double v = a.amp[d][f][t];
        if (v == 0.0) continue;
        if (!drawn) {
          xi = lgd_standard_normal(key);
          drawn = true;
        }
        x[f] = x[f] + ((a.sqrt_dt * v) * xi);
      }
    } else {
      // One fire per cell, Bernoulli(1 − e^{−λ·dt}) against the intensity the
      // model wrote; every moved field takes its increment (driver.rs:159-182).
      const double lambda = a.rate[d][b * a.ncell + cell];
      if (!(lambda <= 0.0)) {
        const double p = -expm1(-lambda * a.dt);
        if (lgd_unit_open(key) < p) {
#pragma unroll
          for (int f = 0; f < kMaxFields; ++f)
            if (f < a.nf && a.amp[d][f]) x[f] = x[f] + a.amp[d][f][t];
        }
      }
    }
  }
#pragma unroll
  for (int f = 0; f < kMaxFields; ++f)
    if (f < a.nf) a.x[f][t] = x[f];
}

bool moves(const lgd_buf *amp) { return amp && amp->len > 0; }

}  // namespace

using namespace lgd;

extern "C" {

int lgd_apply_drivers(lgd_ctx *ctx, const lgd_grid *grid, size_t nfields, lgd_buf *const *x,
                      const uint32_t *ghost, size_t ndrivers, const lgd_driver *drivers, double dt,
                      uint64_t seed, uint64_t salt) {
  if (!ctx || !grid) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (nfields == 0 || ndrivers == 0) return LGD_OK;
  if (!x || !ghost || !drivers) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null array");
  if (nfields > (size_t)kMaxFields || ndrivers > (size_t)kMaxDrivers)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "at most 8 fields and 8 drivers per update");
  int rc;
  for (size_t f = 0; f < nfields; ++f) {
    if ((rc = check_field(ctx, grid, x[f], ghost[f], "state field"))) return rc;
  }
  for (size_t d = 0; d < ndrivers; ++d) {
    const lgd_driver &dr = drivers[d];
    if (dr.kind != LGD_DRIVER_TIME && dr.kind != LGD_DRIVER_WIENER && dr.kind != LGD_DRIVER_JUMP)
      return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "unknown driver kind");
    if (!dr.amp) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "driver without a field array");
    for (size_t f = 0; f < nfields; ++f)
      if (moves(dr.amp[f]) && (rc = check_field(ctx, grid, dr.amp[f], ghost[f], "driver field")))
        return rc;
    if (dr.kind == LGD_DRIVER_JUMP) {
      // Driver::Jump's kernel expects an intensity slab (driver.rs:160).
      if (!moves(dr.rate))
        return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "a jump driver must be given an intensity slab");
      if ((rc = check_field(ctx, grid, dr.rate, 0, "jump intensity"))) return rc;
    }
  }
  const lgd_cart &c = grid->cart;
  // One launch per distinct ghost width, each carrying that width's fields.
  bool done[kMaxFields] = {};
  for (size_t f0 = 0; f0 < nfields; ++f0) {
    if (done[f0]) continue;
    const uint32_t g = ghost[f0];
    DriverArgs a{};
    a.g = g;
    a.dim = c.dim;
    for (int d = 0; d < 3; ++d) a.e[d] = d < c.dim ? c.n[d] + 2 * (int64_t)g : 1;
    a.slab = slab_len(c, g);
    a.ncell = slab_len(c, 0);
    a.tiles = (uint32_t)((a.slab + kWG - 1) / kWG);
    a.nd = (int32_t)ndrivers;
    a.dt = dt;
    a.sqrt_dt = std::sqrt(dt);
    a.seed = seed;
    a.salt = salt;
    a.salt_src = ctx->salt_src;
    int nf = 0;
    for (size_t f = f0; f < nfields; ++f) {
      if (done[f] || ghost[f] != g) continue;
      done[f] = true;
      a.x[nf] = x[f]->data;
      for (size_t d = 0; d < ndrivers; ++d)
        a.amp[d][nf] = moves(drivers[d].amp[f]) ? drivers[d].amp[f]->data : nullptr;
      ++nf;
    }
    a.nf = nf;
    for (size_t d = 0; d < ndrivers; ++d) {
      const lgd_driver &dr = drivers[d];
      a.kind[d] = dr.kind;
      // Driver::stream (src/core/driver.rs:55-61).
      a.stream[d] = dr.kind == LGD_DRIVER_TIME ? 0ull : ((uint64_t)dr.kind << 32) | dr.index;
      a.rate[d] = dr.kind == LGD_DRIVER_JUMP ? dr.rate->data : nullptr;
    }
    if (a.slab == 0 || c.nblocks == 0) continue;
    // A launch covers fewer than 2^32 work-items.
    if ((uint64_t)c.nblocks * a.tiles * kWG > 0xFFFFFFFFull)
      return fail(ctx, LGD_ERR_UNSUPPORTED, "more than 2^32 work-items in one update");
    drivers_kernel<<<(unsigned)(c.nblocks * a.tiles), kWG, 0, ctx->stream>>>(a);
    if ((rc = launch_check(ctx, "apply_drivers"))) return rc;
  }
  return LGD_OK;
}

}  // extern "C"
