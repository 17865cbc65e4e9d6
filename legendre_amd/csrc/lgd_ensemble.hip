// Batched ensemble of independent Model C realizations.
//
// Every realization is its own Simulation (BlockIds from 0, its own seed and
// parameters, mirror faces); all of them share the block shape, so their slabs
// concatenate into one field buffer and the streaming Model C kernel advances
// every block of every realization in one launch per step, reading each
// block's domain (constants, topology, noise key) from a small device table.
#include <cmath>
#include <cstring>
#include <new>
#include <vector>

#include "lgd_internal.h"

using namespace lgd;

struct lgd_ensemble {
  lgd_ctx *ctx = nullptr;
  int64_t nx = 0, ny = 0;
  double h[2] = {0.0, 0.0};
  double anisotropy_tol = 1e-8, fv_tol = 1e-8;
  std::vector<lgd_realization> real;
  std::vector<lgd_domain> domains;     // host copy (noise scale set per dt)
  std::vector<int64_t> block0;         // first block per realization, n + 1
  lgd_domain *d_domains = nullptr;
  int32_t *d_block_domain = nullptr;
  unsigned long long *d_counts = nullptr;
  lgd_buf *state[2] = {nullptr, nullptr};
  lgd_buf *next[2] = {nullptr, nullptr};
  double dt_uploaded = -1.0;
  double t = 0.0;
};

namespace {

// Count interior cells with phi > 0 per realization (integer atomics: the
// result is order-independent).
__global__ void solid_count_kernel(const double *phi, const int32_t *block_domain, int64_t nx,
                                   int64_t ny, unsigned long long *counts) {
  const int64_t b = blockIdx.x;
  const int64_t sx = nx + 2, slab = sx * (ny + 2);
  unsigned long long c = 0;
  for (int64_t t = threadIdx.x; t < nx * ny; t += blockDim.x) {
    int64_t j = t / nx, i = t - j * nx;
    c += phi[b * slab + (j + 1) * sx + (i + 1)] > 0.0 ? 1ull : 0ull;
  }
  for (int off = 32; off > 0; off >>= 1) c += __shfl_down(c, off, 64);
  if ((threadIdx.x & 63) == 0 && c) atomicAdd(&counts[block_domain[b]], c);
}

int upload_domains(lgd_ensemble *e, double dt) {
  if (e->dt_uploaded == dt) return LGD_OK;
  for (size_t r = 0; r < e->real.size(); ++r) {
    const lgd_realization &R = e->real[r];
    e->domains[r].noise_scale = R.noise_amplitude != 0.0 ? std::sqrt(dt) * R.noise_amplitude : 0.0;
  }
  hipError_t err = hipMemcpyAsync(e->d_domains, e->domains.data(),
                                  e->domains.size() * sizeof(lgd_domain), hipMemcpyHostToDevice,
                                  e->ctx->stream);
  if (err != hipSuccess) return hip_fail(e->ctx, err, "upload domain table");
  err = hipStreamSynchronize(e->ctx->stream);
  if (err != hipSuccess) return hip_fail(e->ctx, err, "upload domain table");
  e->dt_uploaded = dt;
  return LGD_OK;
}

void destroy(lgd_ensemble *e) {
  for (int f = 0; f < 2; ++f) {
    lgd_buf_free(e->state[f]);
    lgd_buf_free(e->next[f]);
  }
  (void)hipFree(e->d_domains);
  (void)hipFree(e->d_block_domain);
  (void)hipFree(e->d_counts);
  delete e;
}

}  // namespace

extern "C" {

int lgd_ensemble_create(lgd_ctx *ctx, const lgd_realization *r, size_t n,
                        const int64_t *block_cells2, const double *spacing2, double anisotropy_tol,
                        double fv_tol, lgd_ensemble **out) {
  if (!ctx || !r || !n || !block_cells2 || !spacing2 || !out)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null or empty argument");
  *out = nullptr;
  if (block_cells2[0] < 1 || block_cells2[1] < 1 || !(spacing2[0] > 0.0) || !(spacing2[1] > 0.0))
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "invalid block shape or spacing");
  if (block_cells2[0] + 2 > 128)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "ensemble blocks are at most 126 cells wide");
  lgd_ensemble *e = new (std::nothrow) lgd_ensemble();
  if (!e) return fail(ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  e->ctx = ctx;
  e->nx = block_cells2[0];
  e->ny = block_cells2[1];
  e->h[0] = spacing2[0];
  e->h[1] = spacing2[1];
  e->anisotropy_tol = anisotropy_tol;
  e->fv_tol = fv_tol;
  e->real.assign(r, r + n);
  e->block0.assign(1, 0);
  std::vector<int32_t> block_domain;
  lgd_cart c{};
  c.dim = 2;
  c.h[0] = spacing2[0];
  c.h[1] = spacing2[1];
  for (size_t i = 0; i < n; ++i) {
    const lgd_realization &R = r[i];
    if (R.cells[0] <= 0 || R.cells[1] <= 0 || R.cells[0] % e->nx || R.cells[1] % e->ny) {
      destroy(e);
      return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "realization cells must be multiples of the block");
    }
    lgd_modelc_params p{R.eps4, R.lambda, R.d_thermal, R.noise_amplitude, anisotropy_tol, fv_tol};
    lgd_domain d{};
    d.k = lgd_make_kr_consts(p, c);
    d.noise_scale = 0.0;
    d.seed = R.seed;
    d.block0 = e->block0.back();
    d.nbx = (int32_t)(R.cells[0] / e->nx);
    d.nby = (int32_t)(R.cells[1] / e->ny);
    d.per_x = 0;
    d.per_y = 0;
    e->domains.push_back(d);
    int64_t nb = (int64_t)d.nbx * d.nby;
    for (int64_t b = 0; b < nb; ++b) block_domain.push_back((int32_t)i);
    e->block0.push_back(e->block0.back() + nb);
  }
  int64_t nblocks = e->block0.back();
  size_t len = (size_t)(nblocks * (e->nx + 2) * (e->ny + 2));
  int rc = LGD_OK;
  for (int f = 0; f < 2 && !rc; ++f) {
    rc = lgd_buf_alloc(ctx, len, &e->state[f]);
    if (!rc) rc = lgd_buf_alloc(ctx, len, &e->next[f]);
  }
  hipError_t err = hipSuccess;
  if (!rc) err = hipMalloc((void **)&e->d_domains, n * sizeof(lgd_domain));
  if (!rc && err == hipSuccess) err = hipMalloc((void **)&e->d_block_domain, block_domain.size() * sizeof(int32_t));
  if (!rc && err == hipSuccess) err = hipMalloc((void **)&e->d_counts, n * sizeof(unsigned long long));
  if (!rc && err == hipSuccess)
    err = hipMemcpy(e->d_block_domain, block_domain.data(), block_domain.size() * sizeof(int32_t),
                    hipMemcpyHostToDevice);
  if (!rc && err != hipSuccess) rc = hip_fail(ctx, err, "ensemble tables");
  if (rc) {
    destroy(e);
    return rc;
  }
  *out = e;
  return LGD_OK;
}

int lgd_ensemble_destroy(lgd_ensemble *e) {
  if (!e) return LGD_OK;
  (void)hipStreamSynchronize(e->ctx->stream);
  destroy(e);
  return LGD_OK;
}

int64_t lgd_ensemble_num_blocks(const lgd_ensemble *e) { return e ? e->block0.back() : -1; }

int64_t lgd_ensemble_block_start(const lgd_ensemble *e, int64_t r) {
  if (!e || r < 0 || r > (int64_t)e->real.size()) return -1;
  return e->block0[(size_t)r];
}

int lgd_ensemble_field(lgd_ensemble *e, int32_t field, lgd_buf **out) {
  if (!e || !out || field < 0 || field > 1) return LGD_ERR_INVALID_ARGUMENT;
  *out = e->state[field];
  return LGD_OK;
}

int lgd_ensemble_initialize_grains(lgd_ensemble *e, int64_t r, const double *grains, size_t n,
                                   double u0) {
  if (!e || (!grains && n) || r < 0 || r >= (int64_t)e->real.size()) return LGD_ERR_INVALID_ARGUMENT;
  // ModelC::initialize_grains (src/physics/phasefield.rs:227-286) on this
  // realization's blocks, computed on the host and uploaded.
  const lgd_domain &d = e->domains[(size_t)r];
  const int64_t nx = e->nx, ny = e->ny, sx = nx + 2, slab = sx * (ny + 2);
  const int64_t nb = (int64_t)d.nbx * d.nby;
  std::vector<double> phi((size_t)(nb * slab), 0.0), u((size_t)(nb * slab), 0.0);
  const double sqrt2 = std::sqrt(2.0), cutoff = 27.0;
  std::vector<size_t> near;
  for (int64_t b = 0; b < nb; ++b) {
    int64_t bx = b % d.nbx, by = b / d.nbx;
    auto center = [&](int64_t i, int64_t j, double *xy) {
      xy[0] = std::fma((double)(bx * nx) + (double)i + 0.5, e->h[0], 0.0);
      xy[1] = std::fma((double)(by * ny) + (double)j + 0.5, e->h[1], 0.0);
    };
    double c0[2], c1[2];
    center(0, 0, c0);
    center(nx - 1, ny - 1, c1);
    double lo[2] = {std::fma(0.5, -e->h[0], c0[0]), std::fma(0.5, -e->h[1], c0[1])};
    double hi[2] = {std::fma(0.5, e->h[0], c1[0]), std::fma(0.5, e->h[1], c1[1])};
    near.clear();
    for (size_t g = 0; g < n; ++g) {
      double cx = grains[3 * g], cy = grains[3 * g + 1];
      double dx = std::fmax(std::fmax(lo[0] - cx, cx - hi[0]), 0.0);
      double dy = std::fmax(std::fmax(lo[1] - cy, cy - hi[1]), 0.0);
      if (std::hypot(dx, dy) <= grains[3 * g + 2] + cutoff) near.push_back(g);
    }
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        double v = -1.0;
        if (!near.empty()) {
          double xy[2];
          center(i, j, xy);
          for (size_t g : near) {
            double rr = std::hypot(xy[0] - grains[3 * g], xy[1] - grains[3 * g + 1]);
            v = std::fmax(v, -std::tanh((rr - grains[3 * g + 2]) / sqrt2));
          }
        }
        phi[(size_t)(b * slab + (j + 1) * sx + i + 1)] = v;
        u[(size_t)(b * slab + (j + 1) * sx + i + 1)] = -u0;
      }
  }
  int rc = lgd_buf_upload(e->ctx, e->state[0], (size_t)(d.block0 * slab), phi.data(), phi.size());
  if (rc) return rc;
  return lgd_buf_upload(e->ctx, e->state[1], (size_t)(d.block0 * slab), u.data(), u.size());
}

int lgd_ensemble_step(lgd_ensemble *e, double dt, int64_t nsteps, int32_t per_block) {
  if (!e) return LGD_ERR_INVALID_ARGUMENT;
  int rc = upload_domains(e, dt);
  if (rc) return rc;
  bool noise = false;
  for (const lgd_realization &R : e->real) noise = noise || R.noise_amplitude != 0.0;
  const int64_t nblocks = e->block0.back();
  // Block pairs (kernel 3) when every realization's block rows pair up (then every
  // realization's first block is even too); by default with one strip per block
  // and one row loaded ahead, else one block per wave, one strip, two rows ahead.
  bool pairable = true;
  for (const lgd_domain &D : e->domains) pairable = pairable && lgd_stream_pairable(e->nx, D.nbx);
  const lgd_launch_shape dflt = pairable ? lgd_launch_shape{3, 1, 1} : lgd_launch_shape{2, 1, 2};
  const lgd_launch_shape shape = lgd_modelc_shape(e->ctx, e->nx, e->ny, nblocks, noise, dflt);
  const bool pair = shape.kernel == 3 && pairable;
  const uint64_t wiener0 = (1ull << 32) | 0ull;
  for (int64_t s = 0; s < nsteps; ++s) {
    uint64_t salt;
    std::memcpy(&salt, &e->t, sizeof salt);  // t.to_bits(), shared by every realization
    if (!per_block) {
      rc = lgd_modelc_stream_launch(e->ctx, e->nx, e->ny, e->domains[0], e->d_domains,
                                    e->d_block_domain, 0, nblocks, e->state[0]->data,
                                    e->state[1]->data, e->next[0]->data, e->next[1]->data, dt,
                                    noise, salt, wiener0, shape.strips, shape.prefetch, pair);
    } else {
      // The per-block dispatch baseline: one launch per block, each in its
      // fastest single-block shape (one wave per row).
      for (int64_t b = 0; b < nblocks && !rc; ++b)
        rc = lgd_modelc_stream_launch(e->ctx, e->nx, e->ny, e->domains[0], e->d_domains,
                                      e->d_block_domain, b, 1, e->state[0]->data,
                                      e->state[1]->data, e->next[0]->data, e->next[1]->data, dt,
                                      noise, salt, wiener0, (int)e->ny, shape.prefetch);
    }
    if (rc) return rc;
    std::swap(e->state[0], e->next[0]);
    std::swap(e->state[1], e->next[1]);
    e->t = e->t + dt;
  }
  return LGD_OK;
}

double lgd_ensemble_time(const lgd_ensemble *e) { return e ? e->t : 0.0; }

int lgd_ensemble_solid_fraction(lgd_ensemble *e, double *out) {
  if (!e || !out) return LGD_ERR_INVALID_ARGUMENT;
  const size_t n = e->real.size();
  hipError_t err = hipMemsetAsync(e->d_counts, 0, n * sizeof(unsigned long long), e->ctx->stream);
  if (err != hipSuccess) return hip_fail(e->ctx, err, "clear counts");
  solid_count_kernel<<<(unsigned)e->block0.back(), 256, 0, e->ctx->stream>>>(
      e->state[0]->data, e->d_block_domain, e->nx, e->ny, e->d_counts);
  int rc = launch_check(e->ctx, "solid_count");
  if (rc) return rc;
  std::vector<unsigned long long> counts(n);
  err = hipMemcpyAsync(counts.data(), e->d_counts, n * sizeof(unsigned long long),
                       hipMemcpyDeviceToHost, e->ctx->stream);
  if (err == hipSuccess) err = hipStreamSynchronize(e->ctx->stream);
  if (err != hipSuccess) return hip_fail(e->ctx, err, "download counts");
  for (size_t i = 0; i < n; ++i)
    out[i] = (double)counts[i] / (double)(e->real[i].cells[0] * e->real[i].cells[1]);
  return LGD_OK;
}

}  // extern "C"
