// Per-intent kernels: each mirrors one reference function over every block of a
// field buffer.  They are the building blocks of the reference's own call
// sequence (LGD_MODE_PER_INTENT) and of the AMR/level paths; the fused step
// kernels in lgd_fused.hip replace the whole sequence on the hot path.
#include <cmath>
#include <cstdio>

#include "lgd_internal.h"

namespace {

constexpr int kWG = 256;

// Slab extents for ghost width g (unused dims 1).
struct Ext {
  int64_t e[3];
  int64_t g;
};

__host__ __device__ inline Ext make_ext(const lgd_cart &c, int64_t g) {
  Ext x;
  for (int d = 0; d < 3; ++d) x.e[d] = d < c.dim ? c.n[d] + 2 * g : 1;
  x.g = g;
  return x;
}

// linearize (src/geometry/cartesian.rs:189-204) for idx relative to the interior origin.
__device__ inline int64_t lin(const lgd_cart &c, const Ext &x, const int64_t idx[3]) {
  int64_t off = 0, stride = 1;
  for (int d = 0; d < c.dim; ++d) {
    off += (idx[d] + x.g) * stride;
    stride *= x.e[d];
  }
  return off;
}

__device__ inline void block_coords(const lgd_cart &c, int64_t b, int64_t out[3]) {
  for (int d = 0; d < 3; ++d) {
    out[d] = b % c.nb[d];
    b /= c.nb[d];
  }
}

// face_neighbor (src/geometry/cartesian.rs:162-181); -1 for None.
__device__ inline int64_t face_neighbor(const lgd_cart &c, int64_t block, int dim, int dir) {
  int64_t co[3];
  block_coords(c, block, co);
  int64_t stepped = co[dim] + dir;
  int64_t nb = c.nb[dim];
  if (stepped < 0 || stepped >= nb) {
    if (!c.periodic[dim]) return -1;
    stepped = ((stepped % nb) + nb) % nb;
  }
  co[dim] = stepped;
  return co[0] + c.nb[0] * (co[1] + c.nb[1] * co[2]);
}

// Decode an interior-cell ordinal (dimension 0 fastest) within a block.
__device__ inline void interior_idx(const lgd_cart &c, int64_t r, int64_t idx[3]) {
  for (int d = 0; d < 3; ++d) {
    idx[d] = r % c.n[d];
    r /= c.n[d];
  }
}

// One dimension sweep of fill_ghosts_generic (src/geometry/cartesian.rs:464-525):
// every (block, side) ghost strip in dimension d gets the neighbour's d-interior
// (exchange / periodic wrap, self included) or the face rule at a physical face.
// A sweep writes only d-ghosts and reads only d-interiors, so all strips of one
// sweep are independent; sweeps run in dimension order as separate launches.
__global__ void fill_sweep_kernel(lgd_cart c, int d, Ext x, int64_t strip, double *field,
                                  lgd_face_bc bc_lo, lgd_face_bc bc_hi) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t total = c.nblocks * 2 * strip;
  if (t >= total) return;
  int64_t b = t / (2 * strip);
  int64_t r = t - b * 2 * strip;
  int side = r >= strip ? 1 : 0;
  r -= side * strip;
  int64_t n = c.n[d], g = x.g;
  int64_t idx[3] = {0, 0, 0};
  for (int e = 0; e < c.dim; ++e) {
    int64_t ext = e == d ? g : x.e[e];
    int64_t v = r % ext;
    r /= ext;
    if (e == d)
      idx[e] = side ? n + v : -g + v;
    else
      idx[e] = v - g;
  }
  int64_t slab = x.e[0] * x.e[1] * x.e[2];
  double *dst = field + b * slab;
  int64_t nbk = face_neighbor(c, b, d, side ? 1 : -1);
  double val;
  if (nbk >= 0) {
    int64_t sidx[3] = {idx[0], idx[1], idx[2]};
    sidx[d] = side ? idx[d] - n : n + idx[d];
    val = field[nbk * slab + lin(c, x, sidx)];
  } else {
    const lgd_face_bc bc = side ? bc_hi : bc_lo;
    int64_t layer = side ? idx[d] - n : -1 - idx[d];
    auto interior = [&](int64_t inward) {
      int64_t iidx[3] = {idx[0], idx[1], idx[2]};
      iidx[d] = side ? n - 1 - inward : inward;
      return dst[lin(c, x, iidx)];
    };
    // FaceBc::ghost_value (src/geometry/boundary.rs:63-73).
    switch (bc.kind) {
      case LGD_BC_DIRICHLET:
        val = (2.0 * bc.value) - interior(layer);
        break;
      case LGD_BC_FLUX:
        val = interior(layer) + (((double)(2 * layer + 1) * c.h[d]) * bc.value);
        break;
      case LGD_BC_EXTRAPOLATE:
        val = ((double)(layer + 2) * interior(0)) - ((double)(layer + 1) * interior(1));
        break;
      default:
        val = interior(layer);
    }
  }
  dst[lin(c, x, idx)] = val;
}

// central_laplacian_kernel (src/discretization/finite_difference.rs:49-67),
// optionally followed by the heat model's `* kappa` (examples/heat3d.rs:76-78).
template <bool kScale>
__global__ void laplacian_kernel(lgd_cart c, Ext x, double ih0, double ih1, double ih2,
                                 double kappa, const double *in, double *out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t ncell = c.n[0] * c.n[1] * c.n[2];
  if (t >= c.nblocks * ncell) return;
  int64_t b = t / ncell;
  int64_t idx[3];
  interior_idx(c, t - b * ncell, idx);
  int64_t slab = x.e[0] * x.e[1] * x.e[2];
  const double *s = in + b * slab;
  int64_t o = lin(c, x, idx);
  double center = s[o];
  double inv[3] = {ih0, ih1, ih2};
  double acc = 0.0;
  int64_t stride = 1;
  for (int d = 0; d < c.dim; ++d) {
    acc = acc + (((s[o + stride] - (center + center)) + s[o - stride]) * inv[d]);
    stride *= x.e[d];
  }
  out[b * slab + o] = kScale ? acc * kappa : acc;
}

// KarmaRappelFlux::kernel (src/discretization/finite_volume.rs:248-284), 2-D, ghost 1.
__global__ void kr_flux_kernel(lgd_cart c, lgd_kr_consts k, const double *phi, double *out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t ncell = c.n[0] * c.n[1];
  if (t >= c.nblocks * ncell) return;
  int64_t b = t / ncell;
  int64_t r = t - b * ncell;
  int64_t i = r % c.n[0], j = r / c.n[0];
  int64_t sx = c.n[0] + 2;
  int64_t slab = sx * (c.n[1] + 2);
  const double *s = phi + b * slab + (j + 1) * sx + (i + 1);
  auto p = [&](int di, int dj) { return s[dj * sx + di]; };
  const double q = 0.25;
  double derx_r = p(1, 0) - p(0, 0);
  double derx_l = p(0, 0) - p(-1, 0);
  double derx_t = q * (((p(1, 1) - p(-1, 1)) + p(1, 0)) - p(-1, 0));
  double derx_b = q * (((p(1, 0) - p(-1, 0)) + p(1, -1)) - p(-1, -1));
  double dery_t = p(0, 1) - p(0, 0);
  double dery_b = p(0, 0) - p(0, -1);
  double dery_r = q * (((p(1, 1) - p(1, -1)) + p(0, 1)) - p(0, -1));
  double dery_l = q * (((p(0, 1) - p(0, -1)) + p(-1, 1)) - p(-1, -1));
  double a_r, da_r, a_l, da_l, a_t, da_t, a_b, da_b;
  lgd_kr_face(k, derx_r, dery_r, a_r, da_r);
  lgd_kr_face(k, derx_l, dery_l, a_l, da_l);
  lgd_kr_face(k, derx_t, dery_t, a_t, da_t);
  lgd_kr_face(k, derx_b, dery_b, a_b, da_b);
  double j_r = a_r * ((a_r * derx_r) - (da_r * dery_r));
  double j_l = a_l * ((a_l * derx_l) - (da_l * dery_l));
  double j_t = a_t * ((a_t * dery_t) + (da_t * derx_t));
  double j_b = a_b * ((a_b * dery_b) + (da_b * derx_b));
  out[b * slab + (j + 1) * sx + (i + 1)] = (((j_r - j_l) + j_t) - j_b) * k.inv_h2;
}

// ModelC::vector_field_block(Time) (src/physics/phasefield.rs:367-421): each
// thread evaluates one interior cell's dφ and du from ghost-filled slabs.
// cs0 / sn0 non-null: the multi-grain field, (cos θ₀, sin θ₀) per cell (ghost-0).
__global__ void modelc_rhs_kernel(lgd_cart c, lgd_kr_consts k, const double *phi,
                                  const double *u, double *kphi, double *ku, const double *cs0,
                                  const double *sn0) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t ncell = c.n[0] * c.n[1];
  if (t >= c.nblocks * ncell) return;
  int64_t b = t / ncell;
  int64_t r = t - b * ncell;
  int64_t i = r % c.n[0], j = r / c.n[0];
  int64_t sx = c.n[0] + 2;
  int64_t o = b * sx * (c.n[1] + 2) + (j + 1) * sx + (i + 1);
  double p[3][3];
  for (int dj = -1; dj <= 1; ++dj)
    for (int di = -1; di <= 1; ++di) p[dj + 1][di + 1] = phi[o + dj * sx + di];
  double dphi, du;
  if (cs0)
    lgd_modelc_cell_oriented(k, p, u[o], u[o + 1], u[o - 1], u[o + sx], u[o - sx], cs0[t], sn0[t],
                             dphi, du);
  else
    lgd_modelc_cell(k, p, u[o], u[o + 1], u[o - 1], u[o + sx], u[o - sx], dphi, du);
  kphi[o] = dphi;
  ku[o] = du;
}

// eval_tendency(Wiener(0)) for Model C: zero the slab, interior = b
// (src/integrators/mod.rs:263; src/physics/phasefield.rs:361-366).
__global__ void modelc_amp_kernel(lgd_cart c, Ext x, double amp, double *out) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t slab = x.e[0] * x.e[1] * x.e[2];
  if (t >= c.nblocks * slab) return;
  int64_t r = t % slab;
  bool interior = true;
  for (int d = 0; d < c.dim; ++d) {
    int64_t cc = r % x.e[d];
    r /= x.e[d];
    interior = interior && cc >= x.g && cc < c.n[d] + x.g;
  }
  out[t] = interior ? amp : 0.0;
}

// block_axpy / Driver::Time: x += alpha*y over every entry (src/core/state.rs:494-496).
__global__ void axpy_kernel(int64_t n, double alpha, double *x, const double *y) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  x[t] = x[t] + alpha * y[t];
}

__global__ void fill_zero_kernel(int64_t n, double *x) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  x[t] = 0.0;
}

// Driver::Wiener arm of apply_slab (src/core/driver.rs:139-158) with the ghost
// gate of Grid::cell_key (src/geometry/cartesian.rs:599-619).
__global__ void wiener_kernel(lgd_cart c, Ext x, double scale, uint64_t seed, uint64_t salt,
                              const double *salt_src, uint64_t stream, double *state,
                              const double *amp) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  int64_t slab = x.e[0] * x.e[1] * x.e[2];
  if (t >= c.nblocks * slab) return;
  double v = amp[t];
  if (v == 0.0) return;
  int64_t b = t / slab;
  int64_t r = t - b * slab;
  int64_t key = 0, stride = 1;
  for (int d = 0; d < c.dim; ++d) {
    int64_t cc = r % x.e[d];
    r /= x.e[d];
    if (cc < x.g || cc >= c.n[d] + x.g) return;
    key += (cc - x.g) * stride;
    stride *= c.n[d];
  }
  uint64_t block_key = lgd_mix_key3(seed, lgd_salt(salt, salt_src), stream, (uint64_t)b);
  uint64_t k = lgd_splitmix64(block_key ^ (uint64_t)key);
  state[t] = state[t] + ((scale * v) * lgd_standard_normal(k));
}

}  // namespace

using namespace lgd;

extern "C" {

int lgd_buf_fill_zero(lgd_ctx *ctx, lgd_buf *buf) {
  if (!ctx || !buf) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (buf->len == 0) return LGD_OK;
  fill_zero_kernel<<<blocks_for((int64_t)buf->len, kWG), kWG, 0, ctx->stream>>>((int64_t)buf->len,
                                                                              buf->data);
  return launch_check(ctx, "fill_zero");
}

int lgd_fill_ghosts(lgd_ctx *ctx, const lgd_grid *grid, lgd_buf *field, uint32_t ghost,
                    const lgd_face_bc *bcs) {
  if (!ctx || !grid) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  int rc = check_field(ctx, grid, field, ghost, "field");
  if (rc) return rc;
  if (ghost == 0) return LGD_OK;
  if ((rc = check_ghost(ctx, grid, ghost))) return rc;
  const lgd_cart &c = grid->cart;
  if (bcs) {
    for (int d = 0; d < c.dim; ++d)
      for (int s = 0; s < 2; ++s) {
        int k = bcs[d * 2 + s].kind;
        if (k < LGD_BC_MIRROR || k > LGD_BC_EXTRAPOLATE)
          return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "unknown FaceBc kind");
        // Linear extrapolation reads two interior cells (cartesian.rs:418-421).
        if (k == LGD_BC_EXTRAPOLATE && c.n[d] < 2)
          return fail(ctx, LGD_ERR_INVALID_ARGUMENT,
                      "linear extrapolation reads two interior cells per dimension");
      }
  }
  Ext x = make_ext(c, ghost);
  for (int d = 0; d < c.dim; ++d) {
    int64_t strip = ghost;
    for (int e = 0; e < c.dim; ++e)
      if (e != d) strip *= x.e[e];
    lgd_face_bc lo = {LGD_BC_MIRROR, 0.0}, hi = {LGD_BC_MIRROR, 0.0};
    if (bcs) {
      lo = bcs[d * 2];
      hi = bcs[d * 2 + 1];
    }
    int64_t total = c.nblocks * 2 * strip;
    fill_sweep_kernel<<<blocks_for(total, kWG), kWG, 0, ctx->stream>>>(c, d, x, strip, field->data,
                                                                     lo, hi);
    if ((rc = launch_check(ctx, "fill_ghosts"))) return rc;
  }
  return LGD_OK;
}

static int laplacian_common(lgd_ctx *ctx, const lgd_grid *grid, const lgd_buf *in, lgd_buf *out,
                            uint32_t ghost, bool scale, double kappa) {
  if (!ctx || !grid) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  int rc;
  if ((rc = check_field(ctx, grid, in, ghost, "in"))) return rc;
  if ((rc = check_field(ctx, grid, out, ghost, "out"))) return rc;
  if (ghost < 1) return fail(ctx, LGD_ERR_GHOST_TOO_WIDE, "the Laplacian needs ghost width >= 1");
  const lgd_cart &c = grid->cart;
  Ext x = make_ext(c, ghost);
  double ih[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < c.dim; ++d) ih[d] = 1.0 / (c.h[d] * c.h[d]);
  int64_t total = c.nblocks * c.n[0] * c.n[1] * c.n[2];
  if (scale)
    laplacian_kernel<true><<<blocks_for(total, kWG), kWG, 0, ctx->stream>>>(
        c, x, ih[0], ih[1], ih[2], kappa, in->data, out->data);
  else
    laplacian_kernel<false><<<blocks_for(total, kWG), kWG, 0, ctx->stream>>>(
        c, x, ih[0], ih[1], ih[2], kappa, in->data, out->data);
  return launch_check(ctx, "laplacian");
}

int lgd_central_laplacian(lgd_ctx *ctx, const lgd_grid *grid, const lgd_buf *in, lgd_buf *out,
                          uint32_t ghost) {
  return laplacian_common(ctx, grid, in, out, ghost, false, 1.0);
}

int lgd_heat_rhs(lgd_ctx *ctx, const lgd_grid *grid, double kappa, const lgd_buf *u, lgd_buf *k) {
  return laplacian_common(ctx, grid, u, k, 1, true, kappa);
}

int lgd_karma_rappel_flux(lgd_ctx *ctx, const lgd_grid *grid, const lgd_buf *phi, lgd_buf *out,
                          double eps4, double tol) {
  if (!ctx || !grid) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (grid->cart.dim != 2) return fail(ctx, LGD_ERR_UNSUPPORTED, "the Karma–Rappel flux is 2-D");
  int rc;
  if ((rc = check_field(ctx, grid, phi, 1, "phi"))) return rc;
  if ((rc = check_field(ctx, grid, out, 1, "out"))) return rc;
  lgd_modelc_params p{eps4, 0.0, 0.0, 0.0, tol, tol};
  lgd_kr_consts k = lgd_make_kr_consts(p, grid->cart);
  int64_t total = grid->cart.nblocks * grid->cart.n[0] * grid->cart.n[1];
  kr_flux_kernel<<<blocks_for(total, kWG), kWG, 0, ctx->stream>>>(grid->cart, k, phi->data,
                                                                out->data);
  return launch_check(ctx, "karma_rappel_flux");
}

int lgd_modelc_rhs(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                   const lgd_buf *phi, const lgd_buf *u, lgd_buf *kphi, lgd_buf *ku) {
  if (!ctx || !grid || !params) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (grid->cart.dim != 2) return fail(ctx, LGD_ERR_UNSUPPORTED, "Model C is 2-D");
  int rc;
  if ((rc = check_field(ctx, grid, phi, 1, "phi"))) return rc;
  if ((rc = check_field(ctx, grid, u, 1, "u"))) return rc;
  if ((rc = check_field(ctx, grid, kphi, 1, "kphi"))) return rc;
  if ((rc = check_field(ctx, grid, ku, 1, "ku"))) return rc;
  lgd_kr_consts k = lgd_make_kr_consts(*params, grid->cart);
  int64_t total = grid->cart.nblocks * grid->cart.n[0] * grid->cart.n[1];
  modelc_rhs_kernel<<<blocks_for(total, kWG), kWG, 0, ctx->stream>>>(
      grid->cart, k, phi->data, u->data, kphi->data, ku->data, nullptr, nullptr);
  return launch_check(ctx, "modelc_rhs");
}

int lgd_modelc_rhs_oriented(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                            const lgd_buf *phi, const lgd_buf *u, const lgd_buf *cos0,
                            const lgd_buf *sin0, lgd_buf *kphi, lgd_buf *ku) {
  if (!ctx || !grid || !params) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (grid->cart.dim != 2) return fail(ctx, LGD_ERR_UNSUPPORTED, "Model C is 2-D");
  int rc;
  if ((rc = check_field(ctx, grid, phi, 1, "phi"))) return rc;
  if ((rc = check_field(ctx, grid, u, 1, "u"))) return rc;
  if ((rc = check_field(ctx, grid, cos0, 0, "cos0"))) return rc;
  if ((rc = check_field(ctx, grid, sin0, 0, "sin0"))) return rc;
  if ((rc = check_field(ctx, grid, kphi, 1, "kphi"))) return rc;
  if ((rc = check_field(ctx, grid, ku, 1, "ku"))) return rc;
  lgd_kr_consts k = lgd_make_kr_consts(*params, grid->cart);
  int64_t total = grid->cart.nblocks * grid->cart.n[0] * grid->cart.n[1];
  modelc_rhs_kernel<<<blocks_for(total, kWG), kWG, 0, ctx->stream>>>(
      grid->cart, k, phi->data, u->data, kphi->data, ku->data, cos0->data, sin0->data);
  return launch_check(ctx, "modelc_rhs_oriented");
}

int lgd_modelc_amplitude(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                         lgd_buf *amp_phi) {
  if (!ctx || !grid || !params) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  int rc;
  if ((rc = check_field(ctx, grid, amp_phi, 1, "amp_phi"))) return rc;
  Ext x = make_ext(grid->cart, 1);
  int64_t total = (int64_t)amp_phi->len;
  modelc_amp_kernel<<<blocks_for(total, kWG), kWG, 0, ctx->stream>>>(
      grid->cart, x, params->noise_amplitude, amp_phi->data);
  return launch_check(ctx, "modelc_amplitude");
}

int lgd_axpy(lgd_ctx *ctx, lgd_buf *x, const lgd_buf *y, double alpha) {
  if (!ctx || !x || !y) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  // A zero-length slab on either side is a field the buffer does not carry
  // (src/core/state.rs:491-493).
  if (x->len == 0 || y->len == 0) return LGD_OK;
  if (x->len != y->len) return fail(ctx, LGD_ERR_SIZE_MISMATCH, "axpy between unequal buffers");
  axpy_kernel<<<blocks_for((int64_t)x->len, kWG), kWG, 0, ctx->stream>>>((int64_t)x->len, alpha,
                                                                       x->data, y->data);
  return launch_check(ctx, "axpy");
}

int lgd_apply_wiener(lgd_ctx *ctx, const lgd_grid *grid, lgd_buf *x, const lgd_buf *amp,
                     uint32_t ghost, double dt, uint64_t seed, uint64_t salt, uint64_t stream) {
  if (!ctx || !grid) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (!x || !amp) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null buffer");
  if (x->len == 0 || amp->len == 0) return LGD_OK;
  int rc;
  if ((rc = check_field(ctx, grid, x, ghost, "x"))) return rc;
  if ((rc = check_field(ctx, grid, amp, ghost, "amp"))) return rc;
  Ext e = make_ext(grid->cart, ghost);
  double scale = std::sqrt(dt);
  wiener_kernel<<<blocks_for((int64_t)x->len, kWG), kWG, 0, ctx->stream>>>(
      grid->cart, e, scale, seed, salt, ctx->salt_src, stream, x->data, amp->data);
  return launch_check(ctx, "apply_wiener");
}

}  // extern "C"
