// Fused hot-path kernels: one launch per time step.
//
// Halos are not materialised: a cell outside the block is read straight from the
// neighbour's interior (exchange / periodic wrap) or from this block's own
// interior (mirror), which is the value fill_ghosts_mirror would have written
// there (src/geometry/cartesian.rs:439-526) — for copy-type rules the dimension
// sweep resolves every ghost, corners included, to the cell at the per-dimension
// reflected/wrapped global coordinate.  The RHS and the stage update are fused, and
// the result goes to a second buffer pair (neighbours still read the old values).
#include <cmath>

#include <type_traits>

#include "lgd_internal.h"

namespace {

// Offset from a cell to its neighbour one step along one dimension (dir = ±1): the
// next slab entry inside the block, the adjacent block's edge cell across a block face,
// the wrapped block's edge cell at a periodic face, or the cell itself at a mirror face
// (reflect_coord, src/geometry/boundary.rs:82-90). `l` is the cell's index in its block
// of n cells, `bc` the block's index among nb, `stride` / `bstride` the element
// distances of one cell / one block along the dimension.
__device__ __forceinline__ int64_t neighbour_delta(int64_t l, int64_t n, int64_t bc, int64_t nb,
                                                   int periodic, int64_t stride, int64_t bstride,
                                                   int dir) {
  const int64_t span = (n - 1) * stride;
  if (dir > 0) {
    if (l + 1 < n) return stride;
    if (bc + 1 < nb) return bstride - span;
    return periodic ? -(nb - 1) * bstride - span : 0;
  }
  if (l > 0) return -stride;
  if (bc > 0) return -bstride + span;
  return periodic ? (nb - 1) * bstride + span : 0;
}

// The Model C vector field (dφ, du) of interior cell (i, j) of block b from (phi, u),
// halos read in place; also the cell's element offset and its φ and u.
template <bool kOriented>
__device__ __forceinline__ void cell_vector_field(const lgd_cart &c, const lgd_kr_consts &k,
                                                  int64_t b, int64_t i, int64_t j,
                                                  const double *phi, const double *u,
                                                  const double *cs0, const double *sn0,
                                                  int64_t &oc, double &pc, double &uc,
                                                  double &dphi, double &du) {
  const int64_t bx = b % c.nb[0], by = b / c.nb[0];
  // Element offsets of the neighbours, one dimension at a time (the resolved global
  // coordinates of the halo are per-dimension reflections / wraps, so a corner
  // neighbour composes the two one-dimensional moves).
  const int64_t sx = c.n[0] + 2, slab = sx * (c.n[1] + 2);
  oc = b * slab + (j + 1) * sx + (i + 1);
  const int64_t dxm = neighbour_delta(i, c.n[0], bx, c.nb[0], c.periodic[0], 1, slab, -1);
  const int64_t dxp = neighbour_delta(i, c.n[0], bx, c.nb[0], c.periodic[0], 1, slab, +1);
  const int64_t dym = neighbour_delta(j, c.n[1], by, c.nb[1], c.periodic[1], sx, c.nb[0] * slab, -1);
  const int64_t dyp = neighbour_delta(j, c.n[1], by, c.nb[1], c.periodic[1], sx, c.nb[0] * slab, +1);
  const int64_t dx[3] = {dxm, 0, dxp}, dy[3] = {dym, 0, dyp};
  double p[3][3];
  for (int dj = 0; dj < 3; ++dj)
    for (int di = 0; di < 3; ++di) p[dj][di] = phi[oc + dx[di] + dy[dj]];
  uc = u[oc];
  const double ue = u[oc + dxp], uw = u[oc + dxm];
  const double un = u[oc + dyp], us = u[oc + dym];
  pc = p[1][1];
  if (kOriented) {
    const int64_t ot = b * c.n[0] * c.n[1] + j * c.n[0] + i;
    lgd_modelc_cell_oriented(k, p, uc, ue, uw, un, us, cs0[ot], sn0[ot], dphi, du);
  } else {
    lgd_modelc_cell(k, p, uc, ue, uw, un, us, dphi, du);
  }
}

constexpr int kTX = 64;
constexpr int kTY = 4;

// Model C Euler–Maruyama step, one thread per interior cell (2-D).
// Workgroups tile one block at a time: blockIdx.x = block * tiles + tile.
// kOriented: the multi-grain vector field with (cos θ₀, sin θ₀) per cell in ghost-0
// slabs (cs0, sn0).
template <bool kNoise, bool kOriented>
__global__ __launch_bounds__(kTX *kTY) void modelc_step_kernel(
    lgd_cart c, lgd_kr_consts k, int64_t tiles_x, int64_t tiles_per_block, const double *phi,
    const double *u, double *phi_n, double *u_n, double dt, double noise_scale, uint64_t seed,
    uint64_t salt, const double *salt_src, uint64_t stream, const double *cs0, const double *sn0) {
  int64_t b = blockIdx.x / tiles_per_block;
  int64_t tile = blockIdx.x - b * tiles_per_block;
  int64_t ty = tile / tiles_x, tx = tile - ty * tiles_x;
  int64_t i = tx * kTX + threadIdx.x;
  int64_t j = ty * kTY + threadIdx.y;
  if (i >= c.n[0] || j >= c.n[1]) return;
  int64_t oc;
  double pc, uc, dphi, du;
  cell_vector_field<kOriented>(c, k, b, i, j, phi, u, cs0, sn0, oc, pc, uc, dphi, du);
  // apply_step_with: Time first (x += dt*k), then Wiener(0) on phi
  // (src/core/state.rs:656-699, src/core/driver.rs:131-158).
  double pn = pc + dt * dphi;
  if (kNoise) {
    uint64_t block_key = lgd_mix_key3(seed, lgd_salt(salt, salt_src), stream, (uint64_t)b);
    uint64_t key = lgd_splitmix64(block_key ^ (uint64_t)(i + j * c.n[0]));
    pn = pn + noise_scale * lgd_standard_normal(key);
  }
  phi_n[oc] = pn;
  u_n[oc] = uc + dt * du;
}

// One stage of RungeKutta4::step on Model C (src/integrators/rk4.rs:30-80) in one pass,
// reading the stage input's halos in place.  The four axpys y + w1·k1 + w2·k2 +
// w3·k3 + w4·k4 (w = dt/6, dt/3, dt/3, dt/6) are summed as the stages go, in their
// order, in a running sum A:
// stage 1 writes A = y + w1·k1, stages 2-3 add w_s·k_s, stage 4 writes
// y = A + w4·k4.  That is the reference's fold, operation for operation, without
// keeping k1..k3.
struct Rk4Args {
  const double *pin, *uin;    // the stage's input state (stage 1: y itself)
  double *p0, *u0;            // y (read by stages 2-3; stage 4 writes the result here)
  double *ap, *au;            // the running sum A
  double *pout, *uout;        // stages 1-3: the next stage's input y + c_s·k_s
  double coef;                // stages 1-3: c_s
  double w;                   // w_s
};

template <int kStage, bool kOriented>
__global__ __launch_bounds__(kTX *kTY) void rk4_stage_kernel(lgd_cart c, lgd_kr_consts k,
                                                             int64_t tiles_x, int64_t tiles_per_block,
                                                             Rk4Args a, const double *cs0,
                                                             const double *sn0) {
  int64_t b = blockIdx.x / tiles_per_block;
  int64_t tile = blockIdx.x - b * tiles_per_block;
  int64_t ty = tile / tiles_x, tx = tile - ty * tiles_x;
  int64_t i = tx * kTX + threadIdx.x;
  int64_t j = ty * kTY + threadIdx.y;
  if (i >= c.n[0] || j >= c.n[1]) return;
  int64_t oc;
  double pc, uc, dphi, du;
  cell_vector_field<kOriented>(c, k, b, i, j, a.pin, a.uin, cs0, sn0, oc, pc, uc, dphi, du);
  if (kStage == 1) {  // the input is y: (pc, uc) are y at this cell
    a.ap[oc] = pc + a.w * dphi;
    a.au[oc] = uc + a.w * du;
    a.pout[oc] = pc + a.coef * dphi;
    a.uout[oc] = uc + a.coef * du;
  } else if (kStage < 4) {
    a.ap[oc] = a.ap[oc] + a.w * dphi;
    a.au[oc] = a.au[oc] + a.w * du;
    a.pout[oc] = a.p0[oc] + a.coef * dphi;
    a.uout[oc] = a.u0[oc] + a.coef * du;
  } else {
    a.p0[oc] = a.ap[oc] + a.w * dphi;
    a.u0[oc] = a.au[oc] + a.w * du;
  }
}

// The face rules of the six physical faces, low then high per dimension.
struct FaceRules {
  lgd_face_bc bc[6];
};

// FaceBc::ghost_value at layer 0 (src/geometry/boundary.rs:63-73) from the edge cell
// i0 and the next cell inward i1 (read only by Extrapolate).
__device__ __forceinline__ double face_rule(const lgd_face_bc &bc, double h, double i0,
                                            const double *i1) {
  switch (bc.kind) {
    case LGD_BC_DIRICHLET:
      return (2.0 * bc.value) - i0;
    case LGD_BC_FLUX:
      return i0 + ((1.0 * h) * bc.value);
    case LGD_BC_EXTRAPOLATE:
      return (2.0 * i0) - (1.0 * *i1);
    default:
      return i0;
  }
}

// Heat<D> forward-Euler step (examples/heat3d.rs:63-79, src/integrators/euler.rs:128-129):
// u' = u + dt*(κ·∇²u), one thread per interior cell, any dimension 1..3.
// kRules: a neighbour across a non-periodic domain face is the face rule's ghost
// value. The stencil reads face neighbours only, so the value is the one the
// dimension sweep writes there from the edge cells along the normal.
// Workgroups tile one block (blockIdx.x = block · tiles + tile), so block-local indices
// use 32-bit arithmetic and the block's coordinates are wave-uniform. A neighbour is
// the slab entry one stride away inside the block; across a block face it is the
// adjacent block's edge cell (the next block, or the wrapped one at a periodic face);
// at a domain face it is the face rule's ghost value (the edge cell itself under
// Mirror). These are the cells the reflected / wrapped global coordinate resolves to.
template <bool kRules>
__global__ __launch_bounds__(256) void heat_step_kernel(lgd_cart c, double ih0, double ih1,
                                                        double ih2, double kappa, double dt,
                                                        const double *u, double *u_n,
                                                        FaceRules rules, uint32_t tiles) {
  const uint32_t b = blockIdx.x / tiles;
  const uint32_t n0 = (uint32_t)c.n[0], n1 = (uint32_t)c.n[1], n2 = (uint32_t)c.n[2];
  const uint32_t r = (blockIdx.x - b * tiles) * 256u + threadIdx.x;
  if (r >= n0 * n1 * n2) return;
  const uint32_t q = r / n0;
  const uint32_t l[3] = {r - q * n0, q % n1, q / n1};
  // Slab strides of dimensions 0, 1, 2 (unused dimensions have no ghost layers).
  const int64_t sx = c.dim > 0 ? (int64_t)n0 + 2 : 1;
  const int64_t sy = c.dim > 1 ? (int64_t)n1 + 2 : 1;
  const int64_t sz = c.dim > 2 ? (int64_t)n2 + 2 : 1;
  const int64_t slab = sx * sy * sz;
  const int64_t stride[3] = {1, sx, sx * sy};
  const int64_t bstride[3] = {slab, c.nb[0] * slab, c.nb[0] * c.nb[1] * slab};  // one block over
  int64_t oc = (int64_t)b * slab + (int64_t)(l[0] + 1);
  if (c.dim > 1) oc += (int64_t)(l[1] + 1) * stride[1];
  if (c.dim > 2) oc += (int64_t)(l[2] + 1) * stride[2];
  const double center = u[oc];
  const int64_t bc[3] = {(int64_t)b % c.nb[0], ((int64_t)b / c.nb[0]) % c.nb[1],
                         (int64_t)b / (c.nb[0] * c.nb[1])};
  const double inv[3] = {ih0, ih1, ih2};
  double acc = 0.0;
  for (int d = 0; d < c.dim; ++d) {
    const int64_t nd = d == 0 ? n0 : d == 1 ? n1 : n2;
    const int64_t span = (nd - 1) * stride[d];  // edge cell to edge cell across the block
    double up, um;
    if (l[d] + 1 < nd) {
      up = u[oc + stride[d]];
    } else if (bc[d] + 1 < c.nb[d]) {
      up = u[oc + bstride[d] - span];
    } else if (c.periodic[d]) {
      up = u[oc - (c.nb[d] - 1) * bstride[d] - span];
    } else if (kRules) {
      const double i1 = rules.bc[2 * d + 1].kind == LGD_BC_EXTRAPOLATE ? u[oc - stride[d]] : 0.0;
      up = face_rule(rules.bc[2 * d + 1], c.h[d], center, &i1);
    } else {
      up = center;  // Mirror
    }
    if (l[d] > 0) {
      um = u[oc - stride[d]];
    } else if (bc[d] > 0) {
      um = u[oc - bstride[d] + span];
    } else if (c.periodic[d]) {
      um = u[oc + (c.nb[d] - 1) * bstride[d] + span];
    } else if (kRules) {
      const double i1 = rules.bc[2 * d].kind == LGD_BC_EXTRAPOLATE ? u[oc + stride[d]] : 0.0;
      um = face_rule(rules.bc[2 * d], c.h[d], center, &i1);
    } else {
      um = center;  // Mirror
    }
    acc = acc + (((up - (center + center)) + um) * inv[d]);
  }
  u_n[oc] = center + dt * (acc * kappa);
}

}  // namespace

using namespace lgd;

int lgd_modelc_step_impl(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                         const lgd_buf *phi, const lgd_buf *u, lgd_buf *phi_next, lgd_buf *u_next,
                         double dt, uint64_t seed, uint64_t salt, int32_t wiener,
                         const double *cs0, const double *sn0) {
  if (!ctx || !grid || !params) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  const lgd_cart &c = grid->cart;
  if (c.dim != 2) return fail(ctx, LGD_ERR_UNSUPPORTED, "Model C is 2-D");
  int rc;
  if ((rc = check_field(ctx, grid, phi, 1, "phi"))) return rc;
  if ((rc = check_field(ctx, grid, u, 1, "u"))) return rc;
  if ((rc = check_field(ctx, grid, phi_next, 1, "phi_next"))) return rc;
  if ((rc = check_field(ctx, grid, u_next, 1, "u_next"))) return rc;
  if (phi_next == phi || u_next == u || phi_next == u || u_next == phi || phi_next == u_next)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT,
                "the fused step reads neighbours' old values: output buffers must be distinct "
                "from the inputs");
  lgd_kr_consts k = lgd_make_kr_consts(*params, c);
  // Wiener amplitude b on interior cells; a zero amplitude draws nothing
  // (src/core/driver.rs:144-146).
  const bool noise_on = wiener && params->noise_amplitude != 0.0;
  const double noise_scale = std::sqrt(dt) * params->noise_amplitude;
  const uint64_t wiener0 = (1ull << 32) | 0ull;  // Driver::Wiener(0).stream()
  const lgd_launch_shape shape = lgd_modelc_shape(ctx, c.n[0], c.n[1], c.nblocks, noise_on,
                                                  lgd_modelc_default_shape(noise_on));
  const bool oriented = cs0 != nullptr;
  if (shape.kernel != 1 && !oriented) {
    lgd_domain d0;
    d0.k = k;
    d0.noise_scale = noise_on ? noise_scale : 0.0;
    d0.seed = seed;
    d0.block0 = 0;
    d0.nbx = (int32_t)c.nb[0];
    d0.nby = (int32_t)c.nb[1];
    d0.per_x = c.periodic[0];
    d0.per_y = c.periodic[1];
    // Block pairs where the grid allows them (kernel 3), else one block per wave.
    const bool pair = shape.kernel == 3 && lgd_stream_pairable(c.n[0], c.nb[0]);
    rc = lgd_modelc_stream_launch(ctx, c.n[0], c.n[1], d0, nullptr, nullptr, 0, c.nblocks,
                                  phi->data, u->data, phi_next->data, u_next->data, dt, noise_on,
                                  salt, wiener0, shape.strips, shape.prefetch, pair);
    if (rc != LGD_ERR_UNSUPPORTED) return rc;
    if (ctx->modelc_kernel >= 2)
      return fail(ctx, LGD_ERR_UNSUPPORTED, "streaming kernel needs blocks at most 126 cells wide");
  }
  int64_t tiles_x = (c.n[0] + kTX - 1) / kTX;
  int64_t tiles_y = (c.n[1] + kTY - 1) / kTY;
  int64_t tiles = tiles_x * tiles_y;
  int64_t nwg = c.nblocks * tiles;
  if (nwg > (int64_t)INT32_MAX) return fail(ctx, LGD_ERR_UNSUPPORTED, "grid too large for one launch");
  dim3 wg(kTX, kTY);
  auto go = [&](auto noise_t, auto orient_t) {
    modelc_step_kernel<decltype(noise_t)::value, decltype(orient_t)::value>
        <<<(unsigned)nwg, wg, 0, ctx->stream>>>(c, k, tiles_x, tiles, phi->data, u->data,
                                                phi_next->data, u_next->data, dt, noise_scale,
                                                seed, salt, noise_on ? ctx->salt_src : nullptr,
                                                wiener0, cs0, sn0);
  };
  using T = std::true_type;
  using F = std::false_type;
  if (noise_on && oriented)
    go(T{}, T{});
  else if (noise_on)
    go(T{}, F{});
  else if (oriented)
    go(F{}, T{});
  else
    go(F{}, F{});
  return launch_check(ctx, "modelc_step");
}

// One RungeKutta4 step of Model C as four fused stage launches. y = (phi, u);
// kbuf[0..1] hold the running sum A of the four axpys (φ, u); t0, t1 alternate as
// stage inputs. cs0 / sn0: orientations.
int lgd_modelc_rk4_step(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                        lgd_buf *phi, lgd_buf *u, lgd_buf *const *kbuf, lgd_buf *const *tbuf,
                        double dt, const double *cs0, const double *sn0) {
  const lgd_cart &c = grid->cart;
  if (c.dim != 2) return fail(ctx, LGD_ERR_UNSUPPORTED, "Model C is 2-D");
  const lgd_kr_consts k = lgd_make_kr_consts(*params, c);
  const int64_t tiles_x = (c.n[0] + kTX - 1) / kTX;
  const int64_t tiles = tiles_x * ((c.n[1] + kTY - 1) / kTY);
  const int64_t nwg = c.nblocks * tiles;
  if (nwg > (int64_t)INT32_MAX) return fail(ctx, LGD_ERR_UNSUPPORTED, "grid too large for one launch");
  const dim3 wg(kTX, kTY);
  const double half_dt = 0.5 * dt;
  const double coef[3] = {half_dt, half_dt, dt};
  const double w[4] = {dt / 6.0, dt / 3.0, dt / 3.0, dt / 6.0};
  // kbuf[0], kbuf[1]: the running sum A (φ, u); tbuf: t0φ, t0u, t1φ, t1u.
  Rk4Args a{};
  a.p0 = phi->data;
  a.u0 = u->data;
  a.ap = kbuf[0]->data;
  a.au = kbuf[1]->data;
  const bool oriented = cs0 != nullptr;
  // Unoriented grids with blocks at most 126 wide run the stages on the streaming
  // kernel (face sharing; bitwise the same vector field), unless the per-cell
  // kernel is asked for.
  const lgd_launch_shape shape = lgd_modelc_shape(ctx, c.n[0], c.n[1], c.nblocks, false,
                                                  lgd_launch_shape{});
  if (!oriented && shape.kernel != 1 && c.n[0] + 2 <= 128) {
    lgd_domain d0;
    d0.k = k;
    d0.noise_scale = 0.0;
    d0.seed = 0;
    d0.block0 = 0;
    d0.nbx = (int32_t)c.nb[0];
    d0.nby = (int32_t)c.nb[1];
    d0.per_x = c.periodic[0];
    d0.per_y = c.periodic[1];
    lgd_rk4_stage r{};
    r.y_phi = phi->data;
    r.y_u = u->data;
    r.acc_phi = kbuf[0]->data;
    r.acc_u = kbuf[1]->data;
    for (int st = 1; st <= 4; ++st) {
      const lgd_buf *in_p = st == 1 ? phi : tbuf[(st % 2 == 0) ? 0 : 2];
      const lgd_buf *in_u = st == 1 ? u : tbuf[(st % 2 == 0) ? 1 : 3];
      double *out_p = nullptr, *out_u = nullptr;
      r.w = w[st - 1];
      if (st < 4) {
        out_p = tbuf[(st % 2 == 1) ? 0 : 2]->data;
        out_u = tbuf[(st % 2 == 1) ? 1 : 3]->data;
        r.coef = coef[st - 1];
      }
      int rc = lgd_modelc_stream_rk4_launch(ctx, c.n[0], c.n[1], d0, c.nblocks, in_p->data,
                                            in_u->data, out_p, out_u, st, r, shape.strips,
                                            shape.prefetch);
      if (rc) return rc;
    }
    return LGD_OK;
  }
  for (int st = 1; st <= 4; ++st) {
    const lgd_buf *in_p = st == 1 ? phi : tbuf[(st % 2 == 0) ? 0 : 2];
    const lgd_buf *in_u = st == 1 ? u : tbuf[(st % 2 == 0) ? 1 : 3];
    a.pin = in_p->data;
    a.uin = in_u->data;
    a.w = w[st - 1];
    if (st < 4) {
      a.pout = tbuf[(st % 2 == 1) ? 0 : 2]->data;
      a.uout = tbuf[(st % 2 == 1) ? 1 : 3]->data;
      a.coef = coef[st - 1];
    }
    auto go = [&](auto stage_t, auto orient_t) {
      rk4_stage_kernel<decltype(stage_t)::value, decltype(orient_t)::value>
          <<<(unsigned)nwg, wg, 0, ctx->stream>>>(c, k, tiles_x, tiles, a, cs0, sn0);
    };
    using S1 = std::integral_constant<int, 1>;
    using S2 = std::integral_constant<int, 2>;
    using S3 = std::integral_constant<int, 3>;
    using S4 = std::integral_constant<int, 4>;
    using T = std::true_type;
    using F = std::false_type;
    if (oriented) {
      if (st == 1) go(S1{}, T{}); else if (st == 2) go(S2{}, T{}); else if (st == 3) go(S3{}, T{}); else go(S4{}, T{});
    } else {
      if (st == 1) go(S1{}, F{}); else if (st == 2) go(S2{}, F{}); else if (st == 3) go(S3{}, F{}); else go(S4{}, F{});
    }
    int rc = launch_check(ctx, "modelc_rk4_stage");
    if (rc) return rc;
  }
  return LGD_OK;
}

extern "C" {

int lgd_modelc_step(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                    const lgd_buf *phi, const lgd_buf *u, lgd_buf *phi_next, lgd_buf *u_next,
                    double dt, uint64_t seed, uint64_t salt, int32_t wiener) {
  return lgd_modelc_step_impl(ctx, grid, params, phi, u, phi_next, u_next, dt, seed, salt, wiener,
                              nullptr, nullptr);
}


int lgd_heat_step(lgd_ctx *ctx, const lgd_grid *grid, double kappa, const lgd_buf *u,
                  lgd_buf *u_next, double dt) {
  return lgd_heat_step_bc(ctx, grid, kappa, nullptr, u, u_next, dt);
}

int lgd_heat_step_bc(lgd_ctx *ctx, const lgd_grid *grid, double kappa, const lgd_face_bc *bcs,
                     const lgd_buf *u, lgd_buf *u_next, double dt) {
  if (!ctx || !grid) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  const lgd_cart &c = grid->cart;
  int rc;
  if ((rc = check_field(ctx, grid, u, 1, "u"))) return rc;
  if ((rc = check_field(ctx, grid, u_next, 1, "u_next"))) return rc;
  if ((rc = check_ghost(ctx, grid, 1))) return rc;
  if (u == u_next)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "the fused step needs a distinct output buffer");
  double ih[3] = {0.0, 0.0, 0.0};
  for (int d = 0; d < c.dim; ++d) ih[d] = 1.0 / (c.h[d] * c.h[d]);
  FaceRules rules{};
  bool any_rule = false;
  if (bcs) {
    for (int d = 0; d < c.dim; ++d)
      for (int e = 0; e < 2; ++e) {
        const lgd_face_bc &bc = bcs[2 * d + e];
        if (bc.kind < LGD_BC_MIRROR || bc.kind > LGD_BC_EXTRAPOLATE)
          return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "unknown FaceBc kind");
        // Linear extrapolation reads two interior cells (cartesian.rs:418-421).
        if (bc.kind == LGD_BC_EXTRAPOLATE && c.n[d] < 2)
          return fail(ctx, LGD_ERR_INVALID_ARGUMENT,
                      "linear extrapolation reads two interior cells per dimension");
        rules.bc[2 * d + e] = bc;
        any_rule = any_rule || (bc.kind != LGD_BC_MIRROR && !c.periodic[d]);
      }
  }
  const int64_t ncell = c.n[0] * c.n[1] * c.n[2];
  if (ncell > INT32_MAX) return fail(ctx, LGD_ERR_UNSUPPORTED, "a block of more than 2^31 cells");
  const uint32_t tiles = (uint32_t)((ncell + 255) / 256);
  const int64_t wgs = c.nblocks * (int64_t)tiles;
  if (wgs * 256 > (int64_t)UINT32_MAX) return fail(ctx, LGD_ERR_UNSUPPORTED, "grid too large for one launch");
  if (any_rule)
    heat_step_kernel<true><<<(unsigned)wgs, 256, 0, ctx->stream>>>(c, ih[0], ih[1], ih[2], kappa, dt,
                                                                  u->data, u_next->data, rules, tiles);
  else
    heat_step_kernel<false><<<(unsigned)wgs, 256, 0, ctx->stream>>>(c, ih[0], ih[1], ih[2], kappa, dt,
                                                                   u->data, u_next->data, rules, tiles);
  return launch_check(ctx, "heat_step");
}

}  // extern "C"
