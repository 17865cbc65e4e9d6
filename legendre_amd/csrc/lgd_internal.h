// Internal declarations shared by the legendre_gpu translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstring>
#include <map>
#include <string>
#include <tuple>
#include <vector>

#include "../../include/legendre_gpu.h"
#include "lgd_trig.h"

// A launch shape of the fused Model C step: kernel 1 = one thread per cell,
// 2 = streaming (strips per block, rows loaded ahead).
struct lgd_launch_shape {
  int32_t kernel = 2;
  int32_t strips = 4;
  int32_t prefetch = 2;
};

struct lgd_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  bool owns_stream = false;
  int32_t num_cus = 0;  // compute units of the device (0 = unknown)
  // Set while a simulation captures a graph: every noise launch reads its salt from
  // this device time instead of the value argument, so replays draw fresh noise.
  const double *salt_src = nullptr;
  int64_t tuning_epoch = 0;  // bumped by every tuning change (captured graphs go stale)
  std::string last_error;
  // Launch-shape knobs (lgd_ctx_set_tuning); 0 = pick automatically.
  int64_t modelc_kernel = 0;  // 1 = per-cell kernel, 2 = streaming kernel
  int64_t modelc_strips = 0;  // strips (row ranges) per block for the streaming kernel
  int64_t modelc_prefetch = 0;  // rows the streaming kernel loads ahead (1 or 2; 0 = default)
  int64_t amr_per_block = 0;    // 1 = AMR level kernels launched once per patch (baseline)
  int64_t amr_split_fill = 0;   // 1 = every AMR fill and restriction in a launch of its own
  int64_t amr_tile = 0;         // AMR level-stage tile shape (0 = the plans' 64×4)
  int64_t amr_wg = 0;           // AMR fill / restriction workgroup width (0 = 256)
  int64_t amr_kernel = 0;       // AMR Model C level stage: 0/2 = streaming, 1 = per-cell tiles
  int64_t amr_rows = 0;         // rows per strip of the streaming level stage (0 = by level)
  int64_t adjoint_kernel = 0;   // 1 = three-pass reverse step, 2 = streaming required
  int64_t adjoint_rows = 0;     // rows per strip of the streaming reverse step (0 = by device)
  int64_t regrid_host = 0;      // 1 = AMR regrid tags and migrates on the host
  // Pinned host staging of AMR plan uploads (grown on demand) and the event recorded
  // after the last copy out of it.
  char *pinned = nullptr;
  size_t pinned_cap = 0;
  hipEvent_t pinned_done = nullptr;
  // Device arenas of retired AMR hierarchies (pointer, bytes), reused by the next
  // hierarchy's plans in stream order instead of allocating again.
  std::vector<std::pair<void *, size_t>> arenas;
  // Shapes chosen by lgd_autotune_modelc_step, keyed by (nx, ny, blocks, noise).
  std::map<std::tuple<int64_t, int64_t, int64_t, int>, lgd_launch_shape> tuned;
};

// A device arena of at least `bytes` for work on the context's stream: a retired
// one (stream-ordered after its last reader) or a new one with a quarter to spare;
// *cap receives its size.  Returned with lgd_arena_give.
hipError_t lgd_arena_take(lgd_ctx *ctx, size_t bytes, void **out, size_t *cap);
void lgd_arena_give(lgd_ctx *ctx, void *p, size_t cap);

// The launch shape for a Model C step over `nblocks` blocks of nx × ny: explicit
// knobs first (over the fields of `dflt` they leave unset), then the autotuned
// shape for this key, then `dflt`.
lgd_launch_shape lgd_modelc_shape(const lgd_ctx *ctx, int64_t nx, int64_t ny, int64_t nblocks,
                                  bool noise, const lgd_launch_shape &dflt);
// The untuned shape of the single-grid Model C step: streaming, 4 strips per block,
// two rows loaded ahead, or one with noise (VALU-bound: the registers of the second
// row are worth more as headroom).
inline lgd_launch_shape lgd_modelc_default_shape(bool noise) {
  return lgd_launch_shape{2, 4, noise ? 1 : 2};
}

struct lgd_buf {
  lgd_ctx *ctx = nullptr;
  double *data = nullptr;
  size_t len = 0;  // elements in use
  size_t cap = 0;  // elements allocated (>= len; an AMR simulation reuses buffers)
};

// Host and device view of a CartesianGrid<D> (src/geometry/cartesian.rs:28-35).
struct lgd_cart {
  int32_t dim;
  int32_t periodic[3];
  int64_t n[3];    // block_cells, 1 for unused dims
  int64_t nb[3];   // blocks_per_dim, 1 for unused dims
  double h[3];     // spacing
  double origin[3];
  int64_t nblocks;
};

struct lgd_grid {
  lgd_ctx *ctx = nullptr;
  lgd_cart cart{};
};

namespace lgd {

// Record an error message on the context and return the status.
int fail(lgd_ctx *ctx, int status, const std::string &msg);
int hip_fail(lgd_ctx *ctx, hipError_t e, const char *what);

inline int64_t slab_len(const lgd_cart &c, uint32_t ghost) {
  int64_t len = 1;
  for (int d = 0; d < c.dim; ++d) len *= c.n[d] + 2 * (int64_t)ghost;
  return len;
}

// Validate that `buf` holds exactly one slab of ghost width `ghost` per block.
int check_field(lgd_ctx *ctx, const lgd_grid *grid, const lgd_buf *buf, uint32_t ghost,
                const char *name);
int check_ghost(lgd_ctx *ctx, const lgd_grid *grid, uint32_t ghost);

// Number of workgroups for `n` items at `wg` items each.
inline unsigned blocks_for(int64_t n, int wg) { return (unsigned)((n + wg - 1) / wg); }

inline int launch_check(lgd_ctx *ctx, const char *what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(ctx, e, what);
  return LGD_OK;
}

}  // namespace lgd

// ---------------------------------------------------------------- device RNG
// SplitMix64 finaliser, key chain, 53-bit open uniform and Box–Muller normal
// (src/util/rng.rs:13-49).  Integer arithmetic wraps mod 2^64 as in the reference.
__host__ __device__ inline uint64_t lgd_splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

__host__ __device__ inline uint64_t lgd_mix_key3(uint64_t seed, uint64_t a, uint64_t b,
                                                 uint64_t c) {
  uint64_t s = lgd_splitmix64(seed);
  s = lgd_splitmix64(s ^ a);
  s = lgd_splitmix64(s ^ b);
  return lgd_splitmix64(s ^ c);
}

// The noise salt, t.to_bits() (src/core/simulation.rs:271-281): the value passed,
// or the bits of the time at `src` when a graph-captured step keeps its time in
// device memory.
__device__ inline uint64_t lgd_salt(uint64_t salt, const double *src) {
  return src ? (uint64_t)__double_as_longlong(*src) : salt;
}

__host__ __device__ inline double lgd_unit_open(uint64_t bits) {
  return (double)((bits >> 11) + 1) * (1.0 / 9007199254740992.0);
}

// The logarithm and the angle's cosine are the reduced-range lgd_log_unit and
// lgd_cos_angle (lgd_trig.h): u is a positive normal number no greater than 1 and
// the angle never exceeds 2π, so the library functions' special-case and
// large-argument paths are dead code that costs instructions and registers.
__device__ inline double lgd_standard_normal(uint64_t key) {
  uint64_t a = lgd_splitmix64(key);
  uint64_t b = lgd_splitmix64(a ^ 0xD1B54A32D192ED03ull);
  double u = lgd_unit_open(a);
  double v = lgd_unit_open(b);
  return sqrt(-2.0 * lgd_log_unit(u)) * lgd_cos_angle(6.283185307179586 * v);
}

// --------------------------------------------------- Karma–Rappel constants
// KarmaRappelFlux::new / ModelC::new, computed once on the host
// (src/discretization/finite_volume.rs:113-122, src/physics/phasefield.rs:100-116).
struct lgd_kr_consts {
  double a_bar;      // fma(3, -eps4, 1)
  double eps_prime;  // 3*eps4 / a_bar
  double a12;        // 4*a_bar*eps_prime
  double fv_tol;     // FiniteVolume::tol
  double model_tol;  // ModelC::anisotropy_tol
  double inv_h2;     // 1/(hx*hy)
  double inv_hx2;    // 1/(hx*hx)  (CentralLaplacian on u)
  double inv_hy2;    // 1/(hy*hy)
  double lambda;
  double d_thermal;
};

lgd_kr_consts lgd_make_kr_consts(const lgd_modelc_params &p, const lgd_cart &c);

// One independent Cartesian domain of the streaming Model C kernel: its
// constants, Wiener amplitude scale (sqrt(dt)*b, 0 = none), noise seed, first
// block and block topology (mirror or periodic faces).
struct lgd_domain {
  lgd_kr_consts k;
  double noise_scale;
  uint64_t seed;
  int64_t block0;
  int32_t nbx, nby, per_x, per_y;
};

// The fused Model C step (lgd_modelc_step) with an optional orientation: cs0, sn0 =
// (cos θ₀, sin θ₀) per cell in ghost-0 slabs (null: axis-aligned).
// The snapshot observer (lgd_observer.hip) as the simulation drives it: whether
// `step` snapshots, steps from `step` to the next snapshot step, and the simulation
// fields the observer reads.
namespace lgd {
bool observer_due(const lgd_observer *o, uint64_t step);
int64_t observer_steps_to_due(const lgd_observer *o, uint64_t step);
const std::vector<int32_t> &observer_fields(const lgd_observer *o);
const std::vector<int32_t> &observer_statics(const lgd_observer *o);
int observer_prepare(lgd_observer *o, const lgd_grid *grid, const lgd_amr *amr, uint64_t epoch);
}  // namespace lgd

// One RungeKutta4 step of Model C in four fused stage launches (lgd_fused.hip):
// kbuf = the running sum of the four axpys (φ, u); tbuf = two alternating stage
// inputs (φ, u, φ, u).
int lgd_modelc_rk4_step(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                        lgd_buf *phi, lgd_buf *u, lgd_buf *const *kbuf, lgd_buf *const *tbuf,
                        double dt, const double *cs0, const double *sn0);

int lgd_modelc_step_impl(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                         const lgd_buf *phi, const lgd_buf *u, lgd_buf *phi_next, lgd_buf *u_next,
                         double dt, uint64_t seed, uint64_t salt, int32_t wiener,
                         const double *cs0, const double *sn0);

// The ghost fill of an AMR level at α = 0 from the level a streaming launch steps
// (its parent): the level's fill ops with their four corner slots each (lgd_amr.h)
// and its current buffers (fields 0, 1).  Copies read the level's own buffers,
// prolonged ghosts the corner sum of the launch's inputs (phi, u).
struct lgd_fill_op;
struct lgd_corner;
struct lgd_stream_side_fill {
  const lgd_fill_op *ops;
  const lgd_corner *corners;
  int64_t nops;
  double *lvl[2];
};

// Streaming Model C step (lgd_modelc_stream.hip) over blocks
// [block_begin, block_begin + nblocks); LGD_ERR_UNSUPPORTED when the block is
// wider than 126 interior cells.  `pair` runs two x-neighbouring blocks per wave
// (blocks at most 64 wide; see lgd_stream_pairable).  With `side` (one domain, no
// domain table) the launch's first workgroups write that fill.
int lgd_modelc_stream_launch(lgd_ctx *ctx, int64_t nx, int64_t ny, const lgd_domain &d0,
                             const lgd_domain *domains, const int32_t *block_domain,
                             int64_t block_begin, int64_t nblocks, const double *phi,
                             const double *u, double *phi_n, double *u_n, double dt, bool noise,
                             uint64_t salt, uint64_t stream, int strips, int prefetch,
                             bool pair = false, const lgd_stream_side_fill *side = nullptr);

// Whether a domain of nbx blocks of nx interior columns per row can run as block
// pairs: nx even, 4 <= nx <= 64, nbx even.
bool lgd_stream_pairable(int64_t nx, int64_t nbx);

// Operands of one RungeKutta4 stage on the streaming kernel: y (read by stages
// 2-3; stage 4 writes the result into it), the running sum A of the four axpys
// (stage 1 writes y + w1·k1, stages 2-3 add w_s·k_s, stage 4 reads it), the stage
// coefficient c_s and the stage weight w_s (dt/6, dt/3, dt/3, dt/6).
struct lgd_rk4_stage {
  double *y_phi, *y_u;
  double *acc_phi, *acc_u;
  double coef;
  double w;
};

// RK4 stage `stage` (1..4) of the Model C vector field on the streaming kernel
// over every block of one domain (blocks at most 126 wide, no orientation): the
// stencil reads (pin, uin; stage 1: y); stages 1-3 update the running sum and
// write the next input to (pout, uout). Interior cells only. LGD_ERR_UNSUPPORTED
// when the block is too wide.
int lgd_modelc_stream_rk4_launch(lgd_ctx *ctx, int64_t nx, int64_t ny, const lgd_domain &d0,
                                 int64_t nblocks, const double *pin, const double *uin,
                                 double *pout, double *uout, int stage, const lgd_rk4_stage &rk,
                                 int strips, int prefetch);

// Streaming reverse Model C step (lgd_adjoint_stream.hip): adjoint (APi, AUi) of the
// state after the step → (APo, AUo) before it, global N0 × N1 arrays, each output
// with kAdjStreamPad more doubles after them that the kernel overwrites; θ partials
// (λ, D, ε₄ centre, ε₄ faces) added to acc[4·wave + n]; `rows` rows per strip (at
// least 2; 0 = the height that balances the waves over the device's SIMDs).  With acc == nullptr only *nwaves is set.  LGD_ERR_UNSUPPORTED for a
// grid that is not 2-D or has fewer than 2 cells along a dimension.
constexpr int64_t kAdjStreamPad = 64;
int lgd_adjoint_stream_launch(lgd_ctx *ctx, const lgd_cart &c, const lgd_kr_consts &k, double dt,
                              const double *phi, const double *u, const double *APi,
                              const double *AUi, double *APo, double *AUo, double *acc,
                              int rows, int64_t *nwaves);

// face(dx, dy) (src/discretization/finite_volume.rs:129-142): anisotropy A and its
// angular derivative with the |∇φ|² normalisation folded in.
__device__ __forceinline__ void lgd_kr_face(const lgd_kr_consts &k, double dx, double dy,
                                            double &a, double &da) {
  double dx2 = dx * dx;
  double dy2 = dy * dy;
  double sum = dx2 + dy2;
  double mag2 = sum * sum;
  if (mag2 <= k.fv_tol) {
    a = k.a_bar;
    da = 0.0;
  } else {
    a = k.a_bar * (1.0 + (k.eps_prime * (dx2 * dx2 + dy2 * dy2)) / mag2);
    da = ((((-k.a12) * dx) * dy) * (dx2 - dy2)) / mag2;
  }
}

// ModelC::center_anisotropy (src/physics/phasefield.rs:159-175).
__device__ __forceinline__ double lgd_center_anisotropy(const lgd_kr_consts &k, double pe,
                                                        double pw, double pn, double ps) {
  double dx = 0.5 * (pe - pw);
  double dy = 0.5 * (pn - ps);
  double dx2 = dx * dx;
  double dy2 = dy * dy;
  double sum = dx2 + dy2;
  double mag2 = sum * sum;
  if (mag2 <= k.model_tol) return k.a_bar;
  return k.a_bar * (1.0 + (k.eps_prime * (dx2 * dx2 + dy2 * dy2)) / mag2);
}

// Multi-grain Model C at one interior cell (KarmaRappelFlux::apply_oriented,
// finite_volume.rs:188-235, and the oriented branch of ModelC::vector_field_block,
// phasefield.rs:378-392): every face's gradient, and the central difference, are
// rotated into the crystal frame by θ₀ of this cell — given as (c0, s0) = (cos θ₀,
// sin θ₀), computed once per cell on the host with libm as the reference computes
// them per step — for (A, A′); the fluxes keep the lab-frame components; the centre
// anisotropy is the flux's face() (FV tolerance) of the rotated difference.
__device__ __forceinline__ void lgd_rotate(double c0, double s0, double dx, double dy,
                                           double &rx, double &ry) {
  rx = c0 * dx + s0 * dy;
  ry = c0 * dy - s0 * dx;
}

__device__ __forceinline__ void lgd_modelc_cell_oriented(const lgd_kr_consts &k,
                                                         const double p[3][3], double uc,
                                                         double ue, double uw, double un,
                                                         double us, double c0, double s0,
                                                         double &dphi, double &du) {
  const double q = 0.25;
  double derx_r = p[1][2] - p[1][1];
  double derx_l = p[1][1] - p[1][0];
  double derx_t = q * (((p[2][2] - p[2][0]) + p[1][2]) - p[1][0]);
  double derx_b = q * (((p[1][2] - p[1][0]) + p[0][2]) - p[0][0]);
  double dery_t = p[2][1] - p[1][1];
  double dery_b = p[1][1] - p[0][1];
  double dery_r = q * (((p[2][2] - p[0][2]) + p[2][1]) - p[0][1]);
  double dery_l = q * (((p[2][1] - p[0][1]) + p[2][0]) - p[0][0]);
  double rx, ry, a_r, da_r, a_l, da_l, a_t, da_t, a_b, da_b;
  lgd_rotate(c0, s0, derx_r, dery_r, rx, ry);
  lgd_kr_face(k, rx, ry, a_r, da_r);
  lgd_rotate(c0, s0, derx_l, dery_l, rx, ry);
  lgd_kr_face(k, rx, ry, a_l, da_l);
  lgd_rotate(c0, s0, derx_t, dery_t, rx, ry);
  lgd_kr_face(k, rx, ry, a_t, da_t);
  lgd_rotate(c0, s0, derx_b, dery_b, rx, ry);
  lgd_kr_face(k, rx, ry, a_b, da_b);
  double j_r = a_r * ((a_r * derx_r) - (da_r * dery_r));
  double j_l = a_l * ((a_l * derx_l) - (da_l * dery_l));
  double j_t = a_t * ((a_t * dery_t) + (da_t * derx_t));
  double j_b = a_b * ((a_b * dery_b) + (da_b * derx_b));
  double flux = (((j_r - j_l) + j_t) - j_b) * k.inv_h2;
  double pc = p[1][1];
  double g_prime = fma(pc * pc, pc, -pc);
  double one_m_p2 = fma(pc, -pc, 1.0);
  double p_prime = one_m_p2 * one_m_p2;
  double cdx = 0.5 * (p[1][2] - p[1][0]);
  double cdy = 0.5 * (p[2][1] - p[0][1]);
  double crx, cry, a, da;
  lgd_rotate(c0, s0, cdx, cdy, crx, cry);
  lgd_kr_face(k, crx, cry, a, da);
  dphi = fma(k.lambda * uc, -p_prime, flux - g_prime) / (a * a);
  double acc = 0.0;
  acc = acc + (((ue - (uc + uc)) + uw) * k.inv_hx2);
  acc = acc + (((un - (uc + uc)) + us) * k.inv_hy2);
  du = fma(acc, k.d_thermal, 0.5 * dphi);
}

// Model C right-hand side at one interior cell from its 3x3 phi neighbourhood
// (p[j][i], j = row offset +1, i = column offset +1) and 5-point u neighbourhood:
// the composition KarmaRappelFlux::kernel → pointwise dφ → CentralLaplacian(u) → du
// (finite_volume.rs:258-282, phasefield.rs:394-421), in the reference's order.
__device__ __forceinline__ void lgd_modelc_cell(const lgd_kr_consts &k, const double p[3][3],
                                                double uc, double ue, double uw, double un,
                                                double us, double &dphi, double &du) {
  const double q = 0.25;
  double derx_r = p[1][2] - p[1][1];
  double derx_l = p[1][1] - p[1][0];
  double derx_t = q * (((p[2][2] - p[2][0]) + p[1][2]) - p[1][0]);
  double derx_b = q * (((p[1][2] - p[1][0]) + p[0][2]) - p[0][0]);
  double dery_t = p[2][1] - p[1][1];
  double dery_b = p[1][1] - p[0][1];
  double dery_r = q * (((p[2][2] - p[0][2]) + p[2][1]) - p[0][1]);
  double dery_l = q * (((p[2][1] - p[0][1]) + p[2][0]) - p[0][0]);
  double a_r, da_r, a_l, da_l, a_t, da_t, a_b, da_b;
  lgd_kr_face(k, derx_r, dery_r, a_r, da_r);
  lgd_kr_face(k, derx_l, dery_l, a_l, da_l);
  lgd_kr_face(k, derx_t, dery_t, a_t, da_t);
  lgd_kr_face(k, derx_b, dery_b, a_b, da_b);
  double j_r = a_r * ((a_r * derx_r) - (da_r * dery_r));
  double j_l = a_l * ((a_l * derx_l) - (da_l * dery_l));
  double j_t = a_t * ((a_t * dery_t) + (da_t * derx_t));
  double j_b = a_b * ((a_b * dery_b) + (da_b * derx_b));
  double flux = (((j_r - j_l) + j_t) - j_b) * k.inv_h2;
  double pc = p[1][1];
  double g_prime = fma(pc * pc, pc, -pc);
  double one_m_p2 = fma(pc, -pc, 1.0);
  double p_prime = one_m_p2 * one_m_p2;
  double a = lgd_center_anisotropy(k, p[1][2], p[1][0], p[2][1], p[0][1]);
  dphi = fma(k.lambda * uc, -p_prime, flux - g_prime) / (a * a);
  double acc = 0.0;
  acc = acc + (((ue - (uc + uc)) + uw) * k.inv_hx2);
  acc = acc + (((un - (uc + uc)) + us) * k.inv_hy2);
  du = fma(acc, k.d_thermal, 0.5 * dphi);
}
