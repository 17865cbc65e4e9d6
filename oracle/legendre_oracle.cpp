// legendre CPU oracle — TEST INFRASTRUCTURE ONLY (see legendre_oracle.h).
//
// Restates the reference crate's serial semantics for the block-structured hot
// path.  Each function cites the reference file:line it follows.  Build with
// -ffp-contract=off: the reference (Rust) never contracts a*b+c, and every
// fused multiply-add below is an explicit std::fma standing for a mul_add.
#include "legendre_oracle.h"

#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr double kTau = 6.283185307179586;  // std::f64::consts::TAU

struct Box3 {
  int64_t lo[3];
  int64_t hi[3];
};

inline int dims(const orc_grid *g) { return g->dim; }

inline int64_t blocks_per_dim(const orc_grid *g, int d) {
  return g->cells[d] / g->block_cells[d];
}

// Ghost-inclusive slab extents (unused dims have extent 1).
inline void slab_ext(const orc_grid *g, uint32_t ghost, int64_t ext[3]) {
  for (int d = 0; d < 3; ++d) ext[d] = d < dims(g) ? g->block_cells[d] + 2 * (int64_t)ghost : 1;
}

// linearize (src/geometry/cartesian.rs:189-204).
inline int64_t lin(const orc_grid *g, const int64_t idx[3], uint32_t ghost) {
  int64_t off = 0, stride = 1;
  for (int d = 0; d < dims(g); ++d) {
    off += (idx[d] + (int64_t)ghost) * stride;
    stride *= g->block_cells[d] + 2 * (int64_t)ghost;
  }
  return off;
}

// block_coords (src/geometry/cartesian.rs:145-154).
inline void block_coords(const orc_grid *g, int64_t block, int64_t out[3]) {
  int64_t rem = block;
  for (int d = 0; d < 3; ++d) {
    if (d < dims(g)) {
      int64_t nb = blocks_per_dim(g, d);
      out[d] = rem % nb;
      rem /= nb;
    } else {
      out[d] = 0;
    }
  }
}

// for_each_box, dimension 0 fastest (src/geometry/cartesian.rs:338-358).
template <class F>
void for_each_box(const orc_grid *g, const Box3 &b, F &&f) {
  int D = dims(g);
  for (int d = 0; d < D; ++d)
    if (b.lo[d] >= b.hi[d]) return;
  int64_t lo2 = D > 2 ? b.lo[2] : 0, hi2 = D > 2 ? b.hi[2] : 1;
  int64_t lo1 = D > 1 ? b.lo[1] : 0, hi1 = D > 1 ? b.hi[1] : 1;
  int64_t idx[3];
  for (idx[2] = lo2; idx[2] < hi2; ++idx[2])
    for (idx[1] = lo1; idx[1] < hi1; ++idx[1])
      for (idx[0] = b.lo[0]; idx[0] < b.hi[0]; ++idx[0]) f(idx);
}

// for_each_interior (src/geometry/cartesian.rs:314-334).
template <class F>
void for_each_interior(const orc_grid *g, F &&f) {
  Box3 b;
  for (int d = 0; d < 3; ++d) {
    b.lo[d] = 0;
    b.hi[d] = d < dims(g) ? g->block_cells[d] : 1;
  }
  for_each_box(g, b, f);
}

// FaceBc::ghost_value (src/geometry/boundary.rs:63-73).
template <class F>
inline double ghost_value(const orc_face_bc &bc, int64_t k, double h, F &&interior) {
  switch (bc.kind) {
    case ORC_BC_DIRICHLET:
      return (2.0 * bc.value) - interior(k);
    case ORC_BC_FLUX:
      return interior(k) + (((double)(2 * k + 1) * h) * bc.value);
    case ORC_BC_EXTRAPOLATE:
      return ((double)(k + 2) * interior(0)) - ((double)(k + 1) * interior(1));
    default:
      return interior(k);
  }
}

inline double *slab(const orc_grid *g, double *field, int64_t block, uint32_t ghost) {
  return field + block * orc_block_len(g, ghost);
}
inline const double *slab(const orc_grid *g, const double *field, int64_t block, uint32_t ghost) {
  return field + block * orc_block_len(g, ghost);
}

// face(dx, dy) (src/discretization/finite_volume.rs:129-142).
struct KR {
  double a_bar, eps_prime, a12, tol;
};

inline KR kr_new(double eps4, double tol) {
  // KarmaRappelFlux::new (src/discretization/finite_volume.rs:113-122).
  KR k;
  k.a_bar = std::fma(3.0, -eps4, 1.0);
  k.eps_prime = 3.0 * eps4 / k.a_bar;
  k.a12 = 4.0 * k.a_bar * k.eps_prime;
  k.tol = tol;
  return k;
}

inline void kr_face(const KR &k, double dx, double dy, double &a, double &da) {
  double dx2 = dx * dx;
  double dy2 = dy * dy;
  double sum = dx2 + dy2;
  double mag2 = sum * sum;
  if (mag2 <= k.tol) {
    a = k.a_bar;
    da = 0.0;
  } else {
    a = k.a_bar * (1.0 + k.eps_prime * (dx2 * dx2 + dy2 * dy2) / mag2);
    da = -(k.a12) * dx * dy * (dx2 - dy2) / mag2;
  }
}

// KarmaRappelFlux::kernel on one block (finite_volume.rs:248-284).
void kr_block(const orc_grid *g, const KR &k, const double *in, double *out) {
  double hx = g->spacing[0], hy = g->spacing[1];
  double inv_h2 = 1.0 / (hx * hy);
  double quarter = 0.25;
  int64_t sx = g->block_cells[0] + 2;
  auto p = [&](int64_t i, int64_t j) { return in[(j + 1) * sx + (i + 1)]; };
  for_each_interior(g, [&](const int64_t *idx) {
    int64_t i = idx[0], j = idx[1];
    double derx_r = p(i + 1, j) - p(i, j);
    double derx_l = p(i, j) - p(i - 1, j);
    double derx_t = quarter * (p(i + 1, j + 1) - p(i - 1, j + 1) + p(i + 1, j) - p(i - 1, j));
    double derx_b = quarter * (p(i + 1, j) - p(i - 1, j) + p(i + 1, j - 1) - p(i - 1, j - 1));
    double dery_t = p(i, j + 1) - p(i, j);
    double dery_b = p(i, j) - p(i, j - 1);
    double dery_r = quarter * (p(i + 1, j + 1) - p(i + 1, j - 1) + p(i, j + 1) - p(i, j - 1));
    double dery_l = quarter * (p(i, j + 1) - p(i, j - 1) + p(i - 1, j + 1) - p(i - 1, j - 1));
    double a_r, da_r, a_l, da_l, a_t, da_t, a_b, da_b;
    kr_face(k, derx_r, dery_r, a_r, da_r);
    kr_face(k, derx_l, dery_l, a_l, da_l);
    kr_face(k, derx_t, dery_t, a_t, da_t);
    kr_face(k, derx_b, dery_b, a_b, da_b);
    double j_r = a_r * (a_r * derx_r - da_r * dery_r);
    double j_l = a_l * (a_l * derx_l - da_l * dery_l);
    double j_t = a_t * (a_t * dery_t + da_t * derx_t);
    double j_b = a_b * (a_b * dery_b + da_b * derx_b);
    out[(j + 1) * sx + (i + 1)] = (j_r - j_l + j_t - j_b) * inv_h2;
  });
}

// central_laplacian_kernel on one block (finite_difference.rs:49-67).
void laplacian_block(const orc_grid *g, const double *in, double *out, uint32_t ghost) {
  int D = dims(g);
  double inv_h2[3];
  for (int d = 0; d < D; ++d) inv_h2[d] = 1.0 / (g->spacing[d] * g->spacing[d]);
  for_each_interior(g, [&](const int64_t *idx) {
    int64_t c = lin(g, idx, ghost);
    double center = in[c];
    double acc = 0.0;
    for (int d = 0; d < D; ++d) {
      int64_t plus[3] = {idx[0], idx[1], idx[2]};
      int64_t minus[3] = {idx[0], idx[1], idx[2]};
      plus[d] += 1;
      minus[d] -= 1;
      acc = acc + (in[lin(g, plus, ghost)] - (center + center) + in[lin(g, minus, ghost)]) * inv_h2[d];
    }
    out[c] = acc;
  });
}

// ModelC::center_anisotropy (src/physics/phasefield.rs:159-175).
inline double center_anisotropy(const orc_modelc_params *p, double a_bar, double eps_prime,
                                const double *phi, int64_t sx, int64_t i, int64_t j) {
  auto at = [&](int64_t ii, int64_t jj) { return phi[(jj + 1) * sx + (ii + 1)]; };
  double dx = 0.5 * (at(i + 1, j) - at(i - 1, j));
  double dy = 0.5 * (at(i, j + 1) - at(i, j - 1));
  double dx2 = dx * dx;
  double dy2 = dy * dy;
  double sum = dx2 + dy2;
  double mag2 = sum * sum;
  if (mag2 <= p->anisotropy_tol) return a_bar;
  return a_bar * (1.0 + eps_prime * (dx2 * dx2 + dy2 * dy2) / mag2);
}

// ModelC::vector_field_block, Driver::Time, axis-aligned path
// (src/physics/phasefield.rs:367-421) on one block.
void modelc_rhs_block(const orc_grid *g, const orc_modelc_params *p, const double *phi,
                      const double *u, double *kphi, double *ku) {
  // ModelC::new constants (phasefield.rs:100-116).
  double a_bar = std::fma(3.0, -p->eps4, 1.0);
  double eps_prime = 3.0 * p->eps4 / a_bar;
  KR k = kr_new(p->eps4, p->fv_tol);
  int64_t sx = g->block_cells[0] + 2;
  kr_block(g, k, phi, kphi);
  for_each_interior(g, [&](const int64_t *idx) {
    int64_t o = (idx[1] + 1) * sx + (idx[0] + 1);
    double pv = phi[o];
    double g_prime = std::fma(pv * pv, pv, -pv);
    double one_m_p2 = std::fma(pv, -pv, 1.0);
    double p_prime = one_m_p2 * one_m_p2;
    double a = center_anisotropy(p, a_bar, eps_prime, phi, sx, idx[0], idx[1]);
    double val = std::fma(p->lambda * u[o], -p_prime, kphi[o] - g_prime) / (a * a);
    kphi[o] = val;
  });
  laplacian_block(g, u, ku, 1);
  for_each_interior(g, [&](const int64_t *idx) {
    int64_t o = (idx[1] + 1) * sx + (idx[0] + 1);
    ku[o] = std::fma(ku[o], p->d_thermal, 0.5 * kphi[o]);
  });
}

// Driver::Wiener arm of Driver::apply_slab (src/core/driver.rs:139-158) on one slab.
void wiener_block(const orc_grid *g, int64_t block, double *x, const double *amp, uint32_t ghost,
                  double dt, uint64_t seed, uint64_t salt, uint64_t stream) {
  double scale = std::sqrt(dt);
  uint64_t parts[3] = {salt, stream, (uint64_t)block};
  uint64_t block_key = orc_mix_key(seed, parts, 3);
  int64_t len = orc_block_len(g, ghost);
  for (int64_t i = 0; i < len; ++i) {
    double v = amp[i];
    if (v == 0.0) continue;
    int64_t cell = orc_cell_key(g, ghost, i);
    if (cell < 0) continue;
    uint64_t key = orc_splitmix64(block_key ^ (uint64_t)cell);
    x[i] = x[i] + scale * v * orc_standard_normal(key);
  }
}

}  // namespace

extern "C" {

// ---------------------------------------------------------------- rng.rs:13-49
uint64_t orc_splitmix64(uint64_t z) {
  z += 0x9E3779B97F4A7C15ull;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

uint64_t orc_mix_key(uint64_t seed, const uint64_t *parts, size_t n) {
  uint64_t s = orc_splitmix64(seed);
  for (size_t i = 0; i < n; ++i) s = orc_splitmix64(s ^ parts[i]);
  return s;
}

double orc_unit_open(uint64_t bits) {
  return (double)((bits >> 11) + 1) * (1.0 / 9007199254740992.0);
}

double orc_standard_normal(uint64_t key) {
  uint64_t a = orc_splitmix64(key);
  uint64_t b = orc_splitmix64(a ^ 0xD1B54A32D192ED03ull);
  double u = orc_unit_open(a);
  double v = orc_unit_open(b);
  return std::sqrt(-2.0 * std::log(u)) * std::cos(kTau * v);
}

// ------------------------------------------------------ cartesian.rs topology
int orc_grid_check(const orc_grid *g) {
  // CartesianGrid::new validation (cartesian.rs:48-82).
  if (g->dim < 1 || g->dim > 3) return -1;
  for (int d = 0; d < g->dim; ++d) {
    if (!(g->cells[d] > 0 && g->block_cells[d] > 0)) return -1;
    if (g->cells[d] % g->block_cells[d] != 0) return -1;
    if (!(g->spacing[d] > 0.0) || !std::isfinite(g->spacing[d])) return -1;
  }
  return 0;
}

int64_t orc_num_blocks(const orc_grid *g) {
  int64_t n = 1;
  for (int d = 0; d < g->dim; ++d) n *= blocks_per_dim(g, d);
  return n;
}

int64_t orc_block_len(const orc_grid *g, uint32_t ghost) {
  int64_t ext[3];
  slab_ext(g, ghost, ext);
  return ext[0] * ext[1] * ext[2];
}

// face_neighbor (cartesian.rs:162-181); -1 stands for None.
int64_t orc_face_neighbor(const orc_grid *g, int64_t block, int32_t dim, int32_t dir) {
  int64_t c[3];
  block_coords(g, block, c);
  int64_t stepped = c[dim] + (dir > 0 ? 1 : (dir < 0 ? -1 : 0));
  int64_t nb = blocks_per_dim(g, dim);
  int64_t next;
  if (stepped >= 0 && stepped < nb) {
    next = stepped;
  } else if (g->periodic[dim]) {
    next = ((stepped % nb) + nb) % nb;
  } else {
    return -1;
  }
  int64_t id = 0, stride = 1;
  for (int d = 0; d < g->dim; ++d) {
    id += (d == dim ? next : c[d]) * stride;
    stride *= blocks_per_dim(g, d);
  }
  return id;
}

// cell_key (cartesian.rs:599-619); -1 stands for None.
int64_t orc_cell_key(const orc_grid *g, uint32_t ghost, int64_t offset) {
  int64_t gg = ghost;
  int64_t rem = offset, key = 0, stride = 1;
  for (int d = 0; d < g->dim; ++d) {
    int64_t ghosted = g->block_cells[d] + 2 * gg;
    int64_t c = rem % ghosted;
    rem /= ghosted;
    if (c < gg || c >= g->block_cells[d] + gg) return -1;
    key += (c - gg) * stride;
    stride *= g->block_cells[d];
  }
  return key;
}

// cell_center (cartesian.rs:589-597).
void orc_cell_center(const orc_grid *g, int64_t block, const int64_t *idx, double *out) {
  int64_t c[3];
  block_coords(g, block, c);
  for (int d = 0; d < g->dim; ++d) {
    int64_t global = c[d] * g->block_cells[d];
    out[d] = std::fma((double)global + (double)idx[d] + 0.5, g->spacing[d], g->origin[d]);
  }
}

// fill_ghosts_generic (cartesian.rs:439-526).
int orc_fill_ghosts(const orc_grid *g, double *field, uint32_t ghost, const orc_face_bc *bcs) {
  if (ghost == 0) return 0;
  int D = g->dim;
  int64_t gw = ghost;
  for (int d = 0; d < D; ++d)
    if (gw > g->block_cells[d]) return -1;
  int64_t nblocks = orc_num_blocks(g);
  for (int d = 0; d < D; ++d) {
    int64_t n = g->block_cells[d];
    double h = g->spacing[d];
    Box3 box;
    for (int e = 0; e < 3; ++e) {
      box.lo[e] = e < D ? -gw : 0;
      box.hi[e] = e < D ? g->block_cells[e] + gw : 1;
    }
    for (int64_t b = 0; b < nblocks; ++b) {
      for (int dir = -1; dir <= 1; dir += 2) {
        if (dir < 0) {
          box.lo[d] = -gw;
          box.hi[d] = 0;
        } else {
          box.lo[d] = n;
          box.hi[d] = n + gw;
        }
        int64_t nb = orc_face_neighbor(g, b, d, dir);
        double *dst = slab(g, field, b, ghost);
        if (nb >= 0) {
          // Exchange (or periodic self-wrap when nb == b): copy the neighbor's
          // d-interior into this block's d-ghost strip.
          const double *src = slab(g, (const double *)field, nb, ghost);
          for_each_box(g, box, [&](const int64_t *idx) {
            int64_t sidx[3] = {idx[0], idx[1], idx[2]};
            sidx[d] = dir < 0 ? n + idx[d] : idx[d] - n;
            dst[lin(g, idx, ghost)] = src[lin(g, sidx, ghost)];
          });
        } else {
          orc_face_bc bc = {ORC_BC_MIRROR, 0.0};
          if (bcs) bc = bcs[d * 2 + (dir > 0 ? 1 : 0)];
          for_each_box(g, box, [&](const int64_t *idx) {
            int64_t layer = dir < 0 ? -1 - idx[d] : idx[d] - n;
            double val = ghost_value(bc, layer, h, [&](int64_t inward) {
              int64_t iidx[3] = {idx[0], idx[1], idx[2]};
              iidx[d] = dir < 0 ? inward : n - 1 - inward;
              return dst[lin(g, iidx, ghost)];
            });
            dst[lin(g, idx, ghost)] = val;
          });
        }
      }
    }
  }
  return 0;
}

void orc_central_laplacian(const orc_grid *g, const double *in, double *out, uint32_t ghost) {
  int64_t nblocks = orc_num_blocks(g);
  for (int64_t b = 0; b < nblocks; ++b)
    laplacian_block(g, slab(g, in, b, ghost), slab(g, out, b, ghost), ghost);
}

void orc_karma_rappel_flux(const orc_grid *g, const double *phi, double *out, double eps4,
                           double tol) {
  KR k = kr_new(eps4, tol);
  int64_t nblocks = orc_num_blocks(g);
  for (int64_t b = 0; b < nblocks; ++b) kr_block(g, k, slab(g, phi, b, 1), slab(g, out, b, 1));
}

void orc_modelc_rhs(const orc_grid *g, const orc_modelc_params *p, const double *phi,
                    const double *u, double *kphi, double *ku) {
  int64_t nblocks = orc_num_blocks(g);
  for (int64_t b = 0; b < nblocks; ++b)
    modelc_rhs_block(g, p, slab(g, phi, b, 1), slab(g, u, b, 1), slab(g, kphi, b, 1),
                     slab(g, ku, b, 1));
}

void orc_modelc_amplitude(const orc_grid *g, const orc_modelc_params *p, double *amp_phi) {
  // eval_tendency_impl: fill_zero then the model writes the interior
  // (src/integrators/mod.rs:263; phasefield.rs:361-366).
  int64_t nblocks = orc_num_blocks(g);
  int64_t len = orc_block_len(g, 1);
  for (int64_t b = 0; b < nblocks; ++b) {
    double *a = slab(g, amp_phi, b, 1);
    for (int64_t i = 0; i < len; ++i) a[i] = 0.0;
    for_each_interior(g, [&](const int64_t *idx) { a[lin(g, idx, 1)] = p->noise_amplitude; });
  }
}

void orc_heat_rhs(const orc_grid *g, double kappa, const double *u, double *k) {
  int64_t nblocks = orc_num_blocks(g);
  for (int64_t b = 0; b < nblocks; ++b) {
    double *kb = slab(g, k, b, 1);
    laplacian_block(g, slab(g, u, b, 1), kb, 1);
    for_each_interior(g, [&](const int64_t *idx) {
      int64_t o = lin(g, idx, 1);
      kb[o] = kb[o] * kappa;
    });
  }
}

void orc_axpy(double *x, const double *y, double alpha, size_t n) {
  for (size_t i = 0; i < n; ++i) x[i] = x[i] + alpha * y[i];
}

void orc_apply_time(double *x, const double *k, double dt, size_t n) { orc_axpy(x, k, dt, n); }

void orc_apply_wiener(const orc_grid *g, double *x, const double *amp, uint32_t ghost, double dt,
                      uint64_t seed, uint64_t salt, uint64_t stream) {
  int64_t nblocks = orc_num_blocks(g);
  for (int64_t b = 0; b < nblocks; ++b)
    wiener_block(g, b, slab(g, x, b, ghost), slab(g, amp, b, ghost), ghost, dt, seed, salt, stream);
}

int orc_modelc_run(const orc_grid *g, const orc_modelc_params *p, double *phi, double *u,
                   int64_t nsteps, double dt, double *t, uint64_t seed, int32_t wiener,
                   int32_t nthreads) {
  if (g->dim != 2 || orc_grid_check(g) != 0) return -1;
  int64_t nblocks = orc_num_blocks(g);
  int64_t len = orc_block_len(g, 1);
  size_t total = (size_t)(nblocks * len);
  // Stage layout of EulerMaruyama: k = like_for(Time) (phi, u) and, for
  // Wiener<1>, amp = like_for(Wiener(0)) (phi only) — zero-initialized.
  std::vector<double> kphi(total, 0.0), ku(total, 0.0), amp;
  if (wiener) amp.assign(total, 0.0);
  const uint64_t stream = (1ull << 32) | 0ull;  // Driver::Wiener(0).stream()
  (void)nthreads;
  for (int64_t s = 0; s < nsteps; ++s) {
    double tn = *t;
    uint64_t salt;
    std::memcpy(&salt, &tn, sizeof salt);  // t.to_bits()
    // eval_drift: ModelC::fill_ghosts (serial) then per-block RHS.
    if (orc_fill_ghosts(g, phi, 1, nullptr) != 0) return -1;
    if (orc_fill_ghosts(g, u, 1, nullptr) != 0) return -1;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
    for (int64_t b = 0; b < nblocks; ++b)
      modelc_rhs_block(g, p, slab(g, (const double *)phi, b, 1), slab(g, (const double *)u, b, 1),
                       slab(g, kphi.data(), b, 1), slab(g, ku.data(), b, 1));
    if (wiener) {
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
      for (int64_t b = 0; b < nblocks; ++b) {
        double *a = slab(g, amp.data(), b, 1);
        for (int64_t i = 0; i < len; ++i) a[i] = 0.0;
        for_each_interior(g, [&](const int64_t *idx) { a[lin(g, idx, 1)] = p->noise_amplitude; });
      }
    }
    // apply_step_with: Time on every driven field, then Wiener(0) on phi.
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
    for (int64_t b = 0; b < nblocks; ++b) {
      double *pb = slab(g, phi, b, 1), *ub = slab(g, u, b, 1);
      orc_axpy(pb, slab(g, (const double *)kphi.data(), b, 1), dt, (size_t)len);
      orc_axpy(ub, slab(g, (const double *)ku.data(), b, 1), dt, (size_t)len);
      if (wiener) wiener_block(g, b, pb, slab(g, (const double *)amp.data(), b, 1), 1, dt, seed, salt, stream);
    }
    *t = *t + dt;
  }
  return 0;
}

// The branch face() (finite_volume.rs:129-142) takes at each of a cell's four faces
// (r, l, t, b: bits 0..3) and the centre anisotropy (phasefield.rs:159-175, bit 4)
// for a ghost-filled phi, 1 = the flat branch (|∇φ|⁴ <= tol); one byte per interior
// cell, block-major.
void modelc_branch_mask(const orc_grid *g, const orc_modelc_params *p, const double *phi,
                        std::vector<uint8_t> &mask, int32_t nthreads) {
  const int64_t nblocks = orc_num_blocks(g), nx = g->block_cells[0], ny = g->block_cells[1];
  const int64_t sx = nx + 2, len = orc_block_len(g, 1);
  mask.assign((size_t)(nblocks * nx * ny), 0);
  auto flat = [](double dx, double dy, double tol) {
    double dx2 = dx * dx, dy2 = dy * dy, sum = dx2 + dy2;
    return sum * sum <= tol;
  };
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
  for (int64_t b = 0; b < nblocks; ++b) {
    const double *in = phi + b * len;
    auto q = [&](int64_t i, int64_t j) { return in[(j + 1) * sx + (i + 1)]; };
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        const double qq = 0.25;
        double derx_r = q(i + 1, j) - q(i, j), derx_l = q(i, j) - q(i - 1, j);
        double derx_t = qq * (q(i + 1, j + 1) - q(i - 1, j + 1) + q(i + 1, j) - q(i - 1, j));
        double derx_b = qq * (q(i + 1, j) - q(i - 1, j) + q(i + 1, j - 1) - q(i - 1, j - 1));
        double dery_t = q(i, j + 1) - q(i, j), dery_b = q(i, j) - q(i, j - 1);
        double dery_r = qq * (q(i + 1, j + 1) - q(i + 1, j - 1) + q(i, j + 1) - q(i, j - 1));
        double dery_l = qq * (q(i, j + 1) - q(i, j - 1) + q(i - 1, j + 1) - q(i - 1, j - 1));
        double cx = 0.5 * (q(i + 1, j) - q(i - 1, j)), cy = 0.5 * (q(i, j + 1) - q(i, j - 1));
        uint8_t m = (uint8_t)(flat(derx_r, dery_r, p->fv_tol) | flat(derx_l, dery_l, p->fv_tol) << 1 |
                              flat(derx_t, dery_t, p->fv_tol) << 2 | flat(derx_b, dery_b, p->fv_tol) << 3 |
                              flat(cx, cy, p->anisotropy_tol) << 4);
        mask[(size_t)((b * ny + j) * nx + i)] = m;
      }
  }
}

int orc_modelc_branch_crossings(const orc_grid *g, const orc_modelc_params *pm,
                                const orc_modelc_params *pp, const double *phi0, const double *u0,
                                int64_t nsteps, double dt, int32_t nthreads, int64_t *crossings) {
  if (g->dim != 2 || orc_grid_check(g) != 0) return -1;
  const size_t total = (size_t)(orc_num_blocks(g) * orc_block_len(g, 1));
  std::vector<double> phi[2] = {std::vector<double>(phi0, phi0 + total),
                                std::vector<double>(phi0, phi0 + total)};
  std::vector<double> u[2] = {std::vector<double>(u0, u0 + total), std::vector<double>(u0, u0 + total)};
  const orc_modelc_params *ps[2] = {pm, pp};
  std::vector<uint8_t> mask[2];
  int64_t count = 0;
  for (int64_t s = 0; s < nsteps; ++s) {
    for (int r = 0; r < 2; ++r) {
      if (orc_fill_ghosts(g, phi[r].data(), 1, nullptr) != 0) return -1;
      modelc_branch_mask(g, ps[r], phi[r].data(), mask[r], nthreads);
    }
    for (size_t c = 0; c < mask[0].size(); ++c)
      count += __builtin_popcount((unsigned)(mask[0][c] ^ mask[1][c]));
    for (int r = 0; r < 2; ++r) {
      double t = 0.0;
      if (orc_modelc_run(g, ps[r], phi[r].data(), u[r].data(), 1, dt, &t, 0, 0, nthreads) != 0)
        return -1;
    }
  }
  *crossings = count;
  return 0;
}

int orc_heat_run(const orc_grid *g, double kappa, double *u, int64_t nsteps, double dt, double *t,
                 const orc_face_bc *bcs, int32_t nthreads) {
  if (orc_grid_check(g) != 0) return -1;
  int64_t nblocks = orc_num_blocks(g);
  int64_t len = orc_block_len(g, 1);
  std::vector<double> k((size_t)(nblocks * len), 0.0);
  for (int64_t s = 0; s < nsteps; ++s) {
    if (orc_fill_ghosts(g, u, 1, bcs) != 0) return -1;
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
    for (int64_t b = 0; b < nblocks; ++b) {
      double *kb = slab(g, k.data(), b, 1);
      laplacian_block(g, slab(g, (const double *)u, b, 1), kb, 1);
      for_each_interior(g, [&](const int64_t *idx) {
        int64_t o = lin(g, idx, 1);
        kb[o] = kb[o] * kappa;
      });
    }
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
    for (int64_t b = 0; b < nblocks; ++b)
      orc_axpy(slab(g, u, b, 1), slab(g, (const double *)k.data(), b, 1), dt, (size_t)len);
    *t = *t + dt;
  }
  return 0;
}

void orc_initialize_grains(const orc_grid *g, double *phi, double *u, const double *grains,
                           size_t ngrains, double u0) {
  const double cutoff = 27.0;
  int64_t nx = g->block_cells[0], ny = g->block_cells[1];
  int64_t nblocks = orc_num_blocks(g);
  const double sqrt2 = std::sqrt(2.0);
  for (int64_t b = 0; b < nblocks; ++b) {
    double hx = g->spacing[0], hy = g->spacing[1];
    int64_t i00[3] = {0, 0, 0}, i11[3] = {nx - 1, ny - 1, 0};
    double c0[3], c1[3];
    orc_cell_center(g, b, i00, c0);
    orc_cell_center(g, b, i11, c1);
    double lo[2] = {std::fma(0.5, -hx, c0[0]), std::fma(0.5, -hy, c0[1])};
    double hi[2] = {std::fma(0.5, hx, c1[0]), std::fma(0.5, hy, c1[1])};
    auto dist_box = [&](double cx, double cy) {
      double dx = std::fmax(std::fmax(lo[0] - cx, cx - hi[0]), 0.0);
      double dy = std::fmax(std::fmax(lo[1] - cy, cy - hi[1]), 0.0);
      return std::hypot(dx, dy);
    };
    std::vector<size_t> near;
    for (size_t k = 0; k < ngrains; ++k)
      if (dist_box(grains[3 * k], grains[3 * k + 1]) <= grains[3 * k + 2] + cutoff) near.push_back(k);
    double *pb = slab(g, phi, b, 1), *ub = slab(g, u, b, 1);
    for_each_interior(g, [&](const int64_t *idx) {
      double val = -1.0;
      if (!near.empty()) {
        double xy[3];
        orc_cell_center(g, b, idx, xy);
        for (size_t k : near) {
          double r = std::hypot(xy[0] - grains[3 * k], xy[1] - grains[3 * k + 1]);
          double v = -std::tanh((r - grains[3 * k + 2]) / sqrt2);
          val = std::fmax(val, v);
        }
      }
      pb[lin(g, idx, 1)] = val;
    });
    for_each_interior(g, [&](const int64_t *idx) { ub[lin(g, idx, 1)] = -u0; });
  }
}

void orc_seed_positions(uint64_t seed, size_t nseeds, double extent, double radius,
                        double *grains_out) {
  double margin = 2.0 * radius;
  double span = std::fma(2.0, -margin, extent);
  for (size_t i = 0; i < nseeds; ++i) {
    uint64_t px[2] = {(uint64_t)i, 0}, py[2] = {(uint64_t)i, 1};
    grains_out[3 * i] = margin + orc_unit_open(orc_mix_key(seed, px, 2)) * span;
    grains_out[3 * i + 1] = margin + orc_unit_open(orc_mix_key(seed, py, 2)) * span;
    grains_out[3 * i + 2] = radius;
  }
}

}  // extern "C"

// ===========================================================================
// Multi-grain orientations and the classic fourth-order Runge–Kutta.
// ===========================================================================
namespace {

// rotate (src/discretization/finite_volume.rs:241-243): the crystal-frame
// components (c0·dx + s0·dy, c0·dy − s0·dx).
inline void rotate(double c0, double s0, double dx, double dy, double &rx, double &ry) {
  rx = c0 * dx + s0 * dy;
  ry = c0 * dy - s0 * dx;
}

// KarmaRappelFlux::apply_oriented (finite_volume.rs:188-235) and the multi-grain
// branch of ModelC::vector_field_block (phasefield.rs:378-392) on one block:
// θ₀ at the cell (ghost-0 slab) rotates every face's gradient into the crystal
// frame for (A, A′) while the fluxes keep the lab-frame components; the centre
// anisotropy is the flux's face() of the rotated central difference
// (center_anisotropy_rotated, finite_volume.rs:163-177).
void modelc_rhs_oriented_block(const orc_grid *g, const orc_modelc_params *p, const double *phi,
                               const double *u, const double *theta, double *kphi, double *ku) {
  KR k = kr_new(p->eps4, p->fv_tol);
  const double inv_h2 = 1.0 / (g->spacing[0] * g->spacing[1]);
  const double quarter = 0.25;
  const int64_t nx = g->block_cells[0], sx = nx + 2;
  auto pv = [&](int64_t i, int64_t j) { return phi[(j + 1) * sx + (i + 1)]; };
  for_each_interior(g, [&](const int64_t *idx) {
    const int64_t i = idx[0], j = idx[1];
    const double th = theta[j * nx + i];
    const double s0 = std::sin(th), c0 = std::cos(th);
    double derx_r = pv(i + 1, j) - pv(i, j);
    double derx_l = pv(i, j) - pv(i - 1, j);
    double derx_t = quarter * (pv(i + 1, j + 1) - pv(i - 1, j + 1) + pv(i + 1, j) - pv(i - 1, j));
    double derx_b = quarter * (pv(i + 1, j) - pv(i - 1, j) + pv(i + 1, j - 1) - pv(i - 1, j - 1));
    double dery_t = pv(i, j + 1) - pv(i, j);
    double dery_b = pv(i, j) - pv(i, j - 1);
    double dery_r = quarter * (pv(i + 1, j + 1) - pv(i + 1, j - 1) + pv(i, j + 1) - pv(i, j - 1));
    double dery_l = quarter * (pv(i, j + 1) - pv(i, j - 1) + pv(i - 1, j + 1) - pv(i - 1, j - 1));
    double rx, ry, a_r, da_r, a_l, da_l, a_t, da_t, a_b, da_b;
    rotate(c0, s0, derx_r, dery_r, rx, ry);
    kr_face(k, rx, ry, a_r, da_r);
    rotate(c0, s0, derx_l, dery_l, rx, ry);
    kr_face(k, rx, ry, a_l, da_l);
    rotate(c0, s0, derx_t, dery_t, rx, ry);
    kr_face(k, rx, ry, a_t, da_t);
    rotate(c0, s0, derx_b, dery_b, rx, ry);
    kr_face(k, rx, ry, a_b, da_b);
    double j_r = a_r * (a_r * derx_r - da_r * dery_r);
    double j_l = a_l * (a_l * derx_l - da_l * dery_l);
    double j_t = a_t * (a_t * dery_t + da_t * derx_t);
    double j_b = a_b * (a_b * dery_b + da_b * derx_b);
    const int64_t o = (j + 1) * sx + (i + 1);
    const double flux = (j_r - j_l + j_t - j_b) * inv_h2;
    const double pc = phi[o];
    const double g_prime = std::fma(pc * pc, pc, -pc);
    const double one_m_p2 = std::fma(pc, -pc, 1.0);
    const double p_prime = one_m_p2 * one_m_p2;
    const double cdx = 0.5 * (pv(i + 1, j) - pv(i - 1, j));
    const double cdy = 0.5 * (pv(i, j + 1) - pv(i, j - 1));
    double crx, cry, a, da;
    rotate(c0, s0, cdx, cdy, crx, cry);
    kr_face(k, crx, cry, a, da);
    kphi[o] = std::fma(p->lambda * u[o], -p_prime, flux - g_prime) / (a * a);
  });
  laplacian_block(g, u, ku, 1);
  for_each_interior(g, [&](const int64_t *idx) {
    const int64_t o = (idx[1] + 1) * sx + (idx[0] + 1);
    ku[o] = std::fma(ku[o], p->d_thermal, 0.5 * kphi[o]);
  });
}

// eval_drift of Model C (fill_ghosts_mirror on φ and u, then the per-block vector
// field, axis-aligned or oriented) into (kphi, ku).
void modelc_eval(const orc_grid *g, const orc_modelc_params *p, double *phi, double *u,
                 const double *theta, double *kphi, double *ku) {
  orc_fill_ghosts(g, phi, 1, nullptr);
  orc_fill_ghosts(g, u, 1, nullptr);
  const int64_t nb = orc_num_blocks(g), len = orc_block_len(g, 1), tl = orc_block_len(g, 0);
  for (int64_t b = 0; b < nb; ++b) {
    if (theta)
      modelc_rhs_oriented_block(g, p, phi + b * len, u + b * len, theta + b * tl, kphi + b * len,
                                ku + b * len);
    else
      modelc_rhs_block(g, p, phi + b * len, u + b * len, kphi + b * len, ku + b * len);
  }
}

// eval_drift of Heat<D> (fill_ghosts with bcs, κ·∇²u per block) into k.
void heat_eval(const orc_grid *g, double kappa, double *u, const orc_face_bc *bcs, double *k) {
  orc_fill_ghosts(g, u, 1, bcs);
  const int64_t nb = orc_num_blocks(g), len = orc_block_len(g, 1);
  for (int64_t b = 0; b < nb; ++b) {
    double *kb = k + b * len;
    laplacian_block(g, u + b * len, kb, 1);
    for_each_interior(g, [&](const int64_t *idx) {
      const int64_t o = lin(g, idx, 1);
      kb[o] = kb[o] * kappa;
    });
  }
}

// RungeKutta4::step (src/integrators/rk4.rs:30-80) over nf fields of `len` entries
// (whole slabs: copy_from_with and axpy_with sweep ghost entries too; the
// tendencies' ghost entries stay zero).
template <class Eval>
void rk4_step(int nf, size_t len, double *const *state, std::vector<double> (&k)[4][2],
              std::vector<double> (&y)[2], double dt, Eval &&eval) {
  const double half_dt = 0.5 * dt;
  const double c[3] = {half_dt, half_dt, dt};
  auto ptrs = [&](std::vector<double> *v, double **out) {
    for (int f = 0; f < nf; ++f) out[f] = v[f].data();
  };
  double *kp[2], *yp[2];
  ptrs(k[0], kp);
  eval(state, kp);
  for (int s = 1; s < 4; ++s) {
    ptrs(y, yp);
    for (int f = 0; f < nf; ++f) {
      std::copy(state[f], state[f] + len, y[f].begin());
      orc_axpy(y[f].data(), k[s - 1][f].data(), c[s - 1], len);
    }
    ptrs(k[s], kp);
    eval(yp, kp);
  }
  const double sixth = dt / 6.0, third = dt / 3.0;
  const double w[4] = {sixth, third, third, sixth};
  for (int s = 0; s < 4; ++s)
    for (int f = 0; f < nf; ++f) orc_axpy(state[f], k[s][f].data(), w[s], len);
}

}  // namespace

extern "C" {

void orc_modelc_rhs_oriented(const orc_grid *g, const orc_modelc_params *p, const double *phi,
                             const double *u, const double *theta, double *kphi, double *ku) {
  const int64_t nb = orc_num_blocks(g), len = orc_block_len(g, 1), tl = orc_block_len(g, 0);
  for (int64_t b = 0; b < nb; ++b)
    modelc_rhs_oriented_block(g, p, phi + b * len, u + b * len, theta + b * tl, kphi + b * len,
                              ku + b * len);
}

int orc_modelc_rk4_run(const orc_grid *g, const orc_modelc_params *p, double *phi, double *u,
                       const double *theta, int64_t nsteps, double dt, double *t) {
  if (g->dim != 2 || orc_grid_check(g) != 0) return -1;
  const size_t len = (size_t)(orc_num_blocks(g) * orc_block_len(g, 1));
  std::vector<double> k[4][2], y[2];
  for (auto &ks : k)
    for (auto &v : ks) v.assign(len, 0.0);
  for (auto &v : y) v.assign(len, 0.0);
  double *state[2] = {phi, u};
  for (int64_t s = 0; s < nsteps; ++s) {
    rk4_step(2, len, state, k, y, dt, [&](double *const *in, double *const *out) {
      modelc_eval(g, p, in[0], in[1], theta, out[0], out[1]);
    });
    *t = *t + dt;
  }
  return 0;
}

int orc_heat_rk4_run(const orc_grid *g, double kappa, double *u, int64_t nsteps, double dt,
                     double *t, const orc_face_bc *bcs) {
  if (orc_grid_check(g) != 0) return -1;
  const size_t len = (size_t)(orc_num_blocks(g) * orc_block_len(g, 1));
  std::vector<double> k[4][2], y[2];
  for (auto &ks : k) ks[0].assign(len, 0.0);
  y[0].assign(len, 0.0);
  double *state[2] = {u, nullptr};
  for (int64_t s = 0; s < nsteps; ++s) {
    rk4_step(1, len, state, k, y, dt, [&](double *const *in, double *const *out) {
      heat_eval(g, kappa, in[0], bcs, out[0]);
    });
    *t = *t + dt;
  }
  return 0;
}

// Oriented Euler–Maruyama / forward-Euler run (b = 0): as orc_modelc_run with the
// multi-grain vector field.
int orc_modelc_oriented_run(const orc_grid *g, const orc_modelc_params *p, double *phi, double *u,
                            const double *theta, int64_t nsteps, double dt, double *t) {
  if (g->dim != 2 || orc_grid_check(g) != 0) return -1;
  const size_t len = (size_t)(orc_num_blocks(g) * orc_block_len(g, 1));
  std::vector<double> kphi(len, 0.0), ku(len, 0.0);
  for (int64_t s = 0; s < nsteps; ++s) {
    modelc_eval(g, p, phi, u, theta, kphi.data(), ku.data());
    orc_axpy(phi, kphi.data(), dt, len);
    orc_axpy(u, ku.data(), dt, len);
    *t = *t + dt;
  }
  return 0;
}

// The orientation pass of ModelC::initialize_grains (phasefield.rs:286-314):
// nearest-seed Voronoi over the candidates whose closest approach to the block
// beats every grain's farthest corner; first minimum on ties; 0 with no grain.
// grains: [cx, cy, radius, θ₀] × n; theta: ghost-0 slabs.
void orc_initialize_orientations(const orc_grid *g, double *theta, const double *grains,
                                 size_t ngrains) {
  const int64_t nx = g->block_cells[0], ny = g->block_cells[1];
  const int64_t nb = orc_num_blocks(g);
  for (int64_t b = 0; b < nb; ++b) {
    const double hx = g->spacing[0], hy = g->spacing[1];
    int64_t i00[3] = {0, 0, 0}, i11[3] = {nx - 1, ny - 1, 0};
    double c0[3], c1[3];
    orc_cell_center(g, b, i00, c0);
    orc_cell_center(g, b, i11, c1);
    const double lo[2] = {std::fma(0.5, -hx, c0[0]), std::fma(0.5, -hy, c0[1])};
    const double hi[2] = {std::fma(0.5, hx, c1[0]), std::fma(0.5, hy, c1[1])};
    auto dist_box = [&](double cx, double cy) {
      double dx = std::fmax(std::fmax(lo[0] - cx, cx - hi[0]), 0.0);
      double dy = std::fmax(std::fmax(lo[1] - cy, cy - hi[1]), 0.0);
      return std::hypot(dx, dy);
    };
    auto dist_far = [&](double cx, double cy) {
      double dx = std::fmax(std::fabs(cx - lo[0]), std::fabs(cx - hi[0]));
      double dy = std::fmax(std::fabs(cy - lo[1]), std::fabs(cy - hi[1]));
      return std::hypot(dx, dy);
    };
    double threshold = 1.7976931348623157e308;
    for (size_t k = 0; k < ngrains; ++k)
      threshold = std::fmin(threshold, dist_far(grains[4 * k], grains[4 * k + 1]));
    std::vector<size_t> cand;
    for (size_t k = 0; k < ngrains; ++k)
      if (dist_box(grains[4 * k], grains[4 * k + 1]) <= threshold) cand.push_back(k);
    double *tb = theta + b * nx * ny;
    for (int64_t j = 0; j < ny; ++j)
      for (int64_t i = 0; i < nx; ++i) {
        int64_t idx[3] = {i, j, 0};
        double xy[3];
        orc_cell_center(g, b, idx, xy);
        double best = 0.0, bestd = 0.0;
        bool have = false;
        for (size_t k : cand) {
          const double ddx = xy[0] - grains[4 * k], ddy = xy[1] - grains[4 * k + 1];
          const double d = std::fma(ddy, ddy, ddx * ddx);
          if (!have || d < bestd) {
            have = true;
            bestd = d;
            best = grains[4 * k + 3];
          }
        }
        tb[j * nx + i] = have ? best : 0.0;
      }
  }
}

}  // extern "C"

// ===========================================================================
// AMR (2-D): src/geometry/amr/{grid,intergrid,cluster,adapt}.rs and
// src/integrators/subcycling.rs.
// ===========================================================================
#include <algorithm>
#include <array>
#include <memory>

namespace {

struct Box2 {
  int64_t lo[2], hi[2];
  int64_t extent(int d) const { return hi[d] - lo[d]; }
  uint64_t cells() const {
    int64_t a = extent(0), b = extent(1);
    if (a <= 0 || b <= 0) return 0;
    return (uint64_t)a * (uint64_t)b;
  }
  bool contains(int64_t x, int64_t y) const {
    return lo[0] <= x && x < hi[0] && lo[1] <= y && y < hi[1];
  }
  bool intersects(const Box2 &o) const {
    return lo[0] < o.hi[0] && o.lo[0] < hi[0] && lo[1] < o.hi[1] && o.lo[1] < hi[1];
  }
  Box2 grown(int64_t b) const { return Box2{{lo[0] - b, lo[1] - b}, {hi[0] + b, hi[1] + b}}; }
  // CellBox::clipped (cluster.rs:102-109).
  Box2 clipped(const Box2 &dom) const {
    Box2 o = *this;
    for (int d = 0; d < 2; ++d) {
      o.lo[d] = std::max(o.lo[d], dom.lo[d]);
      o.hi[d] = std::max(std::min(o.hi[d], dom.hi[d]), o.lo[d]);
    }
    return o;
  }
  bool operator==(const Box2 &o) const {
    return lo[0] == o.lo[0] && lo[1] == o.lo[1] && hi[0] == o.hi[0] && hi[1] == o.hi[1];
  }
};

inline int64_t div_euclid(int64_t a, int64_t b) {
  int64_t q = a / b;
  if (a % b < 0) q -= 1;
  return q;
}

// reflect_coord (src/geometry/boundary.rs:82-90).
inline int64_t reflect(int64_t c, int64_t lo, int64_t hi) {
  if (c < lo) return 2 * lo - 1 - c;
  if (c >= hi) return 2 * hi - 1 - c;
  return c;
}

}  // namespace

struct orc_amr {
  orc_grid base;
  std::vector<uint32_t> ratios;
  std::vector<int32_t> level;        // per patch
  std::vector<Box2> box;             // per patch
  std::vector<uint32_t> level_start;  // populated levels + 1 entries
  std::vector<std::array<double, 2>> spacing;  // per level (capacity)
  std::vector<std::vector<uint32_t>> lookup;    // per populated level

  int num_levels() const { return (int)level_start.size() - 1; }
  int max_levels() const { return (int)ratios.size() + 1; }
  Box2 domain(int l) const {
    int64_t s = 1;
    for (int i = 0; i < l; ++i) s *= (int64_t)ratios[i];
    return Box2{{0, 0}, {base.cells[0] * s, base.cells[1] * s}};
  }
  int64_t nblocks() const { return (int64_t)box.size(); }
  int64_t ext(int64_t b, int d) const { return box[b].extent(d); }
  int64_t block_len(int64_t b, uint32_t g) const {
    return (ext(b, 0) + 2 * (int64_t)g) * (ext(b, 1) + 2 * (int64_t)g);
  }
  std::vector<int64_t> offs[3];  // prefix slab offsets for ghost 0, 1, 2
  void build_offsets() {
    for (uint32_t g = 0; g < 3; ++g) {
      offs[g].assign(box.size() + 1, 0);
      for (size_t b = 0; b < box.size(); ++b) offs[g][b + 1] = offs[g][b] + block_len((int64_t)b, g);
    }
  }
  int64_t offset(int64_t b, uint32_t g) const {
    if (g < 3) return offs[g][(size_t)b];
    int64_t o = 0;
    for (int64_t i = 0; i < b; ++i) o += block_len(i, g);
    return o;
  }
  int64_t field_len(uint32_t g) const { return offset(nblocks(), g); }
  // AmrGrid::find_patch (src/geometry/amr/grid.rs:264-278); -1 for None.
  int64_t find_patch(int l, int64_t x, int64_t y) const {
    if (l < 0 || l >= (int)lookup.size()) return -1;
    Box2 dom = domain(l);
    if (!dom.contains(x, y)) return -1;
    uint32_t pid = lookup[l][(size_t)(x + y * dom.hi[0])];
    return pid == UINT32_MAX ? -1 : (int64_t)pid;
  }
  // A patch as a one-block Cartesian grid (extent, level spacing) so the
  // per-block kernels above apply unchanged (src/discretization/*: the AMR
  // Stencil impls call the same kernels with grid.spacing(block)).
  orc_grid patch_grid(int64_t b) const {
    orc_grid g{};
    g.dim = 2;
    for (int d = 0; d < 2; ++d) {
      g.cells[d] = ext(b, d);
      g.block_cells[d] = ext(b, d);
      g.spacing[d] = spacing[level[b]][d];
      g.origin[d] = 0.0;
    }
    return g;
  }
};

namespace {

// covered_by (grid.rs:329-356): every cell of `need` in some box of `cover`.
bool covered_by(const Box2 &need, const std::vector<Box2> &cover) {
  if (need.cells() == 0) return true;
  int64_t c[2] = {need.lo[0], need.lo[1]};
  for (;;) {
    const Box2 *hit = nullptr;
    for (const Box2 &b : cover)
      if (b.contains(c[0], c[1])) {
        hit = &b;
        break;
      }
    if (!hit) return false;
    c[0] = hit->hi[0] - 1;
    int d = 0;
    for (;;) {
      c[d] += 1;
      if (c[d] < need.hi[d]) break;
      c[d] = need.lo[d];
      d += 1;
      if (d == 2) return true;
    }
  }
}

// AmrGrid::from_patches (grid.rs:90-196). Returns 0 or an error code:
// 1 ratio, 2 empty patch, 3 not nested, 4 misaligned, 5 overlap.
int amr_build(const orc_grid &base, const std::vector<uint32_t> &ratios,
              const std::vector<std::vector<Box2>> &levels, orc_amr &out) {
  if (levels.size() > ratios.size()) return 1;
  for (uint32_t r : ratios)
    if (r < 2) return 1;
  out.base = base;
  out.ratios = ratios;
  out.level.clear();
  out.box.clear();
  int64_t nb = orc_num_blocks(&base);
  for (int64_t b = 0; b < nb; ++b) {
    int64_t c[3];
    block_coords(&base, b, c);
    Box2 bx{{c[0] * base.block_cells[0], c[1] * base.block_cells[1]},
            {(c[0] + 1) * base.block_cells[0], (c[1] + 1) * base.block_cells[1]}};
    out.level.push_back(0);
    out.box.push_back(bx);
  }
  out.level_start = {0u, (uint32_t)out.box.size()};
  std::vector<Box2> coarser(out.box);
  for (size_t l = 0; l < levels.size(); ++l) {
    int lev = (int)l + 1;
    int64_t r = ratios[l];
    Box2 dom = out.domain(lev);
    const std::vector<Box2> &boxes = levels[l];
    for (size_t i = 0; i < boxes.size(); ++i) {
      const Box2 &bx = boxes[i];
      if (bx.cells() == 0) return 2;
      if (!(bx.clipped(dom) == bx)) return 3;
      for (int d = 0; d < 2; ++d)
        if (bx.lo[d] % r != 0 || bx.hi[d] % r != 0) return 4;
      for (size_t k = 0; k < i; ++k)
        if (bx.intersects(boxes[k])) return 5;
      Box2 coarse{{div_euclid(bx.lo[0], r), div_euclid(bx.lo[1], r)},
                  {div_euclid(bx.hi[0] + r - 1, r), div_euclid(bx.hi[1] + r - 1, r)}};
      Box2 need = coarse.grown(1).clipped(out.domain(lev - 1));
      if (!covered_by(need, coarser)) return 3;
    }
    for (const Box2 &bx : boxes) {
      out.level.push_back(lev);
      out.box.push_back(bx);
    }
    out.level_start.push_back((uint32_t)out.box.size());
    coarser = boxes;
  }
  out.spacing.clear();
  out.spacing.push_back({base.spacing[0], base.spacing[1]});
  for (uint32_t r : ratios) {
    auto prev = out.spacing.back();
    out.spacing.push_back({prev[0] / (double)r, prev[1] / (double)r});
  }
  out.lookup.clear();
  for (int l = 0; l < out.num_levels(); ++l) {
    Box2 dom = out.domain(l);
    std::vector<uint32_t> map((size_t)(dom.hi[0] * dom.hi[1]), UINT32_MAX);
    for (uint32_t p = out.level_start[l]; p < out.level_start[l + 1]; ++p) {
      const Box2 &b = out.box[p];
      for (int64_t y = b.lo[1]; y < b.hi[1]; ++y)
        for (int64_t x = b.lo[0]; x < b.hi[0]; ++x) map[(size_t)(x + y * dom.hi[0])] = p;
    }
    out.lookup.push_back(std::move(map));
  }
  out.build_offsets();
  return 0;
}

inline double &at(const orc_amr &g, double *f, uint32_t ghost, int64_t b, int64_t i, int64_t j) {
  int64_t sx = g.ext(b, 0) + 2 * (int64_t)ghost;
  return f[g.offset(b, ghost) + (j + ghost) * sx + (i + ghost)];
}

// restrict_level (src/geometry/amr/intergrid.rs:110-150).
void amr_restrict_level(const orc_amr &g, double *f, uint32_t ghost, int fine_level) {
  int64_t r = g.ratios[fine_level - 1];
  double inv = 1.0 / std::pow((double)r, 2);
  for (uint32_t fb = g.level_start[fine_level]; fb < g.level_start[fine_level + 1]; ++fb) {
    const Box2 &fp = g.box[fb];
    Box2 covered{{fp.lo[0] / r, fp.lo[1] / r}, {fp.hi[0] / r, fp.hi[1] / r}};
    for (uint32_t cb = g.level_start[fine_level - 1]; cb < g.level_start[fine_level]; ++cb) {
      const Box2 &cp = g.box[cb];
      Box2 ov = covered.clipped(cp);
      if (ov.cells() == 0) continue;
      for (int64_t cy = ov.lo[1]; cy < ov.hi[1]; ++cy)
        for (int64_t cx = ov.lo[0]; cx < ov.hi[0]; ++cx) {
          double sum = 0.0;
          for (int64_t ky = cy * r - fp.lo[1]; ky < (cy + 1) * r - fp.lo[1]; ++ky)
            for (int64_t kx = cx * r - fp.lo[0]; kx < (cx + 1) * r - fp.lo[0]; ++kx)
              sum = sum + at(g, f, ghost, fb, kx, ky);
          at(g, f, ghost, cb, cx - cp.lo[0], cy - cp.lo[1]) = sum * inv;
        }
    }
  }
}

// prolonged (intergrid.rs:414-459): bilinear from level-1 interiors.
double amr_prolonged(const orc_amr &g, const double *f, uint32_t ghost, int level, int64_t cx,
                     int64_t cy) {
  double r = (double)g.ratios[level - 1];
  Box2 cd = g.domain(level - 1);
  int64_t c0[2];
  double w[2];
  int64_t cell[2] = {cx, cy};
  for (int d = 0; d < 2; ++d) {
    double pc = ((double)cell[d] + 0.5) / r - 0.5;
    double fl = std::floor(pc);
    c0[d] = (int64_t)fl;
    w[d] = pc - fl;
  }
  double acc = 0.0;
  for (unsigned corner = 0; corner < 4; ++corner) {
    int64_t cc[2] = {c0[0], c0[1]};
    double weight = 1.0;
    for (int d = 0; d < 2; ++d) {
      if ((corner >> d) & 1u) {
        cc[d] += 1;
        weight *= w[d];
      } else {
        weight *= 1.0 - w[d];
      }
    }
    if (weight == 0.0) continue;
    for (int d = 0; d < 2; ++d) cc[d] = reflect(cc[d], cd.lo[d], cd.hi[d]);
    int64_t qb = g.find_patch(level - 1, cc[0], cc[1]);
    if (qb < 0) std::abort();  // proper nesting guarantees a donor
    const Box2 &q = g.box[qb];
    acc = acc + at(g, const_cast<double *>(f), ghost, qb, cc[0] - q.lo[0], cc[1] - q.lo[1]) * weight;
  }
  return acc;
}

// fill_one_level / fill_level (intergrid.rs:183-223, 358-408): gather every
// ring value of a block, then write.
void amr_fill_level(const orc_amr &g, double *f, const double *old, double alpha, uint32_t ghost,
                    int level) {
  if (ghost == 0) return;
  Box2 dom = g.domain(level);
  int64_t gw = ghost;
  std::vector<std::pair<std::array<int64_t, 2>, double>> staged;
  for (uint32_t pb = g.level_start[level]; pb < g.level_start[level + 1]; ++pb) {
    const Box2 &p = g.box[pb];
    staged.clear();
    auto visit = [&](int64_t x, int64_t y) {
      int64_t tx = reflect(x, dom.lo[0], dom.hi[0]), ty = reflect(y, dom.lo[1], dom.hi[1]);
      int64_t src = g.find_patch(level, tx, ty);
      double v;
      if (src >= 0) {
        const Box2 &sp = g.box[src];
        v = at(g, f, ghost, src, tx - sp.lo[0], ty - sp.lo[1]);
      } else {
        double vn = amr_prolonged(g, f, ghost, level, tx, ty);
        if (old) {
          double vo = amr_prolonged(g, old, ghost, level, tx, ty);
          v = vo * (1.0 - alpha) + vn * alpha;
        } else {
          v = vn;
        }
      }
      staged.push_back({{x - p.lo[0], y - p.lo[1]}, v});
    };
    // for_each_ring_cell (intergrid.rs:73-86).
    for (int d = 0; d < 2; ++d) {
      for (int side = 0; side < 2; ++side) {
        Box2 s = p;
        for (int e = 0; e < d; ++e) {
          s.lo[e] = p.lo[e] - gw;
          s.hi[e] = p.hi[e] + gw;
        }
        s.lo[d] = side ? p.hi[d] : p.lo[d] - gw;
        s.hi[d] = side ? p.hi[d] + gw : p.lo[d];
        for (int64_t y = s.lo[1]; y < s.hi[1]; ++y)
          for (int64_t x = s.lo[0]; x < s.hi[0]; ++x) visit(x, y);
      }
    }
    for (auto &e : staged) at(g, f, ghost, pb, e.first[0], e.first[1]) = e.second;
  }
}

void amr_fill_all(const orc_amr &g, double *f, uint32_t ghost) {
  for (int l = 0; l < g.num_levels(); ++l) amr_fill_level(g, f, nullptr, 0.0, ghost, l);
}

void amr_restrict_all(const orc_amr &g, double *f, uint32_t ghost) {
  for (int l = g.num_levels() - 1; l >= 1; --l) amr_restrict_level(g, f, ghost, l);
}

// ---- Berger–Rigoutsos (src/geometry/amr/cluster.rs:144-281) ----
using Cell = std::array<int64_t, 2>;

Box2 bounding(const Cell *c, size_t n) {
  Box2 b{{c[0][0], c[0][1]}, {c[0][0], c[0][1]}};
  for (size_t i = 0; i < n; ++i)
    for (int d = 0; d < 2; ++d) {
      b.lo[d] = std::min(b.lo[d], c[i][d]);
      b.hi[d] = std::max(b.hi[d], c[i][d]);
    }
  b.hi[0] += 1;
  b.hi[1] += 1;
  return b;
}

bool find_hole(const Box2 &bx, const std::vector<uint32_t> sig[2], int &axis, int64_t &plane) {
  bool found = false;
  uint64_t bestc = 0;
  for (int d = 0; d < 2; ++d) {
    size_t n = sig[d].size();
    for (size_t i = 1; i + 1 < n; ++i) {
      if (sig[d][i] == 0) {
        uint64_t c = std::min((uint64_t)i, (uint64_t)(n - 1 - i));
        if (!found || c > bestc) {
          found = true;
          bestc = c;
          axis = d;
          plane = bx.lo[d] + (int64_t)i;
        }
      }
    }
  }
  return found;
}

bool find_inflection(const Box2 &bx, const std::vector<uint32_t> sig[2], size_t min_width,
                     int &axis, int64_t &plane) {
  bool found = false;
  int64_t bests = 0;
  uint64_t bestc = 0;
  for (int d = 0; d < 2; ++d) {
    size_t n = sig[d].size();
    if (n < 4) continue;
    auto delta = [&](size_t i) {
      return (int64_t)sig[d][i + 1] - 2 * (int64_t)sig[d][i] + (int64_t)sig[d][i - 1];
    };
    for (size_t i = 1; i < n - 2; ++i) {
      int64_t a = delta(i), b = delta(i + 1);
      if (a * b < 0) {
        size_t pl = i + 1;
        if (pl < min_width || n - pl < min_width) continue;
        int64_t s = std::llabs(b - a);
        uint64_t c = std::min((uint64_t)pl, (uint64_t)(n - pl));
        if (!found || s > bests || (s == bests && c > bestc)) {
          found = true;
          bests = s;
          bestc = c;
          axis = d;
          plane = bx.lo[d] + (int64_t)pl;
        }
      }
    }
  }
  return found;
}

bool bisection(const Box2 &bx, size_t min_width, int &axis, int64_t &plane) {
  // Iterator::max_by_key keeps the last maximum.
  int d = 0;
  for (int e = 1; e < 2; ++e)
    if (bx.extent(e) >= bx.extent(d)) d = e;
  if ((uint64_t)bx.extent(d) < 2 * (uint64_t)min_width) return false;
  axis = d;
  plane = bx.lo[d] + bx.extent(d) / 2;
  return true;
}

void split_or_accept(Cell *flags, size_t n, double efficiency, size_t min_width,
                     std::vector<Box2> &out) {
  Box2 bx = bounding(flags, n);
  double eff = (double)n / (double)bx.cells();
  if (eff >= efficiency) {
    out.push_back(bx);
    return;
  }
  std::vector<uint32_t> sig[2];
  for (int d = 0; d < 2; ++d) sig[d].assign((size_t)bx.extent(d), 0u);
  for (size_t i = 0; i < n; ++i)
    for (int d = 0; d < 2; ++d) sig[d][(size_t)(flags[i][d] - bx.lo[d])] += 1;
  int axis = 0;
  int64_t plane = 0;
  if (!find_hole(bx, sig, axis, plane) && !find_inflection(bx, sig, min_width, axis, plane) &&
      !bisection(bx, min_width, axis, plane)) {
    out.push_back(bx);
    return;
  }
  size_t lt = 0;
  for (size_t j = 0; j < n; ++j)
    if (flags[j][axis] < plane) {
      std::swap(flags[lt], flags[j]);
      lt += 1;
    }
  split_or_accept(flags, lt, efficiency, min_width, out);
  split_or_accept(flags + lt, n - lt, efficiency, min_width, out);
}

// GradientTagger::tag_level (src/geometry/amr/adapt.rs:78-103).
void tag_level(const orc_amr &g, const double *f, int level, double thr, std::vector<Cell> &flags) {
  for (uint32_t pb = g.level_start[level]; pb < g.level_start[level + 1]; ++pb) {
    const Box2 &p = g.box[pb];
    for (int64_t j = 0; j < p.extent(1); ++j)
      for (int64_t i = 0; i < p.extent(0); ++i) {
        double gm = 0.0;
        double *ff = const_cast<double *>(f);
        double dx = std::fabs((at(g, ff, 1, pb, i + 1, j) - at(g, ff, 1, pb, i - 1, j)) * 0.5);
        gm = std::fmax(gm, dx);
        double dy = std::fabs((at(g, ff, 1, pb, i, j + 1) - at(g, ff, 1, pb, i, j - 1)) * 0.5);
        gm = std::fmax(gm, dy);
        if (gm > thr) flags.push_back({p.lo[0] + i, p.lo[1] + j});
      }
  }
}

void push_cells(const Box2 &b, std::vector<Cell> &out) {
  if (b.lo[0] >= b.hi[0] || b.lo[1] >= b.hi[1]) return;
  for (int64_t y = b.lo[1]; y < b.hi[1]; ++y)
    for (int64_t x = b.lo[0]; x < b.hi[0]; ++x) out.push_back({x, y});
}

// The nesting flags of one box of the level just built (adapt.rs:194-211): the box
// coarsened to the level being built (ratio rf, truncating division) and grown by
// one cell, then the source-level cells under that (ratio rs, floor / ceiling
// division).  Not clipped: the caller clips to the source domain.
Box2 nesting_parents(const Box2 &bx, int64_t rf, int64_t rs) {
  Box2 need{{bx.lo[0] / rf - 1, bx.lo[1] / rf - 1}, {bx.hi[0] / rf + 1, bx.hi[1] / rf + 1}};
  return Box2{{div_euclid(need.lo[0], rs), div_euclid(need.lo[1], rs)},
              {div_euclid(need.hi[0] + rs - 1, rs), div_euclid(need.hi[1] + rs - 1, rs)}};
}

// hierarchy_acceptable (adapt.rs:269-290).
bool hierarchy_acceptable(const orc_amr &g, const std::vector<std::vector<Box2>> &nl) {
  if ((int)nl.size() != g.num_levels() - 1) return false;
  for (size_t l = 0; l < nl.size(); ++l) {
    std::vector<Box2> cur(g.box.begin() + g.level_start[l + 1], g.box.begin() + g.level_start[l + 2]);
    uint64_t pc = 0, cc = 0;
    for (auto &b : nl[l]) pc += b.cells();
    for (auto &b : cur) cc += b.cells();
    if (cc > pc + pc / 2) return false;
    for (auto &b : nl[l])
      if (!covered_by(b, cur)) return false;
  }
  return true;
}

// sample (adapt.rs:352-402).
double amr_sample(const orc_amr &g, const double *f, uint32_t ghost, int level, int64_t x,
                  int64_t y) {
  if (level < g.num_levels()) {
    int64_t b = g.find_patch(level, x, y);
    if (b >= 0) return at(g, const_cast<double *>(f), ghost, b, x - g.box[b].lo[0], y - g.box[b].lo[1]);
  }
  double r = (double)g.ratios[level - 1];
  Box2 cd = g.domain(level - 1);
  int64_t c0[2];
  double w[2];
  int64_t cell[2] = {x, y};
  for (int d = 0; d < 2; ++d) {
    double pc = ((double)cell[d] + 0.5) / r - 0.5;
    double fl = std::floor(pc);
    c0[d] = (int64_t)fl;
    w[d] = pc - fl;
  }
  double acc = 0.0;
  for (unsigned corner = 0; corner < 4; ++corner) {
    int64_t cc[2] = {c0[0], c0[1]};
    double weight = 1.0;
    for (int d = 0; d < 2; ++d) {
      if ((corner >> d) & 1u) {
        cc[d] += 1;
        weight *= w[d];
      } else {
        weight *= 1.0 - w[d];
      }
    }
    if (weight == 0.0) continue;
    for (int d = 0; d < 2; ++d) cc[d] = reflect(cc[d], cd.lo[d], cd.hi[d]);
    acc = acc + amr_sample(g, f, ghost, level - 1, cc[0], cc[1]) * weight;
  }
  return acc;
}

}  // namespace

// ---- simulation on an AMR hierarchy ----
struct orc_tag_sink {
  std::vector<std::array<int64_t, 2>> *flags;
};

struct orc_amr_sim {
  std::unique_ptr<orc_amr> grid;
  int32_t model = 0;  // 0 heat, 1 Model C
  orc_modelc_params p{};
  double kappa = 0.0;
  int32_t integ = 0;  // 0 global dt (FE / EM), 1 Berger–Oliger subcycling
  uint64_t seed = 0;
  int32_t wiener = 0;
  bool adaptive = false;
  orc_regrid_policy policy{};
  // A caller's TagCells (adapt.rs:32-46) in place of the GradientTagger: its
  // prepare is the trait's default (nothing) and tag_level is the callback.
  orc_tag_level_fn tagger = nullptr;
  void *tagger_user = nullptr;
  std::vector<std::vector<double>> state;  // one buffer per field (ghost 1)
  std::vector<std::vector<double>> k, amp, snaps;
  double t = 0.0;
  int64_t step_index = 0;
  int nfields() const { return model == 1 ? 2 : 1; }
  double stable_dt(double h) const {
    return model == 1 ? 0.2 * std::pow(h, 2) / p.d_thermal : 0.25 * std::pow(h, 2) / kappa;
  }
};

namespace {

void sim_alloc_stages(orc_amr_sim &s) {
  size_t n = (size_t)s.grid->field_len(1);
  s.k.assign((size_t)s.nfields(), std::vector<double>(n, 0.0));
  s.amp.clear();
  if (s.model == 1 && s.wiener) s.amp.assign(1, std::vector<double>(n, 0.0));
  s.snaps.clear();
  if (s.integ == 1) s.snaps.assign((size_t)s.grid->max_levels() * s.nfields(), std::vector<double>(n, 0.0));
}

// ModelC / Heat fill_ghosts on an AmrGrid (phasefield.rs:454-467):
// restrict every field, then fill every field's ghosts.
void sim_fill_global(orc_amr_sim &s) {
  for (int f = 0; f < s.nfields(); ++f) amr_restrict_all(*s.grid, s.state[f].data(), 1);
  for (int f = 0; f < s.nfields(); ++f) amr_fill_all(*s.grid, s.state[f].data(), 1);
}

// vector_field_block(Time) on one patch.
void sim_rhs_block(orc_amr_sim &s, int64_t b) {
  const orc_amr &g = *s.grid;
  orc_grid pg = g.patch_grid(b);
  int64_t off = g.offset(b, 1);
  if (s.model == 1) {
    modelc_rhs_block(&pg, &s.p, s.state[0].data() + off, s.state[1].data() + off,
                     s.k[0].data() + off, s.k[1].data() + off);
  } else {
    double *kb = s.k[0].data() + off;
    laplacian_block(&pg, s.state[0].data() + off, kb, 1);
    for_each_interior(&pg, [&](const int64_t *idx) {
      int64_t o = lin(&pg, idx, 1);
      kb[o] = kb[o] * s.kappa;
    });
  }
}

// apply_step on one patch: Time then Wiener(0) (state.rs:656-753).
void sim_apply_block(orc_amr_sim &s, int64_t b, double dt, uint64_t salt) {
  const orc_amr &g = *s.grid;
  int64_t off = g.offset(b, 1), len = g.block_len(b, 1);
  for (int f = 0; f < s.nfields(); ++f)
    orc_axpy(s.state[f].data() + off, s.k[f].data() + off, dt, (size_t)len);
  if (s.model == 1 && s.wiener) {
    orc_grid pg = g.patch_grid(b);
    double *a = s.amp[0].data() + off;
    // eval_tendency: zero the slab, interior = b.
    for (int64_t i = 0; i < len; ++i) a[i] = 0.0;
    for_each_interior(&pg, [&](const int64_t *idx) { a[lin(&pg, idx, 1)] = s.p.noise_amplitude; });
    wiener_block(&pg, b, s.state[0].data() + off, a, 1, dt, s.seed, salt, (1ull << 32) | 0ull);
  }
}

void sim_step_global(orc_amr_sim &s, double dt) {
  uint64_t salt;
  std::memcpy(&salt, &s.t, sizeof salt);
  sim_fill_global(s);
  for (int64_t b = 0; b < s.grid->nblocks(); ++b) sim_rhs_block(s, b);
  // EM evaluates every amplitude before any update (Itô); the Model C
  // amplitude does not depend on the state, so evaluation order is moot.
  for (int64_t b = 0; b < s.grid->nblocks(); ++b) sim_apply_block(s, b, dt, salt);
}

// Subcycle::advance (src/integrators/subcycling.rs:195-267).
void sim_advance(orc_amr_sim &s, int level, double t, double dt, double alpha) {
  const orc_amr &g = *s.grid;
  int nf = s.nfields();
  for (int f = 0; f < nf; ++f) s.snaps[(size_t)level * nf + f] = s.state[f];
  for (int f = 0; f < nf; ++f) {
    const double *old = level > 0 ? s.snaps[(size_t)(level - 1) * nf + f].data() : nullptr;
    amr_fill_level(g, s.state[f].data(), old, alpha, 1, level);
  }
  uint64_t salt;
  std::memcpy(&salt, &t, sizeof salt);
  for (uint32_t b = g.level_start[level]; b < g.level_start[level + 1]; ++b) sim_rhs_block(s, b);
  for (uint32_t b = g.level_start[level]; b < g.level_start[level + 1]; ++b)
    sim_apply_block(s, b, dt, salt);
  int child = level + 1;
  if (child < g.num_levels()) {
    double dtc = s.stable_dt(std::fmin(g.spacing[level][0], g.spacing[level][1]));
    double dtf = s.stable_dt(std::fmin(g.spacing[child][0], g.spacing[child][1]));
    double nn = std::fmax(std::round(dtc / dtf), 1.0);
    uint64_t n = (uint64_t)nn;
    double dt_child = dt / (double)n;
    for (uint64_t k = 0; k < n; ++k)
      sim_advance(s, child, std::fma((double)k, dt_child, t), dt_child, (double)k / (double)n);
    for (int f = 0; f < nf; ++f) amr_restrict_level(g, s.state[f].data(), 1, child);
  }
}

// BergerOliger::regrid (adapt.rs:152-237) with GradientTagger; returns 1 when the
// hierarchy was replaced, 0 when kept, -1 when the tagger failed or the proposal is
// not a valid hierarchy (the reference panics at adapt.rs:227-230).
int sim_regrid(orc_amr_sim &s) {
  const orc_regrid_policy &pol = s.policy;
  if (s.step_index % pol.every != 0) return 0;
  orc_amr &g = *s.grid;
  double *tf = s.state[(size_t)pol.tag_field].data();
  if (!s.tagger) {  // GradientTagger::prepare (adapt.rs:72-76)
    amr_restrict_all(g, tf, 1);
    amr_fill_all(g, tf, 1);
  }
  int capacity = g.max_levels() - 1;
  std::vector<std::vector<Box2>> nl((size_t)capacity);
  for (int target = capacity; target >= 1; --target) {
    int src = target - 1;
    std::vector<Cell> flags;
    if (src < g.num_levels()) {
      if (s.tagger) {
        orc_tag_sink sink{&flags};
        if (s.tagger(s.tagger_user, &g, src, &sink) != 0) return -1;
      } else {
        tag_level(g, tf, src, pol.threshold, flags);
      }
    }
    if (pol.buffer > 0) {
      std::vector<Cell> seeds;
      seeds.swap(flags);
      for (const Cell &c : seeds)
        push_cells(Box2{{c[0] - pol.buffer, c[1] - pol.buffer},
                        {c[0] + pol.buffer + 1, c[1] + pol.buffer + 1}},
                   flags);
    }
    if (target < capacity && !nl[(size_t)target].empty()) {
      int64_t rf = g.ratios[(size_t)target], rs = g.ratios[(size_t)target - 1];
      for (const Box2 &bx : nl[(size_t)target]) push_cells(nesting_parents(bx, rf, rs), flags);
    }
    Box2 dom = g.domain(src);
    std::vector<Cell> kept;
    for (const Cell &c : flags)
      if (dom.contains(c[0], c[1])) kept.push_back(c);
    std::sort(kept.begin(), kept.end());
    kept.erase(std::unique(kept.begin(), kept.end()), kept.end());
    if (kept.empty()) continue;
    int64_t r = g.ratios[(size_t)target - 1];
    std::vector<Box2> boxes;
    split_or_accept(kept.data(), kept.size(), pol.efficiency, (size_t)pol.min_width, boxes);
    for (Box2 &b : boxes)
      for (int d = 0; d < 2; ++d) {
        b.lo[d] *= r;
        b.hi[d] *= r;
      }
    nl[(size_t)target - 1] = boxes;
  }
  while (!nl.empty() && nl.back().empty()) nl.pop_back();
  if (hierarchy_acceptable(g, nl)) return 0;
  auto ng = std::make_unique<orc_amr>();
  if (amr_build(g.base, g.ratios, nl, *ng) != 0) return -1;
  // State::reshaped + migrate (adapt.rs:323-347).
  std::vector<std::vector<double>> ns;
  for (int f = 0; f < s.nfields(); ++f) {
    std::vector<double> out((size_t)ng->field_len(1), 0.0);
    for (uint32_t b = 0; b < ng->level_start[1]; ++b) {
      int64_t len = ng->block_len(b, 1);
      std::memcpy(out.data() + ng->offset(b, 1), s.state[f].data() + g.offset(b, 1),
                  sizeof(double) * (size_t)len);
    }
    for (int l = 1; l < ng->num_levels(); ++l)
      for (uint32_t b = ng->level_start[l]; b < ng->level_start[l + 1]; ++b) {
        const Box2 &p = ng->box[b];
        for (int64_t j = 0; j < p.extent(1); ++j)
          for (int64_t i = 0; i < p.extent(0); ++i)
            at(*ng, out.data(), 1, b, i, j) = amr_sample(g, s.state[f].data(), 1, l, p.lo[0] + i, p.lo[1] + j);
      }
    ns.push_back(std::move(out));
  }
  s.grid = std::move(ng);
  s.state = std::move(ns);
  sim_alloc_stages(s);
  return 1;
}

}  // namespace

extern "C" {

int orc_amr_create(const orc_grid *base, const uint32_t *ratios, size_t nratios,
                   const int64_t *boxes, size_t nboxes, orc_amr **out) {
  *out = nullptr;
  if (base->dim != 2 || orc_grid_check(base) != 0) return -1;
  std::vector<uint32_t> r(ratios, ratios + nratios);
  std::vector<std::vector<Box2>> levels;
  for (size_t i = 0; i < nboxes; ++i) {
    int64_t lev = boxes[5 * i];
    if (lev < 1) return -1;
    while ((int64_t)levels.size() < lev) levels.emplace_back();
    levels[(size_t)lev - 1].push_back(
        Box2{{boxes[5 * i + 1], boxes[5 * i + 2]}, {boxes[5 * i + 3], boxes[5 * i + 4]}});
  }
  auto g = std::make_unique<orc_amr>();
  int rc = amr_build(*base, r, levels, *g);
  if (rc != 0) return rc;
  *out = g.release();
  return 0;
}

void orc_amr_destroy(orc_amr *g) { delete g; }
int64_t orc_amr_num_blocks(const orc_amr *g) { return g->nblocks(); }
int32_t orc_amr_num_levels(const orc_amr *g) { return g->num_levels(); }
void orc_amr_patch(const orc_amr *g, int64_t b, int64_t *out) {
  out[0] = g->level[b];
  out[1] = g->box[b].lo[0];
  out[2] = g->box[b].lo[1];
  out[3] = g->box[b].hi[0];
  out[4] = g->box[b].hi[1];
}
int64_t orc_amr_block_offset(const orc_amr *g, int64_t b, uint32_t ghost) { return g->offset(b, ghost); }
int64_t orc_amr_field_len(const orc_amr *g, uint32_t ghost) { return g->field_len(ghost); }
int64_t orc_amr_find_patch(const orc_amr *g, int32_t level, int64_t x, int64_t y) {
  return g->find_patch(level, x, y);
}
void orc_amr_cell_center(const orc_amr *g, int64_t b, int64_t i, int64_t j, double *out) {
  // AmrGrid::cell_center (grid.rs:397-405).
  int l = g->level[b];
  out[0] = std::fma((double)(g->box[b].lo[0] + i) + 0.5, g->spacing[l][0], g->base.origin[0]);
  out[1] = std::fma((double)(g->box[b].lo[1] + j) + 0.5, g->spacing[l][1], g->base.origin[1]);
}
void orc_amr_restrict_level(const orc_amr *g, double *f, uint32_t ghost, int32_t fine_level) {
  amr_restrict_level(*g, f, ghost, fine_level);
}
void orc_amr_restrict(const orc_amr *g, double *f, uint32_t ghost) { amr_restrict_all(*g, f, ghost); }
void orc_amr_fill_level(const orc_amr *g, double *f, const double *old, double alpha,
                        uint32_t ghost, int32_t level) {
  amr_fill_level(*g, f, old, alpha, ghost, level);
}
void orc_amr_fill_ghosts(const orc_amr *g, double *f, uint32_t ghost) { amr_fill_all(*g, f, ghost); }

int orc_amr_sim_create(orc_amr *grid, int32_t model, const orc_modelc_params *p, double kappa,
                       int32_t integ, uint64_t seed, int32_t wiener,
                       const orc_regrid_policy *policy, orc_amr_sim **out) {
  auto s = new orc_amr_sim();
  s->grid.reset(grid);
  s->model = model;
  if (p) s->p = *p;
  s->kappa = kappa;
  s->integ = integ;
  s->seed = seed;
  s->wiener = wiener;
  if (policy) {
    s->adaptive = true;
    s->policy = *policy;
  }
  size_t n = (size_t)s->grid->field_len(1);
  s->state.assign((size_t)s->nfields(), std::vector<double>(n, 0.0));
  sim_alloc_stages(*s);
  *out = s;
  return 0;
}

void orc_amr_sim_destroy(orc_amr_sim *s) { delete s; }
const orc_amr *orc_amr_sim_grid(const orc_amr_sim *s) { return s->grid.get(); }
double *orc_amr_sim_field(orc_amr_sim *s, int32_t f) { return s->state[(size_t)f].data(); }
double orc_amr_sim_time(const orc_amr_sim *s) { return s->t; }

void orc_amr_sim_set_tagger(orc_amr_sim *s, orc_tag_level_fn fn, void *user) {
  s->tagger = fn;
  s->tagger_user = user;
}

int orc_tag_sink_push(orc_tag_sink *sink, const int64_t *cells, size_t n) {
  for (size_t i = 0; i < n; ++i) sink->flags->push_back({cells[2 * i], cells[2 * i + 1]});
  return 0;
}

void orc_nesting_parents(const int64_t *box4, int64_t rf, int64_t rs, int64_t *out4) {
  Box2 p = nesting_parents(Box2{{box4[0], box4[1]}, {box4[2], box4[3]}}, rf, rs);
  out4[0] = p.lo[0];
  out4[1] = p.lo[1];
  out4[2] = p.hi[0];
  out4[3] = p.hi[1];
}

int orc_amr_sim_step(orc_amr_sim *s, double dt, int64_t nsteps) {
  for (int64_t n = 0; n < nsteps; ++n) {
    if (s->adaptive && sim_regrid(*s) < 0) return -1;
    if (s->integ == 1)
      sim_advance(*s, 0, s->t, dt, 0.0);
    else
      sim_step_global(*s, dt);
    s->t = s->t + dt;
    s->step_index += 1;
  }
  return 0;
}

}  // extern "C"

extern "C" size_t orc_cluster(int64_t *flags, size_t n, double efficiency, size_t min_width,
                              int64_t *out, size_t max_out) {
  std::vector<Cell> cells(n);
  for (size_t i = 0; i < n; ++i) cells[i] = {flags[2 * i], flags[2 * i + 1]};
  std::vector<Box2> boxes;
  if (n) split_or_accept(cells.data(), n, efficiency, min_width, boxes);
  for (size_t i = 0; i < n; ++i) {
    flags[2 * i] = cells[i][0];
    flags[2 * i + 1] = cells[i][1];
  }
  for (size_t i = 0; i < boxes.size() && i < max_out; ++i) {
    out[4 * i] = boxes[i].lo[0];
    out[4 * i + 1] = boxes[i].lo[1];
    out[4 * i + 2] = boxes[i].hi[0];
    out[4 * i + 3] = boxes[i].hi[1];
  }
  return boxes.size();
}

// ---- Driver::Jump and the multi-driver explicit update ----

namespace {

// Driver::Jump arm of Driver::apply_slab (src/core/driver.rs:159-182) over the
// fields one jump driver moves, on one block: one fire per interior cell shared by
// every field, thinned against the per-cell intensity (ghost-0 layout, indexed by
// Grid::cell_key) as Bernoulli(1 − e^{−λ·dt}).
void jump_block(const orc_grid *g, int64_t block, size_t nfields, double *const *x,
                const double *const *inc, const uint32_t *ghost, const double *rate, double dt,
                uint64_t seed, uint64_t salt, uint64_t stream) {
  uint64_t parts[3] = {salt, stream, (uint64_t)block};
  uint64_t block_key = orc_mix_key(seed, parts, 3);
  for (size_t f = 0; f < nfields; ++f) {
    if (!inc[f]) continue;
    int64_t len = orc_block_len(g, ghost[f]);
    for (int64_t i = 0; i < len; ++i) {
      int64_t cell = orc_cell_key(g, ghost[f], i);
      if (cell < 0) continue;
      double lambda = rate[cell];
      if (lambda <= 0.0) continue;
      uint64_t key = orc_splitmix64(block_key ^ (uint64_t)cell);
      double p = -std::expm1(-lambda * dt);
      if (orc_unit_open(key) < p) x[f][i] = x[f][i] + inc[f][i];
    }
  }
}

// Driver::stream (src/core/driver.rs:55-61).
uint64_t driver_stream(int32_t kind, uint32_t index) {
  return kind == ORC_DRIVER_TIME ? 0ull : ((uint64_t)kind << 32) | (uint64_t)index;
}

}  // namespace

extern "C" {

int orc_apply_drivers(const orc_grid *g, size_t nfields, double *const *x, const uint32_t *ghost,
                      size_t ndrivers, const orc_driver *drivers, double dt, uint64_t seed,
                      uint64_t salt, int32_t nthreads) {
  if (orc_grid_check(g) != 0) return -1;
  for (size_t d = 0; d < ndrivers; ++d) {
    const int32_t k = drivers[d].kind;
    if (k != ORC_DRIVER_TIME && k != ORC_DRIVER_WIENER && k != ORC_DRIVER_JUMP) return -1;
    if (k == ORC_DRIVER_JUMP && !drivers[d].rate) return -1;  // driver.rs:160
  }
  const int64_t nblocks = orc_num_blocks(g);
  const int64_t ncell = orc_block_len(g, 0);
  // State::apply_step_with (src/core/state.rs:646-699): per block, the drivers in
  // order, each over every field it moves (block_apply_driver, 513-544).
#pragma omp parallel for schedule(static) num_threads(nthreads > 0 ? nthreads : 1) if (nthreads > 1)
  for (int64_t b = 0; b < nblocks; ++b) {
    std::vector<double *> xs(nfields);
    std::vector<const double *> as(nfields);
    for (size_t d = 0; d < ndrivers; ++d) {
      const orc_driver &dr = drivers[d];
      for (size_t f = 0; f < nfields; ++f) {
        const int64_t len = orc_block_len(g, ghost[f]);
        xs[f] = x[f] + b * len;
        as[f] = dr.amp[f] ? dr.amp[f] + b * len : nullptr;
      }
      const uint64_t stream = driver_stream(dr.kind, dr.index);
      if (dr.kind == ORC_DRIVER_TIME) {
        for (size_t f = 0; f < nfields; ++f)
          if (as[f]) orc_axpy(xs[f], as[f], dt, (size_t)orc_block_len(g, ghost[f]));
      } else if (dr.kind == ORC_DRIVER_WIENER) {
        for (size_t f = 0; f < nfields; ++f)
          if (as[f]) wiener_block(g, b, xs[f], as[f], ghost[f], dt, seed, salt, stream);
      } else {
        jump_block(g, b, nfields, xs.data(), as.data(), ghost, dr.rate + b * ncell, dt, seed, salt,
                   stream);
      }
    }
  }
  return 0;
}

int orc_jump_diffusion_run(const orc_grid *g, const double *lam, double *w, double *count,
                           double *jsum, int64_t nsteps, double dt, double *t, uint64_t seed,
                           double mu, double sigma, double payoff, int32_t nthreads) {
  if (orc_grid_check(g) != 0) return -1;
  const size_t n = (size_t)(orc_num_blocks(g) * orc_block_len(g, 0));
  // The vector fields of the jump-diffusion model (tests/jump.rs:48-86): drift μ
  // and amplitude σ of w, increments 1 and the payoff of (count, jsum) on a fire,
  // intensity λ from the static field. Every field has ghost width 0, so every
  // slab entry is an interior cell.
  std::vector<double> drift(n, mu), amp(n, sigma), one(n, 1.0), pay(n, payoff);
  double *x[3] = {w, count, jsum};
  const uint32_t ghost[3] = {0, 0, 0};
  const double *time_f[3] = {drift.data(), nullptr, nullptr};
  const double *wiener_f[3] = {amp.data(), nullptr, nullptr};
  const double *jump_f[3] = {nullptr, one.data(), pay.data()};
  const orc_driver drivers[3] = {{ORC_DRIVER_TIME, 0, time_f, nullptr},
                                 {ORC_DRIVER_WIENER, 0, wiener_f, nullptr},
                                 {ORC_DRIVER_JUMP, 0, jump_f, lam}};
  for (int64_t s = 0; s < nsteps; ++s) {
    uint64_t salt;
    std::memcpy(&salt, t, sizeof salt);
    if (orc_apply_drivers(g, 3, x, ghost, 3, drivers, dt, seed, salt, nthreads) != 0) return -1;
    *t = *t + dt;
  }
  return 0;
}

}  // extern "C"
