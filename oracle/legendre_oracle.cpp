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
