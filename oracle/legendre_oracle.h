/*
 * legendre CPU oracle — TEST INFRASTRUCTURE ONLY.
 *
 * A C++ restatement of the reference crate's SerialScheduler semantics for the
 * block-structured hot path (counter RNG, Cartesian block topology, ghost fill,
 * 2nd-order stencils, Karma–Rappel Model C, forward-Euler / Euler–Maruyama
 * stage updates, time accumulation, initial conditions).  It is the checker the
 * GPU backend is compared against; nothing in the product path may link it.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load it.
 *
 * Layout contract (identical to the device backend): a *field buffer* holds the
 * ghost-inclusive slabs of every block of one field, block-major
 * (block b at offset b * block_len), each slab dimension-0 fastest
 * (src/geometry/cartesian.rs:9-11, 189-204).
 *
 * Arithmetic contract: built with -ffp-contract=off; std::fma exactly where the
 * reference calls mul_add; operation order as the reference writes it.
 */
#ifndef LEGENDRE_ORACLE_H
#define LEGENDRE_ORACLE_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* CartesianGrid<D> (src/geometry/cartesian.rs:28-35); dims >= dim are unused. */
typedef struct {
  int32_t dim;
  int32_t periodic[3];
  int64_t cells[3];
  int64_t block_cells[3];
  double origin[3];
  double spacing[3];
} orc_grid;

/* FaceBc (src/geometry/boundary.rs:36-54). */
enum { ORC_BC_MIRROR = 0, ORC_BC_DIRICHLET = 1, ORC_BC_FLUX = 2, ORC_BC_EXTRAPOLATE = 3 };
typedef struct {
  int32_t kind;
  double value;
} orc_face_bc;

/* ModelC fields (src/physics/phasefield.rs:55-76) + FiniteVolume::tol
 * (src/discretization/finite_volume.rs:41-50). */
typedef struct {
  double eps4;
  double lambda;
  double d_thermal;
  double noise_amplitude;
  double anisotropy_tol;
  double fv_tol;
} orc_modelc_params;

/* ---- src/util/rng.rs ---- */
uint64_t orc_splitmix64(uint64_t z);
uint64_t orc_mix_key(uint64_t seed, const uint64_t *parts, size_t n);
double orc_unit_open(uint64_t bits);
double orc_standard_normal(uint64_t key);

/* ---- src/geometry/cartesian.rs ---- */
int orc_grid_check(const orc_grid *g);
int64_t orc_num_blocks(const orc_grid *g);
int64_t orc_block_len(const orc_grid *g, uint32_t ghost);
int64_t orc_face_neighbor(const orc_grid *g, int64_t block, int32_t dim, int32_t dir);
int64_t orc_cell_key(const orc_grid *g, uint32_t ghost, int64_t offset);
void orc_cell_center(const orc_grid *g, int64_t block, const int64_t *idx, double *out);

/* fill_ghosts_bc / fill_ghosts_mirror (cartesian.rs:377-401, 439-526).
 * bcs: 2*dim entries [d*2 + (side>0)], NULL = Mirror everywhere.
 * Returns 0, or -1 when the ghost width exceeds a block extent. */
int orc_fill_ghosts(const orc_grid *g, double *field, uint32_t ghost, const orc_face_bc *bcs);

/* central_laplacian_kernel (src/discretization/finite_difference.rs:49-67),
 * interior of out; both buffers ghost-width `ghost`. */
void orc_central_laplacian(const orc_grid *g, const double *in, double *out, uint32_t ghost);

/* KarmaRappelFlux::kernel (src/discretization/finite_volume.rs:248-284), ghost 1. */
void orc_karma_rappel_flux(const orc_grid *g, const double *phi, double *out, double eps4,
                           double tol);

/* ModelC::vector_field_block, Time (phasefield.rs:367-421) and Wiener (361-366).
 * Ghost width 1 for phi and u. The Wiener call first zeroes amp (fill_zero,
 * src/integrators/mod.rs:263). */
void orc_modelc_rhs(const orc_grid *g, const orc_modelc_params *p, const double *phi,
                    const double *u, double *kphi, double *ku);
void orc_modelc_amplitude(const orc_grid *g, const orc_modelc_params *p, double *amp_phi);

/* Heat3 / Heat<D>::vector_field_block (examples/heat3d.rs:63-79): out = lap(u)*kappa. */
void orc_heat_rhs(const orc_grid *g, double kappa, const double *u, double *k);

/* block_axpy over whole slabs (src/core/state.rs:485-498). */
void orc_axpy(double *x, const double *y, double alpha, size_t n);

/* Driver::apply_slab (src/core/driver.rs:115-158) for one field buffer.
 * stream: Driver::stream() of the driver (Time = 0, Wiener(j) = (1<<32)|j). */
void orc_apply_time(double *x, const double *k, double dt, size_t n);
void orc_apply_wiener(const orc_grid *g, double *x, const double *amp, uint32_t ghost, double dt,
                      uint64_t seed, uint64_t salt, uint64_t stream);

/* Simulation::step loops (src/core/simulation.rs:255-290) with
 * EulerMaruyama::step (src/integrators/euler_maruyama.rs:34-75), Model C,
 * on a CartesianGrid<2> with mirror ghosts. wiener = 0 runs ModelC<NoNoise>,
 * 1 runs ModelC<Wiener<1>>.  *t is read as the start time and advanced by
 * t += dt per step.  nthreads > 1 dispatches blocks over OpenMP threads
 * (RayonScheduler equivalent; bitwise identical to serial). */
int orc_modelc_run(const orc_grid *g, const orc_modelc_params *p, double *phi, double *u,
                   int64_t nsteps, double dt, double *t, uint64_t seed, int32_t wiener,
                   int32_t nthreads);

/* ForwardEuler::step (src/integrators/euler.rs:104-130) on Heat<D>; bcs as in
 * orc_fill_ghosts. */
/* Face evaluations whose branch differs between two runs from the same state (params
 * pm and pp, b = 0, nsteps steps): at every step's right-hand side, each interior
 * cell's four face() calls (finite_volume.rs:129-142) and its centre anisotropy
 * (phasefield.rs:159-175) take the flat branch (|∇φ|⁴ <= tol) or not; the count sums
 * the evaluations where the two runs disagree.  Zero means L(θ) is smooth between
 * the two parameter sets along the trajectory. */
int orc_modelc_branch_crossings(const orc_grid *g, const orc_modelc_params *pm,
                                const orc_modelc_params *pp, const double *phi0, const double *u0,
                                int64_t nsteps, double dt, int32_t nthreads, int64_t *crossings);
int orc_heat_run(const orc_grid *g, double kappa, double *u, int64_t nsteps, double dt, double *t,
                 const orc_face_bc *bcs, int32_t nthreads);

/* ModelC::initialize_grains with orientation 0 (phasefield.rs:227-286):
 * grains as [cx, cy, radius] triples. phi, u ghost width 1. */
void orc_initialize_grains(const orc_grid *g, double *phi, double *u, const double *grains,
                           size_t ngrains, double u0);

/* Multi-grain Model C (ModelC::with_orientations): the oriented vector field
 * (finite_volume.rs:188-235, phasefield.rs:378-392) with θ₀ in ghost-0 slabs
 * (nx·ny per block), sin/cos from libm per cell as the reference. */
void orc_modelc_rhs_oriented(const orc_grid *g, const orc_modelc_params *p, const double *phi,
                             const double *u, const double *theta, double *kphi, double *ku);
/* Forward-Euler steps of the oriented model (b = 0). */
int orc_modelc_oriented_run(const orc_grid *g, const orc_modelc_params *p, double *phi, double *u,
                            const double *theta, int64_t nsteps, double dt, double *t);
/* RungeKutta4::step (src/integrators/rk4.rs:30-80) on Model C (theta NULL: axis-
 * aligned) and on Heat<D> (bcs as in orc_fill_ghosts). */
int orc_modelc_rk4_run(const orc_grid *g, const orc_modelc_params *p, double *phi, double *u,
                       const double *theta, int64_t nsteps, double dt, double *t);
int orc_heat_rk4_run(const orc_grid *g, double kappa, double *u, int64_t nsteps, double dt,
                     double *t, const orc_face_bc *bcs);
/* The orientation pass of ModelC::initialize_grains (phasefield.rs:286-314):
 * grains as [cx, cy, radius, θ₀] quadruples, θ₀ into ghost-0 slabs. */
void orc_initialize_orientations(const orc_grid *g, double *theta, const double *grains,
                                 size_t ngrains);

/* Drivers (src/core/driver.rs:35-50) for the multi-driver explicit update. */
enum { ORC_DRIVER_TIME = 0, ORC_DRIVER_WIENER = 1, ORC_DRIVER_JUMP = 2 };
typedef struct {
  int32_t kind;               /* ORC_DRIVER_* */
  uint32_t index;             /* j of Wiener(j) / Jump(j) */
  const double *const *amp;   /* per field: dY/dt, the Wiener amplitude or the jump
                                 increment; NULL where the driver does not move it */
  const double *rate;         /* Jump: per-cell intensity in ghost-0 slabs */
} orc_driver;

/* State::apply_step_with (src/core/state.rs:646-699) over nfields field buffers
 * (ghost widths ghost[f]): per block, each driver in order through
 * Driver::apply_slab (src/core/driver.rs:115-182). A WienerJump<W, J> set lists
 * Time, Wiener(0..W), Jump(0..J). Returns -1 on an unknown driver kind or a jump
 * driver without an intensity. nthreads > 1 spreads blocks over OpenMP threads. */
int orc_apply_drivers(const orc_grid *g, size_t nfields, double *const *x, const uint32_t *ghost,
                      size_t ndrivers, const orc_driver *drivers, double dt, uint64_t seed,
                      uint64_t salt, int32_t nthreads);

/* The jump-diffusion model of tests/jump.rs:36-86 (WienerJump<1, 1>, ghost-0
 * fields on a CartesianGrid<1>) stepped by EulerMaruyama: dw = μ dt + σ dW; on a
 * fire count += 1 and jsum += payoff, at intensity lam (ghost-0 slabs). */
int orc_jump_diffusion_run(const orc_grid *g, const double *lam, double *w, double *count,
                           double *jsum, int64_t nsteps, double dt, double *t, uint64_t seed,
                           double mu, double sigma, double payoff, int32_t nthreads);

/* seed_positions recipe (examples/model_c.rs:175-206), multi-seed branch:
 * writes ngrains [cx, cy, radius] triples. */
void orc_seed_positions(uint64_t seed, size_t nseeds, double extent, double radius,
                        double *grains_out);

/* ---- AMR, 2-D (src/geometry/amr, src/integrators/subcycling.rs) ---- */
typedef struct orc_amr orc_amr;
typedef struct orc_amr_sim orc_amr_sim;

/* BergerOliger + GradientTagger knobs (src/geometry/amr/adapt.rs:55-129). */
typedef struct {
  int64_t every;
  int64_t buffer;
  double efficiency;
  int64_t min_width;
  double threshold;
  int32_t tag_field;
} orc_regrid_policy;

/* AmrGrid::from_patches (grid.rs:90-196). boxes: nboxes × [level, lo0, lo1, hi0, hi1],
 * levels >= 1 in level-(l) cell coordinates.  Returns 0, -1 (bad base) or the error
 * kind: 1 ratio, 2 empty patch, 3 not nested, 4 misaligned, 5 overlap. */
int orc_amr_create(const orc_grid *base, const uint32_t *ratios, size_t nratios,
                   const int64_t *boxes, size_t nboxes, orc_amr **out);
void orc_amr_destroy(orc_amr *g);
int64_t orc_amr_num_blocks(const orc_amr *g);
int32_t orc_amr_num_levels(const orc_amr *g);
void orc_amr_patch(const orc_amr *g, int64_t block, int64_t *out5);
int64_t orc_amr_block_offset(const orc_amr *g, int64_t block, uint32_t ghost);
int64_t orc_amr_field_len(const orc_amr *g, uint32_t ghost);
int64_t orc_amr_find_patch(const orc_amr *g, int32_t level, int64_t x, int64_t y);
void orc_amr_cell_center(const orc_amr *g, int64_t block, int64_t i, int64_t j, double *out2);
void orc_amr_restrict_level(const orc_amr *g, double *field, uint32_t ghost, int32_t fine_level);
void orc_amr_restrict(const orc_amr *g, double *field, uint32_t ghost);
/* fill_level (intergrid.rs:183-223) with optional time interpolation from `old`. */
void orc_amr_fill_level(const orc_amr *g, double *field, const double *old, double alpha,
                        uint32_t ghost, int32_t level);
/* fill_ghosts_mirror (intergrid.rs:155-168): every level, global dt. */
void orc_amr_fill_ghosts(const orc_amr *g, double *field, uint32_t ghost);
/* cluster (cluster.rs:144-155): flags n × [x, y] (reordered in place); writes up to
 * max_out boxes [lo0, lo1, hi0, hi1]; returns the number of boxes. */
size_t orc_cluster(int64_t *flags, size_t n, double efficiency, size_t min_width, int64_t *out,
                   size_t max_out);

/* Simulation on an AmrGrid: model 0 = Heat (κ∇²u, FE), 1 = Model C (EM);
 * integ 0 = global dt (fill = restrict + fill_ghosts_mirror), 1 = Subcycling;
 * policy NULL = static hierarchy.  Takes ownership of `grid`. */
int orc_amr_sim_create(orc_amr *grid, int32_t model, const orc_modelc_params *p, double kappa,
                       int32_t integ, uint64_t seed, int32_t wiener,
                       const orc_regrid_policy *policy, orc_amr_sim **out);
void orc_amr_sim_destroy(orc_amr_sim *s);
const orc_amr *orc_amr_sim_grid(const orc_amr_sim *s);
double *orc_amr_sim_field(orc_amr_sim *s, int32_t field);
double orc_amr_sim_time(const orc_amr_sim *s);
int orc_amr_sim_step(orc_amr_sim *s, double dt, int64_t nsteps);
/* A caller's TagCells (adapt.rs:32-46): tag_level(user, grid, level, sink) appends
 * the level-`level` cells needing refinement through orc_tag_sink_push (n × [x, y]);
 * prepare is the trait default (nothing).  NULL restores the GradientTagger. */
typedef struct orc_tag_sink orc_tag_sink;
typedef int (*orc_tag_level_fn)(void *user, const orc_amr *grid, int32_t level, orc_tag_sink *sink);
void orc_amr_sim_set_tagger(orc_amr_sim *s, orc_tag_level_fn fn, void *user);
int orc_tag_sink_push(orc_tag_sink *sink, const int64_t *cells, size_t n);
/* The nesting flags of one box of the level just built (adapt.rs:194-211), box4 =
 * [lo0, lo1, hi0, hi1] in its level's cells; out4 = the source-level box, unclipped. */
void orc_nesting_parents(const int64_t *box4, int64_t rf, int64_t rs, int64_t *out4);

/* ---- Model C derivatives (legendre_oracle_adjoint.cpp; no reference counterpart:
 *      the checkers for the device adjoint).  θ = (λ, D, ε₄); b = 0 only. ---- */
/* Tangent-linear run: advances (phi, u) exactly as orc_modelc_run (values bitwise) and
 * ndir (1..4) tangents at once: direction k has initial field tangents
 * dphi + k*len, du + k*len (len = field length) and parameter tangent
 * dtheta[3k .. 3k+2]; the field tangents are advanced in place.  Interiors only. */
int orc_modelc_tlm_run(const orc_grid *g, const orc_modelc_params *p, double *phi, double *u,
                       int32_t ndir, double *dphi, double *du, const double *dtheta,
                       int64_t nsteps, double dt);
/* Hand-written adjoint of one step at state (phi, u): (pphi, pu) = ∂L/∂(φ', u') on
 * entry, ∂L/∂(φ, u) on return; grad[3] += ∂L/∂θ of the step. */
int orc_modelc_adjoint_step(const orc_grid *g, const orc_modelc_params *p, const double *phi,
                            const double *u, double dt, double *pphi, double *pu, double *grad);
/* Reverse-mode gradient of L = ½‖φ_T − target‖² over nsteps (all states kept). */
int orc_modelc_gradient(const orc_grid *g, const orc_modelc_params *p, const double *phi0,
                        const double *u0, const double *target, int64_t nsteps, double dt,
                        double *loss, double *grad, double *gphi0, double *gu0);

#ifdef __cplusplus
}
#endif
#endif
