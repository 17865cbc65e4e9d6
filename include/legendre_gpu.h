/*
 * legendre_gpu — MI355X (gfx950) execution backend for legendre's per-block hot path.
 *
 * C ABI: plain pointers, sizes and status codes; no C++ or HIP types cross it.
 * Every entry point replaces one reference interface (cited file:line, relative to
 * the legendre crate root) and keeps its argument meaning and error behaviour: the
 * reference's invariant panics come back as LGD_ERR_* codes with a message from
 * lgd_ctx_last_error().
 *
 * Storage contract (src/core/state.rs:219-264, src/geometry/cartesian.rs:9-11):
 * a *field buffer* (lgd_buf) holds the ghost-inclusive slab of one field for every
 * block, block-major (block b at element offset b * lgd_grid_block_len(grid, ghost)),
 * each slab dimension-0 fastest.  Buffers are device-resident and zero-initialised;
 * the host sees them only through lgd_buf_upload / lgd_buf_download.
 *
 * Threading: one context is driven by one host thread.  Calls are stream-ordered
 * and asynchronous until lgd_ctx_sync() or a download.
 */
#ifndef LEGENDRE_GPU_H
#define LEGENDRE_GPU_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LGD_ABI_VERSION 1u

typedef enum {
  LGD_OK = 0,
  LGD_ERR_INVALID_ARGUMENT = 1,
  /* ghost width exceeds a block extent (src/geometry/cartesian.rs:456-462) */
  LGD_ERR_GHOST_TOO_WIDE = 2,
  /* buffers not congruent with the grid / each other (src/core/state.rs:550, 560) */
  LGD_ERR_SIZE_MISMATCH = 3,
  /* stochastic driver declared but no field driven (src/core/simulation.rs:149-160) */
  LGD_ERR_UNDRIVEN_DRIVER = 4,
  /* configuration the backend does not implement (message names it) */
  LGD_ERR_UNSUPPORTED = 5,
  LGD_ERR_OUT_OF_MEMORY = 6,
  LGD_ERR_HIP = 7,
  /* a file could not be written (the reference's ObserverError::Io, src/io/mod.rs) */
  LGD_ERR_IO = 8
} lgd_status;

typedef struct lgd_ctx lgd_ctx;
typedef struct lgd_buf lgd_buf;
typedef struct lgd_grid lgd_grid;
typedef struct lgd_amr lgd_amr;
typedef struct lgd_sim lgd_sim;

/* CartesianGrid<D>::new(cells, block_cells, origin, spacing).with_periodic(periodic)
 * (src/geometry/cartesian.rs:48-97); 1 <= dim <= 3, entries >= dim ignored. */
typedef struct {
  int32_t dim;
  int32_t periodic[3];
  int64_t cells[3];
  int64_t block_cells[3];
  double origin[3];
  double spacing[3];
} lgd_cartesian_desc;

/* FaceBc (src/geometry/boundary.rs:36-54). */
typedef enum {
  LGD_BC_MIRROR = 0,
  LGD_BC_DIRICHLET = 1,
  LGD_BC_FLUX = 2,
  LGD_BC_EXTRAPOLATE = 3
} lgd_bc_kind;

typedef struct {
  int32_t kind; /* lgd_bc_kind */
  double value; /* Dirichlet value / outward normal flux; unused otherwise */
} lgd_face_bc;

/* ModelC public fields (src/physics/phasefield.rs:55-67) plus the finite-volume
 * policy's regularisation threshold FiniteVolume::tol (finite_volume.rs:41-44). */
typedef struct {
  double eps4;
  double lambda;
  double d_thermal;
  double noise_amplitude;
  double anisotropy_tol;
  double fv_tol;
} lgd_modelc_params;

uint32_t lgd_abi_version(void);

/* ---------------------------------------------------------------- context */
/* stream: an existing hipStream_t to launch on, or NULL for a library-owned one. */
int lgd_ctx_create(int32_t device, void *stream, lgd_ctx **out);
/* Every buffer, grid, simulation, ensemble, adjoint and calibration created on the
 * context must be destroyed before it (they synchronise on its stream). */
int lgd_ctx_destroy(lgd_ctx *ctx);
int lgd_ctx_sync(lgd_ctx *ctx);
void *lgd_ctx_stream(lgd_ctx *ctx);
const char *lgd_ctx_last_error(const lgd_ctx *ctx);
/* Launch-shape knobs for this context (no reference counterpart: the reference has
 * no launch shapes).  Keys: "modelc_kernel" (0 auto, 1 per-cell, 2 streaming, 3
 * streaming with two x-neighbouring blocks per wavefront where the blocks are at most
 * 64 cells wide, both even in number per block row and in width, else as 2),
 * "modelc_strips" (row strips per block for the streaming kernel, 0 = default) and
 * "modelc_prefetch" (rows the streaming kernel loads ahead: 1 or 2, 0 = default) and
 * "amr_per_block" (1: the AMR level kernels launch once per patch instead of once per
 * level — the per-block dispatch baseline), "amr_kernel" (the AMR Model C level
 * stage: 0 or 2 streaming — one wavefront per strip of rows of up to 62 columns of a
 * patch —, 1 one thread per cell in tiles), "amr_rows" (rows per strip of the
 * streaming stage, 0 = spread over the device's SIMDs), "amr_tile" and "amr_wg" (the
 * AMR launch shapes lgd_sim_autotune_amr chooses; 0 = 64×4 tiles and 256-wide fills) and
 * "amr_split_fill" (1: every fill and every restriction is a launch of its own; 0,
 * the default: a substep's stage launch on the finest level also fills the ghosts
 * the next substep reads, and on a level with a finer one the restriction ending a
 * substep shares a launch with the next substep's fill).
 * Results are bitwise independent of these knobs.  The adjoint has two more:
 * "adjoint_kernel" (0: the streaming reverse step wherever the grid allows, else the
 * three-pass form; 1: always the three passes; 2: streaming, error if the grid does
 * not allow it) and "adjoint_rows" (rows per strip of the streaming reverse step,
 * 0 = balanced over the device's SIMDs).  The two reverse-step forms sum in
 * different orders and agree to rounding; for a fixed setting every result is
 * bitwise reproducible. */
int lgd_ctx_set_tuning(lgd_ctx *ctx, const char *key, int64_t value);
/* Autotuner (no reference counterpart: the reference has no launch shapes).  Times
 * every launch shape of lgd_modelc_step — one thread per cell, or streaming with
 * 1..32 strips per block and one or two rows loaded ahead, one block per wavefront
 * or (where the grid allows) two — on (phi, u) over this
 * grid (outputs to scratch), `reps` launches each, and keeps the fastest for this
 * (block shape, block count, noise) in the context; lgd_modelc_step and the ensemble
 * step use it whenever no knob is set.  kernel: 1 per-cell, 2 streaming, 3 streaming
 * with block pairs. */
typedef struct {
  int32_t kernel;
  int32_t strips;
  int32_t prefetch;
  int32_t candidates;
  double ms;          /* the chosen shape's mean launch time */
  double default_ms;  /* the untuned default's (streaming, 4 strips, two rows ahead,
                         one with noise) */
} lgd_tuning_result;
int lgd_autotune_modelc_step(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                             const lgd_buf *phi, const lgd_buf *u, double dt, int32_t wiener,
                             int32_t reps, lgd_tuning_result *out);

/* The device's streaming rate, for rooflines (no reference counterpart: the reference
 * launches no kernels): y = x over len f64 (equal, even lengths) with 16-byte
 * nontemporal loads and stores, vec (1, 2, 4 or 8) vectors in flight per lane. */
int lgd_stream_copy(lgd_ctx *ctx, const lgd_buf *x, lgd_buf *y, int32_t vec);

/* ------------------------------------- Allocator / StorageBackend (storage.rs) */
/* Allocator::allocate(len): zero-initialised storage of len f64
 * (src/core/storage.rs:148-154, DenseStorage 158-189). */
int lgd_buf_alloc(lgd_ctx *ctx, size_t len, lgd_buf **out);
int lgd_buf_free(lgd_buf *buf);
size_t lgd_buf_len(const lgd_buf *buf);
/* Device address of element 0 (for callers that launch their own work). */
void *lgd_buf_device_ptr(lgd_buf *buf);
/* StorageBackend::as_slice / as_mut_slice served by explicit copies
 * (src/core/storage.rs:126-138). */
int lgd_buf_upload(lgd_ctx *ctx, lgd_buf *buf, size_t offset, const double *host, size_t n);
int lgd_buf_download(lgd_ctx *ctx, const lgd_buf *buf, size_t offset, double *host, size_t n);
/* State::fill_zero / copy_from over one field buffer (src/core/state.rs:571-597). */
int lgd_buf_fill_zero(lgd_ctx *ctx, lgd_buf *buf);
int lgd_buf_copy(lgd_ctx *ctx, lgd_buf *dst, const lgd_buf *src);

/* ------------------------------------------------ Grid → device block table */
int lgd_grid_create_cartesian(lgd_ctx *ctx, const lgd_cartesian_desc *desc, lgd_grid **out);
int lgd_grid_destroy(lgd_grid *grid);
/* Grid::num_blocks / block_len (src/geometry/cartesian.rs:568-574). */
int64_t lgd_grid_num_blocks(const lgd_grid *grid);
int64_t lgd_grid_block_len(const lgd_grid *grid, uint32_t ghost);
/* CartesianGrid::face_neighbor (cartesian.rs:162-181); returns -1 for None. */
int64_t lgd_grid_face_neighbor(const lgd_grid *grid, int64_t block, int32_t dim, int32_t dir);
/* Grid::cell_key (cartesian.rs:599-619); returns -1 for a ghost entry. */
int64_t lgd_grid_cell_key(const lgd_grid *grid, uint32_t ghost, int64_t offset);

/* ------------------------------------------------------ per-intent kernels */
/* fill_ghosts_bc (cartesian.rs:377-387, sweep 439-526): halo exchange, periodic
 * wrap, then the per-face rule at physical faces.  bcs holds 2*dim entries
 * indexed [d*2 + (side > 0)]; NULL means Mirror on every face
 * (fill_ghosts_mirror, cartesian.rs:392-401). */
int lgd_fill_ghosts(lgd_ctx *ctx, const lgd_grid *grid, lgd_buf *field, uint32_t ghost,
                    const lgd_face_bc *bcs);
/* CentralLaplacian::apply (src/discretization/finite_difference.rs:49-83):
 * interior of out = ∇²in; both buffers ghost width `ghost`. */
int lgd_central_laplacian(lgd_ctx *ctx, const lgd_grid *grid, const lgd_buf *in, lgd_buf *out,
                          uint32_t ghost);
/* KarmaRappelFlux::apply (src/discretization/finite_volume.rs:248-301), ghost 1, 2-D. */
int lgd_karma_rappel_flux(lgd_ctx *ctx, const lgd_grid *grid, const lgd_buf *phi, lgd_buf *out,
                          double eps4, double tol);
/* ModelC::vector_field_block for Driver::Time (src/physics/phasefield.rs:367-421):
 * kphi, ku interiors from ghost-filled phi, u. */
int lgd_modelc_rhs(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                   const lgd_buf *phi, const lgd_buf *u, lgd_buf *kphi, lgd_buf *ku);
/* The multi-grain vector field (KarmaRappelFlux::apply_oriented, finite_volume.rs:188-235;
 * phasefield.rs:378-392): cos0, sin0 = cos θ₀, sin θ₀ per interior cell in ghost-0
 * slabs (the caller takes them with libm). */
int lgd_modelc_rhs_oriented(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                            const lgd_buf *phi, const lgd_buf *u, const lgd_buf *cos0,
                            const lgd_buf *sin0, lgd_buf *kphi, lgd_buf *ku);
/* eval_tendency for Driver::Wiener(0) (src/integrators/mod.rs:242-279 with
 * phasefield.rs:361-366): zero the slab, interior = noise_amplitude. */
int lgd_modelc_amplitude(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                         lgd_buf *amp_phi);
/* Heat3 / Heat<D>::vector_field_block (examples/heat3d.rs:63-79): k = κ·∇²u, ghost 1. */
int lgd_heat_rhs(lgd_ctx *ctx, const lgd_grid *grid, double kappa, const lgd_buf *u, lgd_buf *k);
/* block_axpy over every slab entry: x += alpha*y (src/core/state.rs:485-498, 549-568);
 * also Driver::Time's kernel x += dt*k (src/core/driver.rs:131-138). */
int lgd_axpy(lgd_ctx *ctx, lgd_buf *x, const lgd_buf *y, double alpha);
/* Driver::Wiener arm of apply_slab (src/core/driver.rs:139-158):
 * x += (√dt·amp)·N(splitmix64(mix_key(seed,[salt,stream,block]) ^ cell)) on interior
 * entries with amp != 0.  stream = Driver::stream() (driver.rs:55-61). */
int lgd_apply_wiener(lgd_ctx *ctx, const lgd_grid *grid, lgd_buf *x, const lgd_buf *amp,
                     uint32_t ghost, double dt, uint64_t seed, uint64_t salt, uint64_t stream);

/* Driver (src/core/driver.rs:35-50): what multiplies a vector field. */
typedef enum { LGD_DRIVER_TIME = 0, LGD_DRIVER_WIENER = 1, LGD_DRIVER_JUMP = 2 } lgd_driver_kind;

/* One driver of an explicit update: its kind and index j (Wiener(j) / Jump(j)), and
 * per field the buffer the model evaluated for it — dY/dt for Time, the amplitude
 * for Wiener, the increment ΔJ for Jump — congruent with that field, or NULL (or a
 * zero-length buffer) where the driver does not move the field (the reference's
 * zero-length slabs of State::like_for, src/core/state.rs:317-337).  A jump driver
 * also carries its per-cell intensity λ in ghost-0 slabs (BlockStateMut::
 * intensity_mut, state.rs:806-832). */
typedef struct lgd_driver {
  int32_t kind;              /* lgd_driver_kind */
  uint32_t index;            /* j */
  const lgd_buf *const *amp; /* nfields entries */
  const lgd_buf *rate;       /* Jump only */
} lgd_driver;

/* State::apply_step_with (src/core/state.rs:646-699) for any driver set: x[f] (ghost
 * width ghost[f]) takes every driver in the order given, each through
 * Driver::apply_slab (driver.rs:115-182): Time x += dt·v on every entry; Wiener
 * x += (√dt·v)·ξ on interior entries with v != 0; Jump x += ΔJ on interior cells
 * that fire, each firing with probability 1 − e^{−λ·dt} against the intensity.  A
 * stochastic driver's per-cell draw is keyed by (seed, salt, Driver::stream(),
 * block, cell) and shared by every field it moves.  A WienerJump<W, J> set lists
 * Time, Wiener(0..W), Jump(0..J) (driver.rs:239-252); EulerMaruyama passes
 * salt = t.to_bits().  At most 8 fields and 8 drivers (LGD_ERR_UNSUPPORTED
 * beyond); a jump driver without an intensity is LGD_ERR_INVALID_ARGUMENT (the
 * reference's panic at driver.rs:160).  Jump fires use the device expm1. */
int lgd_apply_drivers(lgd_ctx *ctx, const lgd_grid *grid, size_t nfields, lgd_buf *const *x,
                      const uint32_t *ghost, size_t ndrivers, const lgd_driver *drivers, double dt,
                      uint64_t seed, uint64_t salt);

/* ------------------------------------------------------ fused hot path */
/* One Euler–Maruyama step of Model C (src/integrators/euler_maruyama.rs:34-75 on
 * ModelC<NoNoise | Wiener<1>>, mirror ghosts, src/physics/phasefield.rs:431-439) in one
 * kernel: halos are read from neighbour interiors (no ghost pass), the RHS and the
 * stage update are fused, and the new interiors are written to phi_next/u_next
 * (ghost entries of *_next are unspecified afterwards, as the reference's are until the
 * next ghost fill, src/core/driver.rs:128-130; fill_ghosts restores them if needed).
 * salt = t.to_bits() of the step's start time; wiener = 0 for NoNoise.
 * Interiors equal the reference step's bit for bit (noise: up to the device log/cos). */
int lgd_modelc_step(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                    const lgd_buf *phi, const lgd_buf *u, lgd_buf *phi_next, lgd_buf *u_next,
                    double dt, uint64_t seed, uint64_t salt, int32_t wiener);
/* One forward-Euler step of Heat<D> (src/integrators/euler.rs:104-130,
 * examples/heat3d.rs:54-79) with mirror/periodic ghosts, fused the same way. */
int lgd_heat_step(lgd_ctx *ctx, const lgd_grid *grid, double kappa, const lgd_buf *u,
                  lgd_buf *u_next, double dt);
/* The same step with a FaceBc per physical face (bcs as in lgd_fill_ghosts, NULL =
 * Mirror everywhere): Dirichlet, Flux and Extrapolate halos are evaluated in the
 * kernel from the edge cells (FaceBc::ghost_value, src/geometry/boundary.rs:63-73),
 * bitwise what fill_ghosts would have written at the face-neighbour ghosts the
 * stencil reads. */
int lgd_heat_step_bc(lgd_ctx *ctx, const lgd_grid *grid, double kappa, const lgd_face_bc *bcs,
                     const lgd_buf *u, lgd_buf *u_next, double dt);

/* ------------------------------------- AMR hierarchy (src/geometry/amr), 2-D */
/* AmrGrid::from_patches(base, ratios, levels) (src/geometry/amr/grid.rs:90-196).
 * boxes: nboxes × [level, lo0, lo1, hi0, hi1] in level cell coordinates, level >= 1
 * (level 0 is the base grid's block tiling).  Invalid hierarchies (ratio < 2, empty,
 * overlapping, misaligned or improperly nested patches) return
 * LGD_ERR_INVALID_ARGUMENT with the reason.  Field buffers on the hierarchy hold
 * ghost-1 slabs of every patch, patch order (level-major), lgd_amr_field_len long. */
int lgd_amr_create(lgd_ctx *ctx, const lgd_cartesian_desc *base, const uint32_t *ratios,
                   size_t nratios, const int64_t *boxes, size_t nboxes, lgd_amr **out);
int lgd_amr_destroy(lgd_amr *amr);
int64_t lgd_amr_num_blocks(const lgd_amr *amr);
int32_t lgd_amr_num_levels(const lgd_amr *amr);
/* AmrGrid::patch: [level, lo0, lo1, hi0, hi1] of block `block` (grid.rs:237-240). */
int lgd_amr_patch(const lgd_amr *amr, int64_t block, int64_t *out5);
int64_t lgd_amr_field_len(const lgd_amr *amr);
int64_t lgd_amr_block_offset(const lgd_amr *amr, int64_t block);
/* AmrGrid::find_patch (grid.rs:264-278); -1 for None. */
int64_t lgd_amr_find_patch(const lgd_amr *amr, int32_t level, int64_t x, int64_t y);
/* restrict_level (src/geometry/amr/intergrid.rs:110-150). */
int lgd_amr_restrict_level(lgd_ctx *ctx, const lgd_amr *amr, lgd_buf *field, int32_t fine_level);
/* fill_level (intergrid.rs:183-223): same-level exchange, physical mirror and bilinear
 * coarse→fine prolongation into every ghost of one level; old != NULL interpolates the
 * coarse source in time: old·(1−alpha) + field·alpha. */
int lgd_amr_fill_level(lgd_ctx *ctx, const lgd_amr *amr, lgd_buf *field, const lgd_buf *old,
                       double alpha, int32_t level);

/* ------------------------------------------- batched ensemble of Model C runs */
/* An ensemble of independent 2-D Model C realizations of uneven sizes, each its own
 * Simulation (own BlockIds from 0, own EulerMaruyama seed, mirror faces), sharing
 * one block shape and spacing and advanced together by ONE launch per step over a
 * device table of every block of every realization.  No reference counterpart as a
 * batch: each realization is exactly Simulation::new(CartesianGrid(cells,
 * block_cells), FiniteVolume, ModelC{eps4, lambda, d_thermal, b}, EulerMaruyama{seed}). */
typedef struct {
  int64_t cells[2];
  double eps4;
  double lambda;
  double d_thermal;
  double noise_amplitude; /* b; ModelC<Wiener<1>> when > 0 */
  uint64_t seed;
} lgd_realization;

typedef struct lgd_ensemble lgd_ensemble;

int lgd_ensemble_create(lgd_ctx *ctx, const lgd_realization *r, size_t n,
                        const int64_t *block_cells2, const double *spacing2, double anisotropy_tol,
                        double fv_tol, lgd_ensemble **out);
int lgd_ensemble_destroy(lgd_ensemble *e);
int64_t lgd_ensemble_num_blocks(const lgd_ensemble *e);
/* First block of realization r in the ensemble's field buffers (block-major). */
int64_t lgd_ensemble_block_start(const lgd_ensemble *e, int64_t r);
/* Current state buffer of field 0 (phi) or 1 (u) (valid until the next step). */
int lgd_ensemble_field(lgd_ensemble *e, int32_t field, lgd_buf **out);
/* ModelC::initialize_grains on realization r (host, glibc tanh/hypot). */
int lgd_ensemble_initialize_grains(lgd_ensemble *e, int64_t r, const double *grains, size_t n,
                                   double u0);
/* nsteps × Simulation::step(dt) of every realization; per_block = 1 issues one launch
 * per block instead of one per step (the unbatched dispatch, for comparison). */
int lgd_ensemble_step(lgd_ensemble *e, double dt, int64_t nsteps, int32_t per_block);
double lgd_ensemble_time(const lgd_ensemble *e);
/* Fraction of interior cells with phi > 0, per realization (n doubles). */
int lgd_ensemble_solid_fraction(lgd_ensemble *e, double *out);

/* BergerOliger policy with a GradientTagger (src/geometry/amr/adapt.rs:55-129). */
typedef struct {
  int64_t every;      /* regrid every this many steps (step 0 always) */
  int64_t buffer;     /* flag dilation in cells */
  double efficiency;  /* ClusterParams::efficiency */
  int64_t min_width;  /* ClusterParams::min_width */
  double threshold;   /* GradientTagger::threshold */
  int32_t tag_field;  /* field index the tagger reads */
} lgd_regrid_policy;

/* ---------------------------------------- Simulation (src/core/simulation.rs) */
/* mode: how a step is executed. */
typedef enum {
  /* the reference's call sequence, one kernel per intent: fill_ghosts → RHS →
   * amplitude → apply_step (src/integrators/euler_maruyama.rs:57-74) */
  LGD_MODE_PER_INTENT = 0,
  /* one fused kernel per step over ping-pong state buffers */
  LGD_MODE_FUSED = 1
} lgd_mode;

/* Integrator of a simulation: EulerMaruyama / ForwardEuler (the default) or
 * RungeKutta4 (src/integrators/rk4.rs:30-80; deterministic models only). */
typedef enum { LGD_INTEGRATOR_EULER = 0, LGD_INTEGRATOR_RK4 = 1 } lgd_integrator;

/* Simulation::new(grid, FiniteVolume{fv_tol}, ModelC, EulerMaruyama{seed}, GpuScheduler,
 * DeviceAllocator) (simulation.rs:108-183): registers phi (field 0, Time + Wiener(0)) and
 * u (field 1, Time), ghost 1, allocates state and stage buffers.  wiener = 0 selects
 * ModelC<NoNoise>, 1 ModelC<Wiener<1>>.  Requires a 2-D grid. */
int lgd_sim_create_modelc(lgd_ctx *ctx, const lgd_grid *grid, const lgd_modelc_params *params,
                          int32_t wiener, uint64_t seed, int32_t mode, lgd_sim **out);
/* Simulation::new(grid, FiniteDifference, Heat<D>{kappa}, ForwardEuler, ...): field 0 = u,
 * ghost 1; bcs as in lgd_fill_ghosts (every FaceBc in both modes). */
int lgd_sim_create_heat(lgd_ctx *ctx, const lgd_grid *grid, double kappa,
                        const lgd_face_bc *bcs, int32_t mode, lgd_sim **out);
/* Simulation::adaptive on an AmrGrid (simulation.rs:137-183): model 0 = Heat (κ∇²u,
 * stable dt 0.25·h²/κ), 1 = Model C; subcycle = 0 runs EulerMaruyama with a global dt
 * (fill_ghosts = restrict + fill_ghosts_mirror every field, phasefield.rs:452-467),
 * 1 runs Subcycling (src/integrators/subcycling.rs:75-267).  policy NULL keeps the
 * hierarchy static.  The simulation takes ownership of `amr`; after a regrid the
 * current hierarchy is lgd_sim_amr(). */
int lgd_sim_create_amr(lgd_ctx *ctx, lgd_amr *amr, int32_t model, const lgd_modelc_params *params,
                       double kappa, int32_t wiener, uint64_t seed, int32_t subcycle,
                       const lgd_regrid_policy *policy, lgd_sim **out);
const lgd_amr *lgd_sim_amr(const lgd_sim *sim);
int lgd_sim_destroy(lgd_sim *sim);
/* The current state buffer of field `index` (valid until the next lgd_sim_step).  Its
 * interiors are the state; on an AmrGrid (as after the fused Cartesian step) ghost
 * entries are scratch: every step fills the ghosts it reads. */
int lgd_sim_field(lgd_sim *sim, int32_t index, lgd_buf **out);
/* nsteps × Simulation::step(dt) (simulation.rs:255-290): t advances by t += dt on
 * the host in f64, exactly as the reference accumulates it (the noise salt). */
int lgd_sim_step(lgd_sim *sim, double dt, int64_t nsteps);
/* Whole-step HIP graphs (no reference counterpart: the reference launches no
 * kernels; the caller is any driver of lgd_sim_step).  steps > 0: lgd_sim_step runs
 * `steps` whole steps at a time as one captured graph — every launch of each step,
 * the subcycled level recursion and the time accumulation — captured on first use
 * and replayed while the buffers, dt, AMR hierarchy and tuning stay the same.  The
 * noise salt comes from a device copy of t advanced inside the graph with the same
 * f64 arithmetic, so results are bitwise those of steps = 0.  AMR regrids stay
 * host-driven at the top of their step and no graph spans one.  The context's
 * stream must not be the null stream.  0 turns graphs off (and frees them). */
int lgd_sim_set_graph_steps(lgd_sim *sim, int64_t steps);
double lgd_sim_time(const lgd_sim *sim);
/* Regrid accounting (no reference counterpart; for benchmarks): regrids run so far
 * and their wall seconds by phase — [0] the tagged field prepared and brought to the
 * host, [1] the proposal (tagging, buffer, nesting, clustering, hysteresis), [2] the
 * new hierarchy (patches, plans, uploads), [3] the state migrated. */
int lgd_sim_regrid_stats(const lgd_sim *sim, int64_t *regrids, double *seconds);
int64_t lgd_sim_step_index(const lgd_sim *sim);
/* Autotuner for an AMR simulation's launches (no reference counterpart: the
 * reference has no launch shapes).  Times, on the simulation's current hierarchy and
 * state, the Model C level stage in its streaming form at 2 to 16 rows per strip
 * (the "amr_rows" knob; 0 = the level's own height) and in its per-cell form at every
 * tile shape ("amr_kernel" 1 with "amr_tile": 64×4, 32×8, 16×16, 128×2, 64×2, 32×4,
 * 256×1 threads), every fill / restriction workgroup width ("amr_wg": 64 to 1024)
 * and, for level 0 on the streaming kernel, one block or two per wavefront, 1, 2, 4
 * or 8 strips per block with one or two rows loaded ahead, `reps` launches each,
 * weighted by each level's
 * substeps per coarse step of `dt`; keeps the fastest in the context.  The timed
 * launches write only entries the next step overwrites before reading them (the
 * other buffer of a level, ghost entries), so the simulation's results are bitwise
 * those it would have had untuned; results are bitwise independent of every shape.
 * ms / default_ms: the kernel time per coarse step the timings predict, tuned and
 * untuned (64×4 tiles, 256-wide fills, the untuned level-0 shape). */
typedef struct {
  int32_t tile_x, tile_y; /* the per-cell stage's tile, 0 when the streaming stage won */
  int32_t wg;
  int32_t level0_strips, level0_prefetch; /* 0 when level 0 is not on the streaming kernel */
  int32_t candidates;
  int32_t stage_rows;    /* rows per strip of the streaming stage (finest level), 0 if per-cell */
  int32_t level0_kernel; /* 2: one block per wavefront, 3: block pairs, 0: not streamed */
  double ms;
  double default_ms;
} lgd_amr_tuning_result;
int lgd_sim_autotune_amr(lgd_sim *sim, double dt, int32_t reps, lgd_amr_tuning_result *out);

/* TagCells (src/geometry/amr/adapt.rs:32-46) as a host callback, in place of the
 * policy's GradientTagger: at each regrid BergerOliger::regrid calls tag_level for
 * every existing source level, finest first, and the callback appends the cells of
 * `level` needing refinement (level cell coordinates, n × [x, y]) through
 * lgd_tag_sink_push; the state is readable through lgd_sim_field.  prepare is the
 * trait's default (nothing), so the tag field is not restricted or filled.  Buffer
 * dilation, nesting flags, clustering, hysteresis and migration follow the policy as
 * for the GradientTagger.  A nonzero return fails the step with
 * LGD_ERR_INVALID_ARGUMENT.  fn = NULL restores the GradientTagger. */
typedef struct lgd_tag_sink lgd_tag_sink;
typedef int (*lgd_tag_level_fn)(void *user, const lgd_amr *grid, int32_t level,
                                lgd_tag_sink *sink);
int lgd_sim_set_tagger(lgd_sim *sim, lgd_tag_level_fn fn, void *user);
int lgd_tag_sink_push(lgd_tag_sink *sink, const int64_t *cells, size_t ncells);
/* ModelC::initialize_grains with orientation 0 (phasefield.rs:227-286), computed on the
 * host (glibc tanh/hypot, as the reference) and uploaded: grains = [cx, cy, r] × n. */
int lgd_sim_initialize_grains(lgd_sim *sim, const double *grains, size_t ngrains, double u0);
/* Simulation::new with a RungeKutta4 integrator instead (rk4.rs:23-80): Cartesian
 * Model C without a Wiener driver, or heat; stages k1..k4 and y_tmp are allocated
 * here.  Bitwise the reference's operation order (whole-slab copy and axpy sweeps). */
int lgd_sim_set_integrator(lgd_sim *sim, int32_t integrator);
/* ModelC::with_orientations (phasefield.rs:118-125): θ₀ per interior cell in ghost-0
 * slabs (nblocks·nx·ny values, NULL = all zero).  cos θ₀ and sin θ₀ are taken once
 * here with libm, as the reference takes them per cell and step; the multi-grain
 * vector field (finite_volume.rs:188-235, phasefield.rs:378-392) then runs in every
 * step.  Cartesian Model C only. */
int lgd_sim_set_orientation(lgd_sim *sim, const double *theta0, size_t n);
/* ModelC::initialize_grains with orientations (phasefield.rs:227-314): grains =
 * [cx, cy, r, θ₀] × n; φ and u as lgd_sim_initialize_grains, θ₀ by nearest-seed
 * Voronoi, then lgd_sim_set_orientation. */
int lgd_sim_initialize_grains_oriented(lgd_sim *sim, const double *grains, size_t ngrains,
                                       double u0);

/* ------------------------- observation (src/core/observer.rs, src/io/parquet.rs) */
typedef struct lgd_observer lgd_observer;

/* An observed field: its Parquet column name and, for lgd_sim_attach_observer, the
 * simulation field it reads (observed fields: 0 = phi, 1 = u of Model C, 0 = u of
 * heat; static fields: 0 = theta0 of a multi-grain Model C). */
typedef struct {
  const char *name;
  int32_t field;
} lgd_observed_field;

/* ParquetObserver::new(grid, fields, every, dir).with_static(statics) fed by an
 * AsyncObserver whose snapshot ring holds `ring` >= 1 buffers (src/io/parquet.rs:63-130,
 * src/core/observer.rs:75-135), plus one buffer staging the static file of a grid
 * epoch.  Each buffer is device staging plus pinned host memory, sized on
 * lgd_sim_attach_observer (or on first use) and grown only when an AMR epoch has more
 * rows.  Snapshots fire on step 1 and every `every` steps.  Files:
 * dir/static_<epoch:07>.parquet (x, y[, z], level, patch, static fields) once per
 * grid epoch and dir/snap_<step:07>.parquet (step, t, epoch, fields), rows block-major
 * over interior cells with dimension 0 fastest, row groups of whole blocks cut at
 * 2^22 rows, snappy pages (parquet.rs:1-35, 134-311).  Creates `dir`.  An empty field
 * list is LGD_ERR_INVALID_ARGUMENT (parquet.rs:94-96). */
int lgd_observer_create(lgd_ctx *ctx, const char *dir, uint64_t every, size_t ring,
                        const lgd_observed_field *fields, size_t nfields,
                        const lgd_observed_field *statics, size_t nstatics, lgd_observer **out);
/* Writes every queued snapshot, stops the writer thread and frees the ring. */
int lgd_observer_destroy(lgd_observer *obs);
/* Observer::observe on a Cartesian grid (observer.rs:41-45, 156-176): `fields` and
 * `statics` are device buffers in declaration order with their ghost widths.  At a
 * snapshot step the interiors are gathered on the context's stream and copied to a
 * pinned ring buffer on a copy stream; the call never waits for the disk.  With every
 * ring buffer in flight the snapshot is dropped and counted.  `epoch` names the grid:
 * a new epoch (or grid) writes a new static file before its first snapshot. */
int lgd_observer_observe(lgd_observer *obs, const lgd_grid *grid, uint64_t step, double t,
                         uint64_t epoch, const lgd_buf *const *fields, const uint32_t *field_ghost,
                         const lgd_buf *const *statics, const uint32_t *static_ghost);
/* The same on an AMR hierarchy: ghost-1 fields in patch order; the static file holds
 * the AmrGrid cell centres, each cell's level and patch (parquet.rs:216-235). */
int lgd_observer_observe_amr(lgd_observer *obs, const lgd_amr *amr, uint64_t step, double t,
                             uint64_t epoch, const lgd_buf *const *fields);
/* Blocks until every queued snapshot is on disk (resumes a paused writer). */
int lgd_observer_flush(lgd_observer *obs);
/* Blocking mode (1): with every ring buffer in flight observe() waits for the writer
 * instead of dropping the snapshot — the synchronous ParquetObserver's guarantee that
 * every snapshot step is written (parquet.rs:318-333).  Default 0 (AsyncObserver). */
int lgd_observer_set_blocking(lgd_observer *obs, int32_t blocking);
/* Holds (1) or releases (0) the writer thread: held, snapshots queue until the ring
 * is full and are then dropped (also in blocking mode). */
int lgd_observer_pause(lgd_observer *obs, int32_t paused);
/* Files written, snapshots dropped, failed writes, seconds spent writing. */
int lgd_observer_stats(lgd_observer *obs, uint64_t *written, uint64_t *dropped, uint64_t *failed,
                       double *write_seconds);
/* The last failed write's message (empty if none), truncated to len bytes. */
int lgd_observer_last_error(lgd_observer *obs, char *buf, size_t len);
/* Simulation::attach_observer (simulation.rs:204-208): after every completed step the
 * simulation calls the observer with its step index, time, grid epoch (0, +1 per
 * regrid that changes the hierarchy) and fields.  Graph-replayed runs end a graph at
 * each snapshot step.  The observer must outlive the simulation's stepping. */
int lgd_sim_attach_observer(lgd_sim *sim, lgd_observer *obs);
/* Stops notifying `obs` (before destroying an observer the simulation outlives). */
int lgd_sim_detach_observer(lgd_sim *sim, lgd_observer *obs);

/* Writes host columns as one Parquet file in the snapshot layout: types[c] 0 = DOUBLE,
 * 1 = INT64; columns[c] holds nrows values, or is NULL with constants[c] holding the
 * bits every row shares; row groups of group_rows[0..ngroups) rows (summing to
 * nrows).  LGD_ERR_IO when the file cannot be written (message in err). */
int lgd_parquet_write_file(const char *path, size_t ncols, const char *const *names,
                           const int32_t *types, const void *const *columns,
                           const uint64_t *constants, int64_t nrows, const int64_t *group_rows,
                           size_t ngroups, char *err, size_t errlen);

/* -------------------------------------------------------------- Model C adjoint */
/* The reverse-mode derivative of the explicit Model C step (Euler–Maruyama with
 * b = 0, which is forward Euler: src/integrators/euler_maruyama.rs:34-75,
 * tests/convergence.rs:100-107) for calibrating θ = (λ, D, ε₄) and the initial
 * fields.  No counterpart in the reference, which has no derivative code; the
 * callers are a forward/backward trait in the autograd-Function role on the host
 * side and the Python ModelCCalibration. */
typedef struct lgd_adjoint lgd_adjoint;
/* Workspace for reverse steps on `grid` (2-D; mirror or periodic faces). */
int lgd_adjoint_create(lgd_ctx *ctx, const lgd_grid *grid, lgd_adjoint **out);
int lgd_adjoint_destroy(lgd_adjoint *adj);
/* One reverse step at the state (phi, u) the forward step started from: on entry
 * (pphi, pu) hold dL/d(phi', u') on the interiors, on return dL/d(phi, u); the step's
 * dL/dθ is added to the workspace's accumulator. */
int lgd_modelc_adjoint_step(lgd_adjoint *adj, const lgd_modelc_params *params,
                            const lgd_buf *phi, const lgd_buf *u, double dt, lgd_buf *pphi,
                            lgd_buf *pu);
/* Zero the accumulator / read it: grad = (dL/dλ, dL/dD, dL/dε₄), reduced in a fixed
 * order (bitwise reproducible). */
int lgd_adjoint_reset(lgd_adjoint *adj);
int lgd_adjoint_grad(lgd_adjoint *adj, double *grad);

/* The reverse step without allocation.  No counterpart in the reference: for callers
 * that cannot allocate device memory — code under HIP graph capture, tight
 * calibration loops, harnesses that forbid hipMalloc.  lgd_modelc_adjoint_step on the
 * streaming reverse step with no lgd_adjoint: its workspace (the flattened adjoint in
 * and out and the per-wave θ partials) is a static device array of the library, one
 * per GPU, held until the process ends (about 34 MB).  (pphi, pu) as for
 * lgd_modelc_adjoint_step, bitwise the same; grad (>= 3 f64 on the device) receives
 * this step's dL/dθ = (dL/dλ, dL/dD, dL/dε₄), bitwise what lgd_adjoint_grad reports
 * after lgd_adjoint_reset and the same one step.  Calls no allocator.  Grids above
 * LGD_NOALLOC_ADJOINT_CELLS interior cells, grids the streaming reverse step does not
 * cover, and the "adjoint_kernel" = 1 setting return LGD_ERR_UNSUPPORTED.  Every call
 * on one GPU shares the array: calls that overlap on one GPU (other streams or
 * threads) race. */
#define LGD_NOALLOC_ADJOINT_CELLS (1u << 20)
int lgd_modelc_adjoint_step_noalloc(lgd_ctx *ctx, const lgd_grid *grid,
                                    const lgd_modelc_params *params, const lgd_buf *phi,
                                    const lgd_buf *u, double dt, lgd_buf *pphi, lgd_buf *pu,
                                    lgd_buf *grad);

/* Calibration: L(θ, φ₀, u₀) = ½‖φ_T − target‖² over the interiors after nsteps steps,
 * as a forward and a backward call, reversed by binomial checkpointing in `snapshots`
 * state slots (the initial state included; >= 1).  Results are bitwise independent of
 * `snapshots`. */
typedef struct lgd_calibration lgd_calibration;
int lgd_calibration_create(lgd_ctx *ctx, const lgd_grid *grid, int64_t nsteps, double dt,
                           int32_t snapshots, lgd_calibration **out);
int lgd_calibration_destroy(lgd_calibration *cal);
/* The forward sweep (keeping the schedule's first checkpoints); returns the loss. */
int lgd_calibration_forward(lgd_calibration *cal, const lgd_modelc_params *params,
                            const lgd_buf *phi0, const lgd_buf *u0, const lgd_buf *target,
                            double *loss);
/* The reverse sweep after lgd_calibration_forward: grad = dL/dθ (3 values);
 * gphi0 / gu0 = dL/d(φ₀, u₀) on the interiors (either may be NULL). */
int lgd_calibration_backward(lgd_calibration *cal, double *grad, lgd_buf *gphi0, lgd_buf *gu0);
/* Forward steps executed so far (sweep and recomputation) and reverse steps. */
int lgd_calibration_counts(const lgd_calibration *cal, int64_t *forward_steps,
                           int64_t *reverse_steps);
/* Forward steps of one forward + backward pass under the checkpointing schedule
 * (host arithmetic only, no device work). */
int64_t lgd_checkpoint_forward_steps(int64_t nsteps, int32_t snapshots);

#ifdef __cplusplus
}
#endif
#endif /* LEGENDRE_GPU_H */
