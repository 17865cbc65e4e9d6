// Internal AMR structures shared by lgd_amr.hip and lgd_sim.hip.
#pragma once

#include <array>
#include <vector>

#include "lgd_internal.h"

// A half-open box of cells at one level (CellBox, src/geometry/amr/cluster.rs:29-37).
struct lgd_box {
  int64_t lo[2];
  int64_t hi[2];
  int64_t extent(int d) const { return hi[d] - lo[d]; }
  uint64_t cells() const {
    return extent(0) > 0 && extent(1) > 0 ? (uint64_t)extent(0) * (uint64_t)extent(1) : 0;
  }
  bool contains(int64_t x, int64_t y) const {
    return lo[0] <= x && x < hi[0] && lo[1] <= y && y < hi[1];
  }
  bool intersects(const lgd_box &o) const {
    return lo[0] < o.hi[0] && o.lo[0] < hi[0] && lo[1] < o.hi[1] && o.lo[1] < hi[1];
  }
  lgd_box clipped(const lgd_box &d) const {
    lgd_box o = *this;
    for (int k = 0; k < 2; ++k) {
      o.lo[k] = lo[k] > d.lo[k] ? lo[k] : d.lo[k];
      int64_t h = hi[k] < d.hi[k] ? hi[k] : d.hi[k];
      o.hi[k] = h > o.lo[k] ? h : o.lo[k];
    }
    return o;
  }
  bool operator==(const lgd_box &o) const {
    return lo[0] == o.lo[0] && lo[1] == o.lo[1] && hi[0] == o.hi[0] && hi[1] == o.hi[1];
  }
};

// Device block table entry: interior extent, level, ghost-1 slab offset.
struct lgd_block_desc {
  int32_t nx, ny;
  int32_t level;
  int32_t pad;
  int64_t off;
};

// One workgroup's share of a level: a 64×4 cell tile of one block.
struct lgd_tile {
  int32_t block, x0, y0, pad;
};

// One wavefront's share of a level on the streaming level stage: rows [j0, j1) of
// the interior columns [x0, x0 + w) of one patch (w <= 62: lanes 1..w hold them,
// lanes 0 and w + 1 the columns either side).
struct lgd_sseg {
  int64_t off;     // the patch's ghost-1 slab offset
  int32_t block;   // patch id (noise key)
  int32_t nx;      // patch interior width (slab rows hold nx + 2)
  int32_t x0, w;
  int32_t j0, j1;
};

// Interior columns per wavefront of the streaming level stage.
constexpr int kSegCols = 62;

// One ghost cell of the fill plan: copy `src`, or prolong from corners
// [cstart, cstart + clen).
struct lgd_fill_op {
  int64_t target;
  int64_t src;
  int32_t cstart;
  int32_t clen;
};

struct lgd_corner {
  int64_t src;
  double weight;
};

// Of a copy op whose source cell is covered by the next finer level: the element
// of that cell's first child (child < 0: not covered) and the children's slab row
// stride, so a launch that restricts the finer level can also fill the ghost.
struct lgd_copy_child {
  int64_t child;
  int64_t stride;
};

// One covered coarse cell: target element, child (0,0) element, fine row stride.
struct lgd_restrict_item {
  int64_t target;
  int64_t child;
  int64_t stride;
};

// Device copy of one level's find_patch index: domain extent, bucket grid and the
// level's first entry in the concatenated bucket starts.
struct lgd_dev_level {
  int64_t w, h;
  int64_t nbx, nby;
  int64_t start0;
  int64_t bucket;  // bucket side in cells
};

struct lgd_level_plan {
  void *block = nullptr;  // one device allocation holding the four arrays below
  lgd_tile *tiles = nullptr;
  int64_t ntiles = 0;
  std::vector<int64_t> tile0;  // first tile of each of the level's patches, + end
  lgd_fill_op *ops = nullptr;
  int64_t nops = 0;
  lgd_corner *corners = nullptr;  // four slots per op: op t's corners start at 4t
  int64_t ncorners = 0;
  lgd_copy_child *cchild = nullptr;  // per op, on levels with a finer level (else null)
  lgd_restrict_item *ritems = nullptr;  // restriction of this level onto level-1
  int64_t nritems = 0;
  // First restriction item of each of the level's patches, + end (device): patch p's
  // items are its coarse cells row by row, (nx / r) per row.
  int64_t *rit0 = nullptr;
  // The streaming level stage: column segments of every patch cut into strips of
  // seg_rows rows (patch-major; seg0 = first item of each patch, + end).
  lgd_sseg *segs = nullptr;
  int64_t nsegs = 0;
  int32_t seg_rows = 0;
  int32_t auto_rows = 0;  // lgd_amr_default_seg_rows of the level
  std::vector<int64_t> seg0;
  // The same in a tuned strip height other than seg_rows (allocated on first use).
  mutable lgd_sseg *tsegs = nullptr;
  mutable int64_t ntsegs = 0;
  mutable int32_t trows = 0;
  mutable std::vector<int64_t> tseg0;
  mutable std::vector<lgd_sseg> tsegs_host;  // host copy kept until the plan is freed
  // The level-stage kernels' tiling in a tuned shape other than the plan's 64×4
  // (allocated on first use of that shape; the regrid kernels keep `tiles`).
  mutable lgd_tile *stiles = nullptr;
  mutable int64_t nstiles = 0;
  mutable int32_t stx = 0, sty = 0;
  mutable std::vector<int64_t> stile0;
};

// The level-stage tile shapes the "amr_tile" knob selects (threads x, y per workgroup,
// one cell per thread); shape 0 is the plans' own 64×4 tiling, which the regrid
// kernels share.
constexpr int kAmrTileShapes[][2] = {{64, 4}, {32, 8}, {16, 16}, {128, 2}, {64, 2}, {32, 4}, {256, 1}};

// Rows per strip of the streaming level stage when no "amr_rows" knob is set: the
// level's segment rows spread to about two wavefronts per SIMD of the device, 2 to
// 32 rows per strip.
int lgd_amr_default_seg_rows(const lgd_amr &g, int level);
// Whether `level` steps through the streaming level stage (Model C, batched
// launches, the "amr_kernel" knob not asking for the per-cell tiles).
bool lgd_amr_level_streams(const lgd_ctx *ctx, const lgd_amr &g, int level);
// Whether the first fill of level + 1 may share the stage launch of `level`
// (lgd_amr_modelc_step's `cf`): a streamed level with a finer one (level 0 on the
// Cartesian streaming kernel too), batched launches, the "amr_split_fill" knob unset.
bool lgd_amr_child_fill_fuses(const lgd_ctx *ctx, const lgd_amr &g, int level);
// Whether the restriction of `level` into its parent (and the parent's next fill) may
// share the level's stage launch (lgd_amr_modelc_step's `ps`): a streamed level
// >= 1 refined by 2 over its parent, an even strip height, batched launches, the
// "amr_split_fill" knob unset.
bool lgd_amr_restrict_fuses(const lgd_ctx *ctx, const lgd_amr &g, int level);

// Workgroup width of the AMR fill and restriction launches ("amr_wg"; 256 by default).
inline int lgd_amr_wg(const lgd_ctx *ctx) { return ctx->amr_wg > 0 ? (int)ctx->amr_wg : 256; }

struct lgd_amr {
  lgd_ctx *ctx = nullptr;
  lgd_cart base{};
  int64_t cells[2] = {0, 0};
  std::vector<uint32_t> ratios;
  std::vector<int32_t> level;         // per patch (level-major)
  std::vector<lgd_box> box;           // per patch, level cell coordinates
  std::vector<uint32_t> level_start;  // populated levels + 1
  std::vector<std::array<double, 2>> spacing;  // per level, ratio capacity
  // Per populated level: patches bucketed by b×b-cell squares of the level domain
  // (CSR: bucket k holds ids[start[k] .. start[k+1])), for find_patch.  b is kBucket
  // at level 0 and grows with the refinement ratios, so every level's buckets cover
  // the same area of the domain.
  struct bucket_index {
    int64_t b = 16;
    int64_t nbx = 0, nby = 0;
    std::vector<uint32_t> start, ids;
  };
  static constexpr int64_t kBucket = 16;
  std::vector<bucket_index> lookup;
  std::vector<int64_t> off;           // ghost-1 slab offsets, nblocks + 1
  void *d_arena = nullptr;  // one device allocation: the patch index and the plans of levels >= 1
  size_t d_arena_cap = 0;   // its bytes (it returns to the context's retired arenas)
  lgd_block_desc *d_blocks = nullptr;
  std::vector<lgd_level_plan> plans;  // per populated level
  // find_patch on the device: per-level headers, concatenated bucket starts (into
  // the concatenated ids), patch ids, and each patch's lower corner (x, y).
  lgd_dev_level *d_levels = nullptr;
  uint32_t *d_bstart = nullptr;
  uint32_t *d_bids = nullptr;
  int64_t *d_lo = nullptr;
  int32_t *d_err = nullptr;  // set by the device plan builders when a ghost has no donor

  int num_levels() const { return (int)level_start.size() - 1; }
  int max_levels() const { return (int)ratios.size() + 1; }
  lgd_box domain(int l) const;
  int64_t find_patch(int l, int64_t x, int64_t y) const;
  int64_t index(int64_t b, int64_t i, int64_t j) const;  // element of interior (i, j)
  void free_device();
};

// find_patch on the device (the hierarchy's d_levels / d_bstart / d_bids / d_lo /
// d_blocks): the patch of level l holding cell (x, y), -1 if none.
struct lgd_dev_index {
  const lgd_dev_level *levels;
  const uint32_t *bstart;
  const uint32_t *bids;
  const int64_t *lo;
  const lgd_block_desc *blocks;
  int nlevels;
};

__device__ inline int64_t lgd_dev_find_patch(const lgd_dev_index &ix, int l, int64_t x, int64_t y) {
  if (l >= ix.nlevels) return -1;
  const lgd_dev_level L = ix.levels[l];
  if (x < 0 || y < 0 || x >= L.w || y >= L.h) return -1;
  const int64_t k = L.start0 + x / L.bucket + L.nbx * (y / L.bucket);
  for (uint32_t i = ix.bstart[k]; i < ix.bstart[k + 1]; ++i) {
    const uint32_t p = ix.bids[i];
    const int64_t px = ix.lo[2 * p], py = ix.lo[2 * p + 1];
    const lgd_block_desc bd = ix.blocks[p];
    if (x >= px && x < px + bd.nx && y >= py && y < py + bd.ny) return (int64_t)p;
  }
  return -1;
}

// `reuse0`: a hierarchy on the same base whose level-0 plan (which depends on the
// base tiling alone) g will take over instead of building its own (a regrid's old
// hierarchy): g's level-0 plan is then left empty until lgd_amr_take_level0_plan,
// called once the caller commits to g.
int lgd_amr_build(lgd_ctx *ctx, const lgd_cart &base, const std::vector<uint32_t> &ratios,
                  const std::vector<std::vector<lgd_box>> &levels, lgd_amr &g,
                  const lgd_amr *reuse0 = nullptr);
int lgd_amr_build_plans(lgd_amr &g, const lgd_amr *reuse0 = nullptr);
// Queues the copy of the device plan builders' error flag into *host_flag on the
// context's stream (valid once the stream has synchronised): nonzero when a ghost
// cell of the hierarchy found no donor.
int lgd_amr_plan_error(lgd_ctx *ctx, const lgd_amr &g, int32_t *host_flag);
void lgd_amr_take_level0_plan(lgd_amr &g, lgd_amr &from);

// Device transfers and level kernels (lgd_amr.hip).
int lgd_amr_fill(lgd_ctx *ctx, const lgd_amr &g, double *field, const double *old, double alpha,
                 int level);
int lgd_amr_restrict(lgd_ctx *ctx, const lgd_amr &g, double *field, int fine_level);


// Ping-pong stepping (each level's current values in one of two buffers per
// field).  nf <= 2 fields per launch.
// Fill plan of `level`: lvl = the level's current buffers (copies read, ghosts
// written); par / old = the parent's current and previous buffers (prolongation,
// blended old·(1−α) + new·α when old is given; par may be null at level 0).
int lgd_amr_fill_pp(lgd_ctx *ctx, const lgd_amr &g, int level, int nf, double *const *lvl,
                    const double *const *par, const double *const *old, double alpha);
// Restriction of fine_level into its parent, both in their current buffers.
int lgd_amr_restrict_pp(lgd_ctx *ctx, const lgd_amr &g, int fine_level, int nf,
                        const double *const *fine, double *const *coarse);
// Both in one launch: restriction of level + 1 (`fine`, current) into `level`
// (lvl, current) and the fill of `level` as lgd_amr_fill_pp; a ghost copied from a
// cell the restriction writes takes the restricted value, computed the same way.
// Bitwise the two launches in sequence.
int lgd_amr_restrict_fill_pp(lgd_ctx *ctx, const lgd_amr &g, int level, int nf, double *const *lvl,
                             const double *const *par, const double *const *old, double alpha,
                             const double *const *fine);
// Whether level 0 steps through the Cartesian streaming kernel (which reads
// neighbour interiors itself, so level 0 needs no ghost fill).
bool lgd_amr_level0_streams(const lgd_ctx *ctx, const lgd_amr &g);
// The untuned shape of level 0 on the streaming kernel: two x-neighbouring blocks
// per wavefront where the base tiling allows it (blocks at most 64 wide leave half
// of a wavefront idle otherwise), else one.
lgd_launch_shape lgd_amr_level0_default_shape(const lgd_amr &g, bool noise);
// The ghosts of the next substep of a level without children, written by the stage
// launch that produces that substep's input: the parent's current and previous
// buffers (fields 0, 1) and the next substep's α.  Bitwise the separate fill.
struct lgd_amr_next_fill {
  const double *par[2];
  const double *old[2];
  double alpha;
};
// The first fill of level + 1 at α = 0 in the stage launch of `level`: the child's
// current buffers (fields 0, 1), whose ghosts are written.  A prolonged ghost is
// old·1 + new·0 of the parent; the corner sum of the old values is never −0 (it
// starts from +0), so that is the old sum bit for bit whenever the new values are
// finite: the bits of the fill launched after the stage for every finite state.
struct lgd_amr_child_fill {
  double *lvl[2];
};
// The end of a level's last substep within its parent's substep, in the stage
// launch of that substep: the restriction into the parent (coarse: the parent's
// buffers that take it, fields 0, 1) and, with `fill`, the parent's ghosts for its
// next substep in those buffers (copies, the copies from cells under this level
// restricted from the stage's own new values, and prolongation from the
// grandparent's current (par) and previous (old) buffers blended at alpha).  Each
// wavefront restricts the 2×2 blocks of its strip (ratio 2, an even strip height);
// the fill's threads recompute the four children of a covered source from the
// stage's inputs.  Bitwise the restriction and fill launched after the stage.
struct lgd_amr_parent_step {
  double *coarse[2];
  bool fill;
  const double *par[2];
  const double *old[2];
  double alpha;
};
// One fused Euler–Maruyama stage of `level`: (phi, u) → (phi_n, u_n) interiors, and
// with `nf` the ghosts of (phi_n, u_n) the level's next substep reads, or with `cf`
// (a streamed level with a finer one; LGD_ERR_UNSUPPORTED otherwise) the first fill of
// level + 1.
int lgd_amr_modelc_step(lgd_ctx *ctx, const lgd_amr &g, const lgd_modelc_params &p, int level,
                        double dt, const double *phi, const double *u, double *phi_n, double *u_n,
                        uint64_t seed, uint64_t salt, bool wiener,
                        const lgd_amr_next_fill *nf = nullptr,
                        const lgd_amr_child_fill *cf = nullptr,
                        const lgd_amr_parent_step *ps = nullptr);
int lgd_amr_heat_step(lgd_ctx *ctx, const lgd_amr &g, double kappa, int level, double dt,
                      const double *u, double *u_n, const lgd_amr_next_fill *nf = nullptr);
// Copy one level's range (interiors and ghosts) between buffers of the layout.
int lgd_amr_copy_level(lgd_ctx *ctx, const lgd_amr &g, int level, const double *src, double *dst);
