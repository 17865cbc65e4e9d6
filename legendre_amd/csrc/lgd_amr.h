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
};

struct lgd_level_plan {
  lgd_tile *tiles = nullptr;
  int64_t ntiles = 0;
  std::vector<int64_t> tile0;  // first tile of each of the level's patches, + end
  lgd_fill_op *ops = nullptr;
  int64_t nops = 0;
  lgd_corner *corners = nullptr;
  int64_t ncorners = 0;
  lgd_restrict_item *ritems = nullptr;  // restriction of this level onto level-1
  int64_t nritems = 0;
};

struct lgd_amr {
  lgd_ctx *ctx = nullptr;
  lgd_cart base{};
  int64_t cells[2] = {0, 0};
  std::vector<uint32_t> ratios;
  std::vector<int32_t> level;         // per patch (level-major)
  std::vector<lgd_box> box;           // per patch, level cell coordinates
  std::vector<uint32_t> level_start;  // populated levels + 1
  std::vector<std::array<double, 2>> spacing;  // per level, ratio capacity
  // Per populated level: patches bucketed by 16×16-cell squares of the level domain
  // (CSR: bucket k holds ids[start[k] .. start[k+1])), for find_patch.
  struct bucket_index {
    int64_t nbx = 0, nby = 0;
    std::vector<uint32_t> start, ids;
  };
  static constexpr int64_t kBucket = 16;
  std::vector<bucket_index> lookup;
  std::vector<int64_t> off;           // ghost-1 slab offsets, nblocks + 1
  lgd_block_desc *d_blocks = nullptr;
  std::vector<lgd_level_plan> plans;  // per populated level
  // find_patch on the device: per-level headers, concatenated bucket starts (into
  // the concatenated ids), patch ids, and each patch's lower corner (x, y).
  lgd_dev_level *d_levels = nullptr;
  uint32_t *d_bstart = nullptr;
  uint32_t *d_bids = nullptr;
  int64_t *d_lo = nullptr;

  int num_levels() const { return (int)level_start.size() - 1; }
  int max_levels() const { return (int)ratios.size() + 1; }
  lgd_box domain(int l) const;
  int64_t find_patch(int l, int64_t x, int64_t y) const;
  int64_t index(int64_t b, int64_t i, int64_t j) const;  // element of interior (i, j)
  void free_device();
};

int lgd_amr_build(lgd_ctx *ctx, const lgd_cart &base, const std::vector<uint32_t> &ratios,
                  const std::vector<std::vector<lgd_box>> &levels, lgd_amr &g);
int lgd_amr_build_plans(lgd_amr &g);

// Device transfers and level kernels (lgd_amr.hip).
int lgd_amr_fill(lgd_ctx *ctx, const lgd_amr &g, double *field, const double *old, double alpha,
                 int level);
int lgd_amr_restrict(lgd_ctx *ctx, const lgd_amr &g, double *field, int fine_level);
int lgd_amr_modelc_rhs(lgd_ctx *ctx, const lgd_amr &g, const lgd_modelc_params &p, int level,
                       const double *phi, const double *u, double *kphi, double *ku);
int lgd_amr_heat_rhs(lgd_ctx *ctx, const lgd_amr &g, double kappa, int level, const double *u,
                     double *k);
int lgd_amr_apply(lgd_ctx *ctx, const lgd_amr &g, int level, double *x, const double *k, double dt);
int lgd_amr_wiener(lgd_ctx *ctx, const lgd_amr &g, int level, double *phi, double noise_scale,
                   uint64_t seed, uint64_t salt);
