// Block-structured AMR on the device (2-D): the patch hierarchy as a device
// block table, and the Berger–Oliger grid transfers as gather kernels over
// host-built plans.
//
// Host side (construction / regrid cadence, integer work): AmrGrid validation
// (src/geometry/amr/grid.rs:90-196), the per-level ownership maps behind
// find_patch, and the transfer plans:
//   * fill plan per level — one op per ghost-ring cell of every patch: copy one
//     interior cell (same-level exchange, mirror) or prolong bilinearly from the
//     coarser level's interiors (corner cells + weights), the FillPlan of
//     src/geometry/amr/intergrid.rs:225-351;
//   * restriction items per fine level — one per covered coarse cell
//     (intergrid.rs:110-150);
//   * tiles per level — (block, 64×4 cell tile) work items for the level kernels.
// Device side (every substep): the ops are applied by one launch per field and
// level; sources are always interiors and targets always ghosts (fill) or
// covered coarse interiors (restriction), so one launch has no write conflicts
// and visit order cannot change results (intergrid.rs:23-26).
#include <algorithm>
#include <array>
#include <climits>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <new>
#include <vector>

#include "lgd_amr.h"
#include "lgd_stream_ops.h"

using namespace lgd;

namespace {

constexpr int kTX = 64;
constexpr int kTY = 4;

inline int64_t reflect_coord(int64_t c, int64_t lo, int64_t hi) {
  if (c < lo) return 2 * lo - 1 - c;
  if (c >= hi) return 2 * hi - 1 - c;
  return c;
}

inline int64_t floor_div(int64_t a, int64_t b) {
  int64_t q = a / b;
  return (a % b < 0) ? q - 1 : q;
}

// The streaming level stage's items of `level`: each patch cut into balanced column
// segments of at most kSegCols columns and strips of `rows` rows, patch-major.
void build_segs(const lgd_amr &g, int level, int rows, std::vector<lgd_sseg> &out,
                std::vector<int64_t> &seg0) {
  out.clear();
  seg0.clear();
  for (uint32_t b = g.level_start[level]; b < g.level_start[level + 1]; ++b) {
    seg0.push_back((int64_t)out.size());
    const int64_t nx = g.box[b].extent(0), ny = g.box[b].extent(1);
    const int64_t nseg = (nx + kSegCols - 1) / kSegCols;
    // Balanced widths; an even-width patch cuts at even columns (at most kSegCols
    // wide still), so a ratio-2 block of 2×2 cells never spans two segments.
    const int64_t even = nx % 2 == 0 ? ~int64_t(1) : ~int64_t(0);
    for (int64_t s = 0; s < nseg; ++s) {
      const int64_t x0 = (s * nx / nseg) & even, x1 = ((s + 1) * nx / nseg) & even;
      for (int64_t j0 = 0; j0 < ny; j0 += rows)
        out.push_back(lgd_sseg{g.off[b], (int32_t)b, (int32_t)nx, (int32_t)x0, (int32_t)(x1 - x0),
                               (int32_t)j0, (int32_t)std::min<int64_t>(j0 + rows, ny)});
    }
  }
  seg0.push_back((int64_t)out.size());
}

// Several host arrays uploaded as one device allocation (256-byte aligned parts),
// so a hierarchy costs a handful of allocations, not one per array.  The arrays are
// gathered into the context's pinned staging buffer and copied with one
// stream-ordered copy into one stream-ordered allocation from the device's memory
// pool: a regrid neither synchronises the device nor pays for a fresh allocation,
// and the host returns once the bytes are staged.  The vectors added must live
// until upload returns.
struct Packer {
  struct Part {
    const void *src;
    size_t bytes, off;
    void **dptr;
  };
  std::vector<Part> parts, space;  // uploaded arrays; device-only arrays (no host bytes)
  size_t total = 0;                // bytes of the uploaded arrays
  template <class T>
  void add(const std::vector<T> &v, T **dptr) {
    const size_t off = (total + 255) & ~size_t(255);
    parts.push_back({v.data(), v.size() * sizeof(T), off, (void **)dptr});
    total = off + v.size() * sizeof(T);
  }
  // n elements the device fills in, placed after every uploaded array.
  template <class T>
  void reserve(size_t n, T **dptr) {
    space.push_back({nullptr, n * sizeof(T), 0, (void **)dptr});
  }
  // With `cap`, the device block comes from the context's retired arenas (its size
  // in *cap), else from the stream-ordered pool.
  hipError_t alloc(lgd_ctx *ctx, size_t bytes, void **base, size_t *cap) {
    if (cap) return lgd_arena_take(ctx, bytes, base, cap);
    return hipMallocAsync(base, bytes, ctx->stream);
  }
  int upload(lgd_ctx *ctx, void **base, size_t *cap = nullptr) {
    *base = nullptr;
    for (auto &p : parts) *p.dptr = nullptr;
    for (auto &p : space) *p.dptr = nullptr;
    size_t end = total;
    for (auto &p : space) {
      p.off = (end + 255) & ~size_t(255);
      end = p.off + p.bytes;
    }
    if (end == 0) return LGD_OK;
    if (total == 0) {
      hipError_t e = alloc(ctx, end, base, cap);
      if (e != hipSuccess) {
        *base = nullptr;
        return hip_fail(ctx, e, "hipMallocAsync(amr plan)");
      }
      for (auto &p : space) *p.dptr = (char *)*base + p.off;
      return LGD_OK;
    }
    hipError_t e = hipSuccess;
    // The staging buffer may still feed the previous upload's copy.
    if (ctx->pinned_done) e = hipEventSynchronize(ctx->pinned_done);
    else e = hipEventCreateWithFlags(&ctx->pinned_done, hipEventDisableTiming);
    if (e != hipSuccess) return hip_fail(ctx, e, "amr plan staging event");
    if (ctx->pinned_cap < total) {
      if (ctx->pinned) (void)hipHostFree(ctx->pinned);
      ctx->pinned = nullptr;
      ctx->pinned_cap = 0;
      const size_t cap = std::max(total + total / 2, (size_t)1 << 20);
      e = hipHostMalloc((void **)&ctx->pinned, cap, hipHostMallocDefault);
      if (e != hipSuccess) {
        ctx->pinned = nullptr;
        return hip_fail(ctx, e, "hipHostMalloc(amr plan staging)");
      }
      ctx->pinned_cap = cap;
    }
    for (const auto &p : parts)
      if (p.bytes) std::memcpy(ctx->pinned + p.off, p.src, p.bytes);
    e = alloc(ctx, end, base, cap);
    if (e != hipSuccess) {
      *base = nullptr;
      return hip_fail(ctx, e, "hipMallocAsync(amr plan)");
    }
    e = hipMemcpyAsync(*base, ctx->pinned, total, hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = hipEventRecord(ctx->pinned_done, ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "upload(amr plan)");
    for (auto &p : parts) *p.dptr = (char *)*base + p.off;
    for (auto &p : space) *p.dptr = (char *)*base + p.off;
    return LGD_OK;
  }
};

// Whether every cell of `need` lies in a patch of level l (run-skipping walk along
// axis 0 through the level's bucket index).
bool covered(const lgd_amr &g, int l, const lgd_box &need) {
  if (need.cells() == 0) return true;
  for (int64_t y = need.lo[1]; y < need.hi[1]; ++y) {
    int64_t x = need.lo[0];
    while (x < need.hi[0]) {
      const int64_t p = g.find_patch(l, x, y);
      if (p < 0) return false;
      x = g.box[(size_t)p].hi[0];
    }
  }
  return true;
}

// Patches of one level do not overlap; each 16×16-cell bucket of the level domain
// lists the few that touch it (find_patch's index for level l, appended to lookup).
void build_bucket_index(lgd_amr &g, int l) {
  int64_t B = lgd_amr::kBucket;
  for (int i = 0; i < l; ++i) B *= (int64_t)g.ratios[(size_t)i];
  const lgd_box d = g.domain(l);
  lgd_amr::bucket_index ix;
  ix.b = B;
  ix.nbx = (d.hi[0] + B - 1) / B;
  ix.nby = (d.hi[1] + B - 1) / B;
  ix.start.assign((size_t)(ix.nbx * ix.nby + 1), 0);
  auto each_bucket = [&](uint32_t p, auto &&f) {
    const lgd_box &bx = g.box[p];
    for (int64_t by = bx.lo[1] / B; by <= (bx.hi[1] - 1) / B; ++by)
      for (int64_t bx0 = bx.lo[0] / B; bx0 <= (bx.hi[0] - 1) / B; ++bx0) f(bx0 + ix.nbx * by);
  };
  for (uint32_t p = g.level_start[l]; p < g.level_start[l + 1]; ++p)
    each_bucket(p, [&](int64_t k) { ++ix.start[(size_t)k + 1]; });
  for (size_t k = 1; k < ix.start.size(); ++k) ix.start[k] += ix.start[k - 1];
  ix.ids.resize(ix.start.back());
  std::vector<uint32_t> fillpos(ix.start.begin(), ix.start.end() - 1);
  for (uint32_t p = g.level_start[l]; p < g.level_start[l + 1]; ++p)
    each_bucket(p, [&](int64_t k) { ix.ids[fillpos[(size_t)k]++] = p; });
  if ((int)g.lookup.size() <= l) g.lookup.resize((size_t)l + 1);
  g.lookup[(size_t)l] = std::move(ix);
}

}  // namespace

lgd_box lgd_amr::domain(int l) const {
  int64_t s = 1;
  for (int i = 0; i < l; ++i) s *= (int64_t)ratios[i];
  return lgd_box{{0, 0}, {cells[0] * s, cells[1] * s}};
}

int64_t lgd_amr::find_patch(int l, int64_t x, int64_t y) const {
  if (l < 0 || l >= num_levels()) return -1;
  lgd_box d = domain(l);
  if (!d.contains(x, y)) return -1;
  const bucket_index &ix = lookup[(size_t)l];
  const int64_t k = (x / ix.b) + ix.nbx * (y / ix.b);
  for (uint32_t i = ix.start[(size_t)k]; i < ix.start[(size_t)k + 1]; ++i)
    if (box[ix.ids[i]].contains(x, y)) return (int64_t)ix.ids[i];
  return -1;
}

int64_t lgd_amr::index(int64_t b, int64_t i, int64_t j) const {
  int64_t sx = box[b].extent(0) + 2;
  return off[b] + (j + 1) * sx + (i + 1);
}

void lgd_amr::free_device() {
  // Returned to the pool in stream order, after the kernels that read them.
  hipStream_t st = ctx ? ctx->stream : nullptr;
  if (d_arena) {
    if (ctx) lgd_arena_give(ctx, d_arena, d_arena_cap);
    else (void)hipFreeAsync(d_arena, st);
  }
  d_arena = nullptr;
  d_arena_cap = 0;
  d_blocks = nullptr;
  d_levels = nullptr;
  d_bstart = nullptr;
  d_bids = nullptr;
  d_lo = nullptr;
  d_err = nullptr;
  for (auto &p : plans) {
    if (p.block) (void)hipFreeAsync(p.block, st);
    if (p.stiles) (void)hipFreeAsync(p.stiles, st);
    if (p.tsegs) (void)hipFreeAsync(p.tsegs, st);
  }
  plans.clear();
}

int lgd_amr_default_seg_rows(const lgd_amr &g, int level) {
  int64_t rows = 0;
  for (uint32_t b = g.level_start[level]; b < g.level_start[level + 1]; ++b)
    rows += (g.box[b].extent(0) + kSegCols - 1) / kSegCols * g.box[b].extent(1);
  const int64_t cus = g.ctx && g.ctx->num_cus > 0 ? g.ctx->num_cus : 256;
  const int64_t waves = 2 * 4 * cus;
  // Even, so that a ratio-2 level's 2×2 blocks never straddle two strips.
  const int64_t r = std::min<int64_t>(32, std::max<int64_t>(2, (rows + waves - 1) / waves));
  return (int)(r + (r & 1));
}

// AmrGrid::from_patches (src/geometry/amr/grid.rs:90-196): level 0 is the base
// tiling; finer boxes are validated (non-empty, inside the domain, aligned to
// the ratio, disjoint, properly nested) and appended level-major.
int lgd_amr_build(lgd_ctx *ctx, const lgd_cart &base, const std::vector<uint32_t> &ratios,
                  const std::vector<std::vector<lgd_box>> &levels, lgd_amr &g,
                  const lgd_amr *reuse0) {
  if (base.dim != 2) return fail(ctx, LGD_ERR_UNSUPPORTED, "AMR hierarchies are 2-D");
  if (levels.size() > ratios.size())
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "more refined levels than ratios");
  for (uint32_t r : ratios)
    if (r < 2) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "refinement ratio below 2");
  g.ctx = ctx;
  g.base = base;
  g.cells[0] = base.n[0] * base.nb[0];
  g.cells[1] = base.n[1] * base.nb[1];
  g.ratios = ratios;
  // Level domains stay below 2^31 cells per axis (the regrid's cell coordinates are
  // 32-bit).
  for (int l = 0; l <= (int)ratios.size(); ++l) {
    const lgd_box d = g.domain(l);
    if (d.hi[0] > INT32_MAX || d.hi[1] > INT32_MAX)
      return fail(ctx, LGD_ERR_UNSUPPORTED, "an AMR level domain of 2^31 cells or more per axis");
  }
  g.level.clear();
  g.box.clear();
  for (int64_t by = 0; by < base.nb[1]; ++by)
    for (int64_t bx = 0; bx < base.nb[0]; ++bx) {
      g.level.push_back(0);
      g.box.push_back(lgd_box{{bx * base.n[0], by * base.n[1]},
                              {(bx + 1) * base.n[0], (by + 1) * base.n[1]}});
    }
  g.level_start = {0, (uint32_t)g.box.size()};
  g.lookup.clear();
  build_bucket_index(g, 0);
  for (size_t l = 0; l < levels.size(); ++l) {
    const int lev = (int)l + 1;
    const int64_t r = ratios[l];
    const lgd_box dom = g.domain(lev), cdom = g.domain(lev - 1);
    const std::vector<lgd_box> &boxes = levels[l];
    for (size_t i = 0; i < boxes.size(); ++i) {
      const lgd_box &b = boxes[i];
      char m[160];
      if (b.cells() == 0) {
        snprintf(m, sizeof m, "empty patch at level %d", lev);
        return fail(ctx, LGD_ERR_INVALID_ARGUMENT, m);
      }
      if (b.lo[0] < dom.lo[0] || b.lo[1] < dom.lo[1] || b.hi[0] > dom.hi[0] || b.hi[1] > dom.hi[1]) {
        snprintf(m, sizeof m, "patch outside the level-%d domain", lev);
        return fail(ctx, LGD_ERR_INVALID_ARGUMENT, m);
      }
      for (int d = 0; d < 2; ++d)
        if (b.lo[d] % r != 0 || b.hi[d] % r != 0) {
          snprintf(m, sizeof m, "patch not aligned to the refinement ratio at level %d", lev);
          return fail(ctx, LGD_ERR_INVALID_ARGUMENT, m);
        }
      for (size_t k = 0; k < i; ++k)
        if (b.intersects(boxes[k])) {
          snprintf(m, sizeof m, "overlapping patches at level %d", lev);
          return fail(ctx, LGD_ERR_INVALID_ARGUMENT, m);
        }
      lgd_box need{{floor_div(b.lo[0], r) - 1, floor_div(b.lo[1], r) - 1},
                   {floor_div(b.hi[0] + r - 1, r) + 1, floor_div(b.hi[1] + r - 1, r) + 1}};
      need = need.clipped(cdom);
      if (!covered(g, lev - 1, need)) {
        snprintf(m, sizeof m, "level-%d patch is not properly nested in level %d", lev, lev - 1);
        return fail(ctx, LGD_ERR_INVALID_ARGUMENT, m);
      }
    }
    for (const lgd_box &b : boxes) {
      g.level.push_back(lev);
      g.box.push_back(b);
    }
    g.level_start.push_back((uint32_t)g.box.size());
    build_bucket_index(g, lev);
  }
  g.spacing.assign(1, {base.h[0], base.h[1]});
  for (uint32_t r : ratios) {
    auto p = g.spacing.back();
    g.spacing.push_back({p[0] / (double)r, p[1] / (double)r});
  }
  g.off.assign(g.box.size() + 1, 0);
  for (size_t b = 0; b < g.box.size(); ++b)
    g.off[b + 1] = g.off[b] + (g.box[b].extent(0) + 2) * (g.box[b].extent(1) + 2);
  return lgd_amr_build_plans(g, reuse0);
}

void lgd_amr_take_level0_plan(lgd_amr &g, lgd_amr &from) {
  if (g.plans.empty() || from.plans.empty() || g.plans[0].tiles) return;
  g.plans[0] = from.plans[0];
  from.plans[0] = lgd_level_plan{};
}

namespace {

__device__ inline int64_t plan_mirror(int64_t c, int64_t hi) {
  if (c < 0) return -1 - c;
  if (c >= hi) return 2 * hi - 1 - c;
  return c;
}

// The entry of a sorted prefix array (n + 1 entries, first 0) whose range holds t.
__device__ inline int plan_bucket(const int64_t *pre, int n, int64_t t) {
  int lo = 0, hi = n;
  while (hi - lo > 1) {
    const int mid = (lo + hi) >> 1;
    if (pre[mid] <= t) lo = mid;
    else hi = mid;
  }
  return lo;
}

// Level l's fill plan on the device, one thread per ghost-ring cell: patch b0 + p's
// ring cells are entries [ring0[p], ring0[p + 1]) in the host's visiting order (the
// rows below and above, left to right, then the columns left and right, bottom to
// top).  Each entry is the host plan's: a copy from the same-level patch holding the
// mirrored cell, or the bilinear prolongation stencil of prolong_corners
// (intergrid.rs:307-351) in corner bitmask order with zero weights skipped and the
// corners mirrored into the coarse domain, with the same floating-point operations —
// at most four corners, in the op's own slots 4t..4t+3.  A ghost with no donor (a
// hierarchy that is not properly nested) sets *err.
__global__ void plan_fill_kernel(lgd_dev_index ix, int l, uint32_t b0, int32_t np,
                                 const int64_t *__restrict__ ring0, int64_t total, int64_t dw,
                                 int64_t dh, int64_t cw, int64_t ch, double r, lgd_fill_op *ops,
                                 lgd_corner *corners, int64_t rf, lgd_copy_child *cchild,
                                 int32_t *err) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int p = plan_bucket(ring0, np, t);
  const uint32_t b = b0 + (uint32_t)p;
  const int64_t k = t - ring0[p];
  const lgd_block_desc bd = ix.blocks[b];
  const int64_t px = ix.lo[2 * b], py = ix.lo[2 * b + 1];
  const int64_t W = (int64_t)bd.nx + 2;
  int64_t x, y;
  if (k < 2 * W) {
    x = px - 1 + k / 2;
    y = (k & 1) ? py + bd.ny : py - 1;
  } else {
    const int64_t k2 = k - 2 * W;
    y = py + k2 / 2;
    x = (k2 & 1) ? px + bd.nx : px - 1;
  }
  lgd_fill_op op;
  op.target = bd.off + (y - py + 1) * W + (x - px + 1);
  const int64_t tx = plan_mirror(x, dw), ty = plan_mirror(y, dh);
  const int64_t src = lgd_dev_find_patch(ix, l, tx, ty);
  if (src >= 0) {
    const lgd_block_desc sd = ix.blocks[src];
    op.src = sd.off + (ty - ix.lo[2 * src + 1] + 1) * ((int64_t)sd.nx + 2) + (tx - ix.lo[2 * src] + 1);
    op.cstart = (int32_t)src;
    op.clen = 0;
    ops[t] = op;
    if (cchild) {
      // The source cell's children at level l + 1 (patches there are aligned to rf).
      const int64_t fb = lgd_dev_find_patch(ix, l + 1, tx * rf, ty * rf);
      lgd_copy_child cc{-1, 0};
      if (fb >= 0) {
        const lgd_block_desc fd = ix.blocks[fb];
        cc.stride = (int64_t)fd.nx + 2;
        cc.child = fd.off + (ty * rf - ix.lo[2 * fb + 1] + 1) * cc.stride + (tx * rf - ix.lo[2 * fb] + 1);
      }
      cchild[t] = cc;
    }
    return;
  }
  if (cchild) cchild[t] = lgd_copy_child{-1, 0};
  op.src = -1;
  op.cstart = (int32_t)(4 * t);
  int n = 0;
  if (l == 0) {
    *err = 1;  // a level-0 ghost always has a same-level donor
  } else {
    int64_t c0[2];
    double w[2];
    const int64_t cell[2] = {tx, ty};
    for (int d = 0; d < 2; ++d) {
      const double pc = ((double)cell[d] + 0.5) / r - 0.5;
      const double f = floor(pc);
      c0[d] = (int64_t)f;
      w[d] = pc - f;
    }
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
      cc[0] = plan_mirror(cc[0], cw);
      cc[1] = plan_mirror(cc[1], ch);
      const int64_t q = lgd_dev_find_patch(ix, l - 1, cc[0], cc[1]);
      lgd_corner c;
      c.weight = weight;
      if (q < 0) {
        c.src = -1;
        *err = 1;
      } else {
        const lgd_block_desc qd = ix.blocks[q];
        c.src = qd.off + (cc[1] - ix.lo[2 * q + 1] + 1) * ((int64_t)qd.nx + 2) +
                (cc[0] - ix.lo[2 * q] + 1);
      }
      corners[4 * t + n++] = c;
    }
  }
  op.clen = n;
  ops[t] = op;
}

// Level l's restriction onto level l - 1 on the device, one thread per coarse cell
// under a patch: patch b0 + p covers the coarse cells [lo / r, hi / r) (aligned to
// r), entries [rit0[p], rit0[p + 1]) row by row; the coarse patch holding each is
// unique (patches of a level do not overlap).  Items write distinct coarse cells, so
// their order does not matter.  A coarse cell outside level l - 1 sets *err.
__global__ void plan_restrict_kernel(lgd_dev_index ix, int l, uint32_t b0, int32_t np,
                                     const int64_t *__restrict__ rit0, int64_t total, int64_t r,
                                     lgd_restrict_item *items, int32_t *err) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= total) return;
  const int p = plan_bucket(rit0, np, t);
  const uint32_t fb = b0 + (uint32_t)p;
  const int64_t k = t - rit0[p];
  const lgd_block_desc fd = ix.blocks[fb];
  const int64_t fx = ix.lo[2 * fb], fy = ix.lo[2 * fb + 1];
  const int64_t cwid = fd.nx / r;
  const int64_t cx = fx / r + k % cwid, cy = fy / r + k / cwid;
  const int64_t cb = lgd_dev_find_patch(ix, l - 1, cx, cy);
  lgd_restrict_item it;
  it.stride = (int64_t)fd.nx + 2;
  it.child = fd.off + (cy * r - fy + 1) * it.stride + (cx * r - fx + 1);
  if (cb < 0) {
    *err = 1;
    it.target = -1;
  } else {
    const lgd_block_desc cd = ix.blocks[cb];
    it.target = cd.off + (cy - ix.lo[2 * cb + 1] + 1) * ((int64_t)cd.nx + 2) + (cx - ix.lo[2 * cb] + 1);
  }
  items[t] = it;
}

}  // namespace

int lgd_amr_plan_error(lgd_ctx *ctx, const lgd_amr &g, int32_t *host_flag) {
  *host_flag = 0;
  if (!g.d_err) return LGD_OK;
  hipError_t e = hipMemcpyAsync(host_flag, g.d_err, sizeof(int32_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "amr plan status");
  return LGD_OK;
}

int lgd_amr_build_plans(lgd_amr &g, const lgd_amr *reuse0) {
  lgd_ctx *ctx = g.ctx;
  g.free_device();
  const bool take0 = reuse0 && !reuse0->plans.empty() && reuse0->level_start.size() >= 2 &&
                     reuse0->level_start[1] == g.level_start[1] &&
                     reuse0->base.n[0] == g.base.n[0] && reuse0->base.n[1] == g.base.n[1] &&
                     reuse0->base.nb[0] == g.base.nb[0] && reuse0->base.nb[1] == g.base.nb[1];
  std::vector<lgd_block_desc> blocks(g.box.size());
  for (size_t b = 0; b < g.box.size(); ++b) {
    blocks[b].nx = (int32_t)g.box[b].extent(0);
    blocks[b].ny = (int32_t)g.box[b].extent(1);
    blocks[b].level = g.level[b];
    blocks[b].off = g.off[b];
  }
  int rc;
  // One device allocation for the patch index and the plans of levels >= 1; level 0's
  // plan (the base tiling's, which a regrid's next hierarchy takes over) gets its own.
  Packer pk, pk0;
  // find_patch's index, concatenated over the levels.
  std::vector<lgd_dev_level> lv;
  std::vector<uint32_t> bstart, bids;
  std::vector<int64_t> lo;
  {
    for (int l = 0; l < g.num_levels(); ++l) {
      const lgd_amr::bucket_index &ix = g.lookup[(size_t)l];
      const lgd_box d = g.domain(l);
      lv.push_back(lgd_dev_level{d.hi[0], d.hi[1], ix.nbx, ix.nby, (int64_t)bstart.size(), ix.b});
      const uint32_t base = (uint32_t)bids.size();
      for (uint32_t s : ix.start) bstart.push_back(base + s);
      bids.insert(bids.end(), ix.ids.begin(), ix.ids.end());
    }
    for (const lgd_box &b : g.box) {
      lo.push_back(b.lo[0]);
      lo.push_back(b.lo[1]);
    }
    pk.add(blocks, &g.d_blocks);
    pk.add(lv, &g.d_levels);
    pk.add(bstart, &g.d_bstart);
    pk.add(bids, &g.d_bids);
    pk.add(lo, &g.d_lo);
  }
  g.plans.assign((size_t)g.num_levels(), lgd_level_plan{});
  // Host arrays of every level's plan, alive until the uploads: the stage tilings
  // and the fill / restriction plans' per-patch prefix sums (the plans themselves are
  // built on the device, in arrays that are reserved, not uploaded).
  struct LevelHost {
    std::vector<lgd_tile> tiles;
    std::vector<lgd_sseg> segs;
    std::vector<int64_t> ring0, rit0;
    int64_t *d_ring0 = nullptr, *d_rit0 = nullptr;
  };
  std::vector<LevelHost> host((size_t)g.num_levels());
  const std::vector<int32_t> zero{0};
  pk.add(zero, &g.d_err);
  for (int l = 0; l < g.num_levels(); ++l) {
    lgd_level_plan &P = g.plans[(size_t)l];
    if (l == 0 && take0) continue;  // lgd_amr_take_level0_plan fills it in
    LevelHost &Hl = host[(size_t)l];
    const uint32_t b0 = g.level_start[l], b1 = g.level_start[l + 1];
    // Level kernels: one 64×4 tile of cells per workgroup.
    std::vector<lgd_tile> &tiles = Hl.tiles;
    P.tile0.clear();
    for (uint32_t b = b0; b < b1; ++b) {
      P.tile0.push_back((int64_t)tiles.size());
      for (int64_t y0 = 0; y0 < g.box[b].extent(1); y0 += kTY)
        for (int64_t x0 = 0; x0 < g.box[b].extent(0); x0 += kTX)
          tiles.push_back(lgd_tile{(int32_t)b, (int32_t)x0, (int32_t)y0, 0});
    }
    P.tile0.push_back((int64_t)tiles.size());
    // Fill plan: every ghost-ring cell (ghost width 1) of every patch.
    Hl.ring0.assign((size_t)(b1 - b0) + 1, 0);
    for (uint32_t b = b0; b < b1; ++b)
      Hl.ring0[b - b0 + 1] = Hl.ring0[b - b0] + 2 * (g.box[b].extent(0) + g.box[b].extent(1)) + 4;
    // Restriction of this level onto the coarser one: every coarse cell under a patch.
    Hl.rit0.assign((size_t)(b1 - b0) + 1, 0);
    if (l >= 1) {
      const int64_t r = g.ratios[(size_t)l - 1];
      for (uint32_t b = b0; b < b1; ++b)
        Hl.rit0[b - b0 + 1] = Hl.rit0[b - b0] + (g.box[b].extent(0) / r) * (g.box[b].extent(1) / r);
    }
    // Streaming level-stage items at the level's default strip height.
    std::vector<lgd_sseg> &segs = Hl.segs;
    P.auto_rows = lgd_amr_default_seg_rows(g, l);
    P.seg_rows = ctx->amr_rows > 0 ? (int32_t)ctx->amr_rows : P.auto_rows;
    build_segs(g, l, P.seg_rows, segs, P.seg0);
    P.nsegs = (int64_t)segs.size();
    P.ntiles = (int64_t)tiles.size();
    P.nops = Hl.ring0.back();
    P.ncorners = 4 * P.nops;
    P.nritems = Hl.rit0.back();
    Packer &to = l == 0 ? pk0 : pk;
    to.add(tiles, &P.tiles);
    to.add(segs, &P.segs);
    to.add(Hl.ring0, &Hl.d_ring0);
    to.add(Hl.rit0, &Hl.d_rit0);
    to.reserve((size_t)P.nops, &P.ops);
    to.reserve((size_t)P.ncorners, &P.corners);
    to.reserve((size_t)P.nritems, &P.ritems);
    if (l + 1 < g.num_levels()) to.reserve((size_t)P.nops, &P.cchild);
  }
  if (!take0 && (rc = pk0.upload(ctx, &g.plans[0].block))) return rc;
  if ((rc = pk.upload(ctx, &g.d_arena, &g.d_arena_cap))) return rc;
  const lgd_dev_index ix{g.d_levels, g.d_bstart, g.d_bids, g.d_lo, g.d_blocks, g.num_levels()};
  for (int l = 0; l < g.num_levels(); ++l) {
    if (l == 0 && take0) continue;
    const lgd_level_plan &P = g.plans[(size_t)l];
    const LevelHost &Hl = host[(size_t)l];
    const uint32_t b0 = g.level_start[l];
    const int32_t np = (int32_t)(g.level_start[l + 1] - b0);
    const lgd_box d = g.domain(l);
    if (P.nops > 0) {
      const lgd_box cd = l > 0 ? g.domain(l - 1) : d;
      const double r = l > 0 ? (double)g.ratios[(size_t)l - 1] : 1.0;
      plan_fill_kernel<<<blocks_for(P.nops, 256), 256, 0, ctx->stream>>>(
          ix, l, b0, np, Hl.d_ring0, P.nops, d.hi[0], d.hi[1], cd.hi[0], cd.hi[1], r, P.ops,
          P.corners, l + 1 < g.num_levels() ? (int64_t)g.ratios[(size_t)l] : 0, P.cchild, g.d_err);
      if ((rc = launch_check(ctx, "amr fill plan"))) return rc;
    }
    g.plans[(size_t)l].rit0 = Hl.d_rit0;
    if (l >= 1 && P.nritems > 0) {
      plan_restrict_kernel<<<blocks_for(P.nritems, 256), 256, 0, ctx->stream>>>(
          ix, l, b0, np, Hl.d_rit0, P.nritems, (int64_t)g.ratios[(size_t)l - 1], P.ritems, g.d_err);
      if ((rc = launch_check(ctx, "amr restriction plan"))) return rc;
    }
  }
  return LGD_OK;
}

// ------------------------------------------------------------------ kernels
namespace {

// Apply one level's fill plan (fill_level, intergrid.rs:183-223): copy or
// prolong (acc from 0 in corner order), optionally blended in time between the
// parent's snapshot `old` and its current state: old·(1−α) + new·α.
// Op t's four corner slots (4t .. 4t + 3; the plan's cstart), read beside the op
// itself so the corner loads do not wait for it; slots past clen hold no corner
// and are never followed.
__device__ __forceinline__ void load_corners(const lgd_corner *__restrict__ corners, int64_t t,
                                             lgd_corner (&k)[4]) {
#pragma unroll
  for (int c = 0; c < 4; ++c) k[c] = corners[4 * t + c];
}

// The prolongation of one field at an op (corners in order, acc from 0), blended
// in time: old·(1−α) + new·α.
template <bool kTime>
__device__ __forceinline__ double prolong_value(const lgd_corner (&k)[4], int clen, const double *nw,
                                                const double *old, double one_minus_alpha,
                                                double alpha) {
  double vn = 0.0, vo = 0.0;
#pragma unroll
  for (int c = 0; c < 4; ++c) {
    if (c < clen) {
      vn = vn + nw[k[c].src] * k[c].weight;
      if (kTime) vo = vo + old[k[c].src] * k[c].weight;
    }
  }
  return kTime ? vo * one_minus_alpha + vn * alpha : vn;
}

template <bool kTime>
__global__ void amr_fill_kernel(const lgd_fill_op *__restrict__ ops, int64_t nops,
                                const lgd_corner *__restrict__ corners, double *field,
                                const double *old, double one_minus_alpha, double alpha) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nops) return;
  lgd_corner k[4];
  load_corners(corners, t, k);
  const lgd_fill_op op = ops[t];
  const double v = op.clen == 0 ? field[op.src]
                                : prolong_value<kTime>(k, op.clen, field, old, one_minus_alpha, alpha);
  field[op.target] = v;
}

// restrict_level (intergrid.rs:110-150): coarse = (Σ children, dimension 0
// fastest, from 0) · 1/r².
__global__ void amr_restrict_kernel(const lgd_restrict_item *__restrict__ items, int64_t n,
                                    int r, double inv, double *field) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const lgd_restrict_item it = items[t];
  double sum = 0.0;
  for (int ky = 0; ky < r; ++ky)
    for (int kx = 0; kx < r; ++kx) sum = sum + field[it.child + ky * it.stride + kx];
  field[it.target] = sum * inv;
}

// ---- ping-pong stepping: each level's current values live in one of two buffers
// per field; a substep reads the current buffer and writes the other, so the
// values a level had before its step stay readable (the time-interpolation
// source of the finer level) without a snapshot copy.

// Fill plan of one level for nf fields: copies read the level's current buffer,
// prolongation corners the parent's current (new) and previous (old) buffers.
struct FieldPtrs {
  double *lvl[2];
  const double *par[2];
  const double *old[2];
};

template <bool kTime>
__global__ void amr_fill_pp_kernel(const lgd_fill_op *__restrict__ ops, int64_t nops,
                                   const lgd_corner *__restrict__ corners, FieldPtrs F, int nf,
                                   double one_minus_alpha, double alpha) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= nops) return;
  lgd_corner k[4];
  load_corners(corners, t, k);
  const lgd_fill_op op = ops[t];
  for (int f = 0; f < nf; ++f) {
    const double v = op.clen == 0
                         ? F.lvl[f][op.src]
                         : prolong_value<kTime>(k, op.clen, F.par[f], F.old[f], one_minus_alpha, alpha);
    F.lvl[f][op.target] = v;
  }
}

// Restriction of nf fields: children from the fine level's current buffer into the
// coarse level's current buffer.
struct RestrictPtrs {
  const double *fine[2];
  double *coarse[2];
};

// restrict_level's value at one coarse cell: (Σ children, dimension 0 fastest,
// from 0) · 1/r².
__device__ __forceinline__ double restricted(const double *fine, int64_t child, int64_t stride,
                                             int r, double inv) {
  double sum = 0.0;
  for (int ky = 0; ky < r; ++ky)
    for (int kx = 0; kx < r; ++kx) sum = sum + fine[child + ky * stride + kx];
  return sum * inv;
}

__global__ void amr_restrict_pp_kernel(const lgd_restrict_item *__restrict__ items, int64_t n,
                                       int r, double inv, RestrictPtrs R, int nf) {
  int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= n) return;
  const lgd_restrict_item it = items[t];
  for (int f = 0; f < nf; ++f) R.coarse[f][it.target] = restricted(R.fine[f], it.child, it.stride, r, inv);
}

// Restriction of level + 1 into `level` (threads [0, nitems)) and the fill of
// `level` (threads from nitems on) in one launch.  The restriction writes interior
// cells under the finer level and the fill writes ghost cells, so the two never
// write one element; a ghost copied from a cell under the finer level takes the
// restricted value from the children itself instead of reading the cell.
template <bool kTime>
__global__ void amr_restrict_fill_pp_kernel(const lgd_restrict_item *__restrict__ items,
                                            int64_t nitems, int r, double inv,
                                            const lgd_fill_op *__restrict__ ops, int64_t nops,
                                            const lgd_corner *__restrict__ corners,
                                            const lgd_copy_child *__restrict__ cchild, FieldPtrs F,
                                            RestrictPtrs R, int nf, double one_minus_alpha,
                                            double alpha) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < nitems) {
    const lgd_restrict_item it = items[t];
    for (int f = 0; f < nf; ++f)
      R.coarse[f][it.target] = restricted(R.fine[f], it.child, it.stride, r, inv);
    return;
  }
  const int64_t o = t - nitems;
  if (o >= nops) return;
  lgd_corner k[4];
  load_corners(corners, o, k);
  const lgd_copy_child cc = cchild[o];
  const lgd_fill_op op = ops[o];
  for (int f = 0; f < nf; ++f) {
    double v;
    if (op.clen == 0)
      v = cc.child >= 0 ? restricted(R.fine[f], cc.child, cc.stride, r, inv) : F.lvl[f][op.src];
    else
      v = prolong_value<kTime>(k, op.clen, F.par[f], F.old[f], one_minus_alpha, alpha);
    F.lvl[f][op.target] = v;
  }
}

// Fused Model C stage at one interior cell (element o of block b, interior (i, j)),
// from ghost-filled slabs, in the reference's order: the vector field
// (ModelC::vector_field_block on AmrGrid, phasefield.rs:452-475), the Time update
// x + dt·k, then the Wiener(0) increment (src/core/driver.rs:139-158) with the cell
// key per patch extent (grid.rs:407-425).
template <bool kNoise>
__device__ __forceinline__ void modelc_cell_stage(const lgd_kr_consts &k, double dt,
                                                  const double *phi, const double *u, int64_t o,
                                                  int64_t sx, double noise_scale, uint64_t seed,
                                                  uint64_t salt, int64_t b, int64_t cell,
                                                  double &pn, double &un) {
  double p[3][3];
  for (int dj = -1; dj <= 1; ++dj)
    for (int di = -1; di <= 1; ++di) p[dj + 1][di + 1] = phi[o + dj * sx + di];
  double dphi, du;
  lgd_modelc_cell(k, p, u[o], u[o + 1], u[o - 1], u[o + sx], u[o - sx], dphi, du);
  pn = p[1][1] + dt * dphi;
  if (kNoise) {
    uint64_t bk = lgd_mix_key3(seed, salt, (1ull << 32) | 0ull, (uint64_t)b);
    uint64_t key = lgd_splitmix64(bk ^ (uint64_t)cell);
    pn = pn + noise_scale * lgd_standard_normal(key);
  }
  un = u[o] + dt * du;
}

// Fused Heat forward-Euler stage at one interior cell (κ·∇²u, the Heat model of
// tests/amr.rs, then u + dt·k).
__device__ __forceinline__ double heat_cell_stage(double ihx, double ihy, double kappa, double dt,
                                                  const double *u, int64_t o, int64_t sx) {
  const double c = u[o];
  double acc = 0.0;
  acc = acc + (((u[o + 1] - (c + c)) + u[o - 1]) * ihx);
  acc = acc + (((u[o + sx] - (c + c)) + u[o - sx]) * ihy);
  return c + dt * (acc * kappa);
}

// The ghost ring the next substep of the finest level reads, written by the stage
// launch that produces the substep's input (the fill of intergrid.rs:183-223, one
// launch earlier): a copy ghost takes its source cell's new value, which the
// ghost's thread computes itself from the stage's input (the same operations and
// noise key as the thread that owns the cell, so the same bits); a prolonged ghost
// blends the parent's previous and current values at the next substep's α, which
// no substep of this level changes.
struct NextFill {
  const lgd_fill_op *ops;
  const lgd_corner *corners;
  int64_t nops;
  const double *par[2];
  const double *old[2];
  double one_minus_alpha, alpha;
};

__device__ __forceinline__ double prolong_blend(const NextFill &N, const lgd_fill_op &op, int f) {
  double vn = 0.0, vo = 0.0;
  for (int c = 0; c < op.clen; ++c) {
    const lgd_corner kc = N.corners[op.cstart + c];
    vn = vn + N.par[f][kc.src] * kc.weight;
    vo = vo + N.old[f][kc.src] * kc.weight;
  }
  return vo * N.one_minus_alpha + vn * N.alpha;
}

// The element's interior coordinates in its block (o is an interior element of b).
__device__ __forceinline__ void slab_coords(const lgd_block_desc &bd, int64_t o, int &i, int &j) {
  const int64_t local = o - bd.off, sx = bd.nx + 2;
  j = (int)(local / sx) - 1;
  i = (int)(local - (local / sx) * sx) - 1;
}

// The first fill of the next finer level at α = 0, written by the stage launch of
// its parent: the child's plan and current buffers.
struct ChildFill {
  const lgd_fill_op *ops;
  const lgd_corner *corners;
  int64_t nops;
  double *lvl[2];
};

// The end of a level's substep in its stage launch (lgd_amr_parent_step): the
// restriction into the parent through the level's restriction plan, and the
// parent's fill plan.  b0 / nb: the stepped level's patches.
struct ParentStep {
  double *coarse[2];
  const lgd_restrict_item *items;
  const int64_t *rit0;
  double inv;
  const lgd_fill_op *ops;
  const lgd_corner *corners;
  const lgd_copy_child *cchild;
  int64_t nops;
  const double *par[2];
  const double *old[2];
  double one_minus_alpha, alpha;
  uint32_t b0, nb;
};

// The stepped level's patch holding element o (patches are laid out in order of
// their slab offsets).
__device__ __forceinline__ uint32_t patch_of(const lgd_block_desc *blocks, uint32_t b0, uint32_t nb,
                                             int64_t o) {
  uint32_t lo = b0, hi = b0 + nb;
  while (hi - lo > 1) {
    const uint32_t mid = (lo + hi) >> 1;
    if (blocks[mid].off <= o) lo = mid;
    else hi = mid;
  }
  return lo;
}

// One fused Model C stage of a level: workgroups [0, ntiles) advance the tiles'
// cells; with kFill, workgroups from ntiles on write the next substep's ghosts.
template <bool kNoise, bool kFill>
__global__ __launch_bounds__(kTX *kTY) void amr_modelc_step_kernel(
    const lgd_tile *__restrict__ tiles, int64_t ntiles, const lgd_block_desc *__restrict__ blocks,
    lgd_kr_consts k, double dt, const double *phi, const double *u, double *phi_n, double *u_n,
    double noise_scale, uint64_t seed, uint64_t salt, const double *salt_src, NextFill N) {
  const uint64_t st = kNoise ? lgd_salt(salt, salt_src) : 0;
  if (!kFill || blockIdx.x < ntiles) {
    const lgd_tile tl = tiles[blockIdx.x];
    const lgd_block_desc bd = blocks[tl.block];
    const int i = tl.x0 + threadIdx.x, j = tl.y0 + threadIdx.y;
    if (i >= bd.nx || j >= bd.ny) return;
    const int64_t sx = bd.nx + 2;
    const int64_t o = bd.off + (int64_t)(j + 1) * sx + (i + 1);
    double pn, un;
    modelc_cell_stage<kNoise>(k, dt, phi, u, o, sx, noise_scale, seed, st, tl.block,
                              i + (int64_t)j * bd.nx, pn, un);
    phi_n[o] = pn;
    u_n[o] = un;
    return;
  }
  const int64_t t = (int64_t)(blockIdx.x - ntiles) * (blockDim.x * blockDim.y) +
                    threadIdx.y * blockDim.x + threadIdx.x;
  if (t >= N.nops) return;
  const lgd_fill_op op = N.ops[t];
  if (op.clen == 0) {
    const lgd_block_desc bd = blocks[op.cstart];
    int i, j;
    slab_coords(bd, op.src, i, j);
    double pn, un;
    modelc_cell_stage<kNoise>(k, dt, phi, u, op.src, bd.nx + 2, noise_scale, seed, st, op.cstart,
                              i + (int64_t)j * bd.nx, pn, un);
    phi_n[op.target] = pn;
    u_n[op.target] = un;
  } else {
    phi_n[op.target] = prolong_blend(N, op, 0);
    u_n[op.target] = prolong_blend(N, op, 1);
  }
}

// The streaming form of the fused Model C stage: one wavefront advances rows
// [j0, j1) of one column segment of one patch, lane L holding slab column x0 + L
// (interior column x0 + L - 1; lanes 1..w are the segment's cells, lane 0 and lane
// w + 1 the columns either side, which are ghost cells or another segment's
// cells).  Every operand comes from the patch's own ghost-filled slab.  The wave
// marches up the strip with rows j-1, j, j+1 of both fields in registers while the
// loads of the row kDepth ahead are in flight; each x-face flux is evaluated once,
// by the lane left of it, and reaches its right neighbour by a DPP shift, and each
// y-face flux once, as the top face of one row and the bottom face of the next.
// These are the operands, in the order, of the per-cell evaluation
// (src/discretization/finite_volume.rs:261-280; src/physics/phasefield.rs:395-421),
// so the results are bitwise those of amr_modelc_step_kernel.  With kFill, the
// first fill_wg workgroups write the next substep's ghosts as that kernel does.
constexpr int kSegWaves = 4;

struct SegArgs {
  const lgd_sseg *segs;
  int64_t nsegs;
  int32_t nwg_seg;  // workgroups of the streaming part
  int32_t fill_wg;  // workgroups of the next fill before it (a multiple of 8; 0: none)
  lgd_kr_consts k;
  double dt;
  const double *phi;
  const double *u;
  double *phi_n;
  double *u_n;
  double noise_scale;
  uint64_t seed;
  uint64_t salt;
  const double *salt_src;
};

// kFill: 0 none, 1 the next substep's ghosts (N), 2 the child level's first fill (C),
// 3 the restriction into the parent and its next fill (Q).
template <bool kNoise, int kFill, int kDepth>
__global__ __launch_bounds__(64 * kSegWaves) void amr_stream_kernel(
    SegArgs a, const lgd_block_desc *__restrict__ blocks, NextFill N, ChildFill C, ParentStep Q) {
  using namespace lgd_stream;
  // The fill's workgroups come first, so that they run beside the streaming waves
  // rather than after them (their copy ghosts each recompute a whole cell).
  if (kFill == 2 && (int)blockIdx.x < a.fill_wg) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= C.nops) return;
    const lgd_fill_op op = C.ops[t];
    if (op.clen == 0) {  // a copy within the child level
      C.lvl[0][op.target] = C.lvl[0][op.src];
      C.lvl[1][op.target] = C.lvl[1][op.src];
      return;
    }
    // old·(1 − 0) + new·0 is the old value bit for bit whenever the new one is
    // finite: the corner sum starts from +0, so it is never −0, and adding ±0 to it
    // changes nothing.  The stage being computed is not read.
    lgd_corner k[4];
    load_corners(C.corners, t, k);
    for (int f = 0; f < 2; ++f) {
      const double *old = f == 0 ? a.phi : a.u;
      double vo = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c < op.clen) vo = vo + old[k[c].src] * k[c].weight;
      C.lvl[f][op.target] = vo;
    }
    return;
  }
  if (kFill == 3 && (int)blockIdx.x < a.fill_wg) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= Q.nops) return;
    lgd_corner kk[4];
    load_corners(Q.corners, t, kk);
    const lgd_fill_op op = Q.ops[t];
    if (op.clen != 0) {
      for (int f = 0; f < 2; ++f)
        Q.coarse[f][op.target] =
            prolong_value<true>(kk, op.clen, Q.par[f], Q.old[f], Q.one_minus_alpha, Q.alpha);
      return;
    }
    const lgd_copy_child cc = Q.cchild[t];
    if (cc.child < 0) {
      Q.coarse[0][op.target] = Q.coarse[0][op.src];
      Q.coarse[1][op.target] = Q.coarse[1][op.src];
      return;
    }
    // A source cell under this level: its children's new values, recomputed from
    // the stage's inputs, restricted (children row by row, from 0) · 1/4.
    const uint64_t st = kNoise ? lgd_salt(a.salt, a.salt_src) : 0;
    double sp = 0.0, su = 0.0;
#pragma unroll 1
    for (int c = 0; c < 4; ++c) {  // (kx, ky) = (c & 1, c >> 1)
      const int64_t o = cc.child + (c >> 1) * cc.stride + (c & 1);
      const uint32_t b = patch_of(blocks, Q.b0, Q.nb, o);
      const lgd_block_desc bd = blocks[b];
      int i, j;
      slab_coords(bd, o, i, j);
      double pn, un;
      modelc_cell_stage<kNoise>(a.k, a.dt, a.phi, a.u, o, bd.nx + 2, a.noise_scale, a.seed, st, b,
                                i + (int64_t)j * bd.nx, pn, un);
      sp = sp + pn;
      su = su + un;
    }
    Q.coarse[0][op.target] = sp * Q.inv;
    Q.coarse[1][op.target] = su * Q.inv;
    return;
  }
  if (kFill == 1 && (int)blockIdx.x < a.fill_wg) {
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= N.nops) return;
    const uint64_t st = kNoise ? lgd_salt(a.salt, a.salt_src) : 0;
    const lgd_fill_op op = N.ops[t];
    if (op.clen == 0) {
      const lgd_block_desc bd = blocks[op.cstart];
      int i, j;
      slab_coords(bd, op.src, i, j);
      double pn, un;
      modelc_cell_stage<kNoise>(a.k, a.dt, a.phi, a.u, op.src, bd.nx + 2, a.noise_scale, a.seed,
                                st, op.cstart, i + (int64_t)j * bd.nx, pn, un);
      a.phi_n[op.target] = pn;
      a.u_n[op.target] = un;
    } else {
      a.phi_n[op.target] = prolong_blend(N, op, 0);
      a.u_n[op.target] = prolong_blend(N, op, 1);
    }
    return;
  }
  // XCD-aware order: XCD x (= blockIdx.x % 8, also wg % 8) takes one contiguous range
  // of items, so the strips of a segment, which share their halo rows, meet in one L2.
  const int wg = (int)blockIdx.x - (kFill != 0 ? a.fill_wg : 0);
  const int xcd = wg % 8, idx = wg / 8;
  const int per = a.nwg_seg / 8, rem = a.nwg_seg % 8;
  const int lw = xcd * per + min(xcd, rem) + idx;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int64_t item = (int64_t)lw * kSegWaves + wave;
  if (item >= a.nsegs) return;  // whole wave exits together
  const lgd_sseg S = load_uniform(a.segs + item);
  const lgd_kr_consts &k = a.k;
  const int64_t sx = (int64_t)S.nx + 2;
  // Lanes past w + 1 re-read column w + 1: never stored, never a stored cell's
  // neighbour (w <= 62, so lane 63 is at most column w + 1).
  const int64_t cbase = S.off + S.x0 + min(lane, S.w + 1);
  const bool in = lane >= 1 && lane <= S.w;
  const int j0 = S.j0, j1 = S.j1;
  auto ld = [&](const double *f, int r) { return f[cbase + (int64_t)(r + 1) * sx]; };

  uint64_t block_key = 0;
  if (kNoise)
    block_key = lgd_mix_key3(a.seed, lgd_salt(a.salt, a.salt_src), (1ull << 32) | 0ull,
                             (uint64_t)S.block);
  const uint64_t kc = (uint64_t)(S.x0 + lane - 1);  // interior column (cell key)

  double pm = ld(a.phi, j0 - 1), pc = ld(a.phi, j0), pp = ld(a.phi, j0 + 1);
  double um = ld(a.u, j0 - 1), uc = ld(a.u, j0), up = ld(a.u, j0 + 1);
  const double q = 0.25;
  double pmR = from_right(pm);
  double pcL = from_left(pc), pcR = from_right(pc);
  double ppL = from_left(pp), ppR = from_right(pp);
  double ucL = from_left(uc), ucR = from_right(uc);
  // Flux through the bottom face of row j0 (the top face of row j0 - 1).
  double jb = flux_y(k, q * (((pcR - pcL) + pmR) - from_left(pm)), pc - pm);

  // At row j the loads of row j + 1 + kDepth are issued; the row that becomes the
  // top neighbour of row j + 1 is moved into place after row j is computed (two
  // alternating buffers for kDepth 2, so no register with a load in flight is
  // copied).
  double rp = 0.0, ru = 0.0;  // kFill 3: the even row's new values
  auto body = [&](int j, double &ip, double &iu, double &cp, double &cu) {
    const int rn = min(j + 1 + kDepth, j1);
    ip = ld(a.phi, rn);
    iu = ld(a.u, rn);
    const double jx = flux_x(k, pcR - pc, q * (((ppR - pmR) + pp) - pm));
    const double jxl = from_left(jx);
    const double jt = flux_y(k, q * (((ppR - ppL) + pcR) - pcL), pp - pc);
    const double f = (((jx - jxl) + jt) - jb) * k.inv_h2;
    double n, v, d, e;
    cell_update(k, f, pc, pcR, pcL, pp, pm, uc, ucR, ucL, up, um, a.dt, n, v, d, e);
    if (kNoise && in)
      n = n + a.noise_scale *
                  lgd_standard_normal(lgd_splitmix64(block_key ^ ((uint64_t)j * (uint64_t)S.nx + kc)));
    const int64_t o = cbase + (int64_t)(j + 1) * sx;
    if constexpr (kFill == 3) {
      // Rows j − 1 (even) and j of a 2×2 block: the even column's lane restricts the
      // block (children row by row, from 0) into its parent cell.
      if (j & 1) {
        const double rpR = from_right(rp), ruR = from_right(ru), nR = from_right(n), vR = from_right(v);
        const int64_t ci = (int64_t)S.x0 + lane - 1;
        if (in && (ci & 1) == 0) {
          double sp = 0.0, su = 0.0;
          sp = sp + rp;
          sp = sp + rpR;
          sp = sp + n;
          sp = sp + nR;
          su = su + ru;
          su = su + ruR;
          su = su + v;
          su = su + vR;
          const int64_t t = Q.rit0[S.block - (int32_t)Q.b0] + (int64_t)(j >> 1) * (S.nx >> 1) + (ci >> 1);
          const int64_t tg = Q.items[t].target;
          Q.coarse[0][tg] = sp * Q.inv;
          Q.coarse[1][tg] = su * Q.inv;
        }
      } else {
        rp = n;
        ru = v;
      }
    }
    jb = jt;
    if (j + 1 < j1) {
      pm = pc; pmR = pcR;
      pc = pp; pcL = ppL; pcR = ppR;
      pp = cp;
      ppL = from_left(pp);
      ppR = from_right(pp);
      um = uc;
      uc = up;
      up = cu;
      ucL = from_left(uc);
      ucR = from_right(uc);
    }
    if (in) {
      a.phi_n[o] = n;
      a.u_n[o] = v;
    }
  };
  double ap = 0.0, au = 0.0, bp = 0.0, bu = 0.0;
  if (kDepth == 2) {
    ap = ld(a.phi, min(j0 + 2, j1));
    au = ld(a.u, min(j0 + 2, j1));
  }
  for (int j = j0; j < j1; j += 2) {
    if (kDepth == 2) {
      body(j, bp, bu, ap, au);
      if (j + 1 < j1) body(j + 1, ap, au, bp, bu);
    } else {
      body(j, ap, au, ap, au);
      if (j + 1 < j1) body(j + 1, ap, au, ap, au);
    }
  }
}

// One fused Heat stage of a level, with the next substep's ghosts as above.
template <bool kFill>
__global__ __launch_bounds__(kTX *kTY) void amr_heat_step_kernel(
    const lgd_tile *__restrict__ tiles, int64_t ntiles, const lgd_block_desc *__restrict__ blocks,
    double ihx, double ihy, double kappa, double dt, const double *u, double *u_n, NextFill N) {
  if (!kFill || blockIdx.x < ntiles) {
    const lgd_tile tl = tiles[blockIdx.x];
    const lgd_block_desc bd = blocks[tl.block];
    const int i = tl.x0 + threadIdx.x, j = tl.y0 + threadIdx.y;
    if (i >= bd.nx || j >= bd.ny) return;
    const int64_t sx = bd.nx + 2;
    const int64_t o = bd.off + (int64_t)(j + 1) * sx + (i + 1);
    u_n[o] = heat_cell_stage(ihx, ihy, kappa, dt, u, o, sx);
    return;
  }
  const int64_t t = (int64_t)(blockIdx.x - ntiles) * (blockDim.x * blockDim.y) +
                    threadIdx.y * blockDim.x + threadIdx.x;
  if (t >= N.nops) return;
  const lgd_fill_op op = N.ops[t];
  if (op.clen == 0) {
    const lgd_block_desc bd = blocks[op.cstart];
    u_n[op.target] = heat_cell_stage(ihx, ihy, kappa, dt, u, op.src, bd.nx + 2);
  } else {
    u_n[op.target] = prolong_blend(N, op, 0);
  }
}

}  // namespace

int lgd_amr_fill(lgd_ctx *ctx, const lgd_amr &g, double *field, const double *old, double alpha,
                 int level) {
  const lgd_level_plan &P = g.plans[(size_t)level];
  if (P.nops == 0) return LGD_OK;
  unsigned nb = blocks_for(P.nops, 256);
  if (old)
    amr_fill_kernel<true><<<nb, 256, 0, ctx->stream>>>(P.ops, P.nops, P.corners, field, old,
                                                      1.0 - alpha, alpha);
  else
    amr_fill_kernel<false><<<nb, 256, 0, ctx->stream>>>(P.ops, P.nops, P.corners, field, nullptr,
                                                       0.0, 0.0);
  return launch_check(ctx, "amr_fill");
}

int lgd_amr_restrict(lgd_ctx *ctx, const lgd_amr &g, double *field, int fine_level) {
  const lgd_level_plan &P = g.plans[(size_t)fine_level];
  if (P.nritems == 0) return LGD_OK;
  int r = (int)g.ratios[fine_level - 1];
  double inv = 1.0 / std::pow((double)r, 2);
  amr_restrict_kernel<<<blocks_for(P.nritems, 256), 256, 0, ctx->stream>>>(P.ritems, P.nritems, r,
                                                                         inv, field);
  return launch_check(ctx, "amr_restrict");
}

int lgd_amr_fill_pp(lgd_ctx *ctx, const lgd_amr &g, int level, int nf, double *const *lvl,
                    const double *const *par, const double *const *old, double alpha) {
  const lgd_level_plan &P = g.plans[(size_t)level];
  if (P.nops == 0) return LGD_OK;
  FieldPtrs F{};
  for (int f = 0; f < nf; ++f) {
    F.lvl[f] = lvl[f];
    F.par[f] = par ? par[f] : nullptr;
    F.old[f] = old ? old[f] : nullptr;
  }
  const int wg = lgd_amr_wg(ctx);
  const unsigned nb = blocks_for(P.nops, wg);
  if (old)
    amr_fill_pp_kernel<true><<<nb, wg, 0, ctx->stream>>>(P.ops, P.nops, P.corners, F, nf,
                                                         1.0 - alpha, alpha);
  else
    amr_fill_pp_kernel<false><<<nb, wg, 0, ctx->stream>>>(P.ops, P.nops, P.corners, F, nf, 0.0,
                                                          0.0);
  return launch_check(ctx, "amr_fill");
}

int lgd_amr_restrict_pp(lgd_ctx *ctx, const lgd_amr &g, int fine_level, int nf,
                        const double *const *fine, double *const *coarse) {
  const lgd_level_plan &P = g.plans[(size_t)fine_level];
  if (P.nritems == 0) return LGD_OK;
  const int r = (int)g.ratios[(size_t)fine_level - 1];
  RestrictPtrs R{};
  for (int f = 0; f < nf; ++f) {
    R.fine[f] = fine[f];
    R.coarse[f] = coarse[f];
  }
  const int wg = lgd_amr_wg(ctx);
  amr_restrict_pp_kernel<<<blocks_for(P.nritems, wg), wg, 0, ctx->stream>>>(
      P.ritems, P.nritems, r, 1.0 / (double)(r * r), R, nf);
  return launch_check(ctx, "amr_restrict");
}

int lgd_amr_restrict_fill_pp(lgd_ctx *ctx, const lgd_amr &g, int level, int nf, double *const *lvl,
                             const double *const *par, const double *const *old, double alpha,
                             const double *const *fine) {
  const lgd_level_plan &P = g.plans[(size_t)level];
  const lgd_level_plan &Q = g.plans[(size_t)level + 1];
  if (P.nops > 0 && !P.cchild)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "restrict_fill: the level has no finer level");
  const int64_t total = Q.nritems + P.nops;
  if (total == 0) return LGD_OK;
  FieldPtrs F{};
  RestrictPtrs R{};
  for (int f = 0; f < nf; ++f) {
    F.lvl[f] = lvl[f];
    F.par[f] = par ? par[f] : nullptr;
    F.old[f] = old ? old[f] : nullptr;
    R.fine[f] = fine[f];
    R.coarse[f] = lvl[f];
  }
  const int r = (int)g.ratios[(size_t)level];
  const double inv = 1.0 / (double)(r * r);
  const int wg = lgd_amr_wg(ctx);
  const unsigned nb = blocks_for(total, wg);
  if (old)
    amr_restrict_fill_pp_kernel<true><<<nb, wg, 0, ctx->stream>>>(
        Q.ritems, Q.nritems, r, inv, P.ops, P.nops, P.corners, P.cchild, F, R, nf, 1.0 - alpha, alpha);
  else
    amr_restrict_fill_pp_kernel<false><<<nb, wg, 0, ctx->stream>>>(
        Q.ritems, Q.nritems, r, inv, P.ops, P.nops, P.corners, P.cchild, F, R, nf, 0.0, 0.0);
  return launch_check(ctx, "amr_restrict_fill");
}

bool lgd_amr_level0_streams(const lgd_ctx *ctx, const lgd_amr &g) {
  return ctx->modelc_kernel != 1 && g.base.n[0] + 2 <= 128;
}

lgd_launch_shape lgd_amr_level0_default_shape(const lgd_amr &g, bool noise) {
  lgd_launch_shape s = lgd_modelc_default_shape(noise);
  if (lgd_stream_pairable(g.base.n[0], g.base.nb[0])) s.kernel = 3;
  return s;
}

bool lgd_amr_level_streams(const lgd_ctx *ctx, const lgd_amr &g, int level) {
  if (ctx->amr_kernel == 1) return false;
  return !(level == 0 && lgd_amr_level0_streams(ctx, g));
}

bool lgd_amr_restrict_fuses(const lgd_ctx *ctx, const lgd_amr &g, int level) {
  if (level < 1 || level >= g.num_levels() || !lgd_amr_level_streams(ctx, g, level)) return false;
  if (ctx->amr_per_block || ctx->amr_split_fill || g.ratios[(size_t)level - 1] != 2) return false;
  const lgd_level_plan &P = g.plans[(size_t)level];
  const int rows = ctx->amr_rows > 0 ? (int)ctx->amr_rows : P.auto_rows;
  return rows % 2 == 0;
}

bool lgd_amr_child_fill_fuses(const lgd_ctx *ctx, const lgd_amr &g, int level) {
  const bool streams = lgd_amr_level_streams(ctx, g, level) || (level == 0 && lgd_amr_level0_streams(ctx, g));
  return level + 1 < g.num_levels() && streams && !ctx->amr_per_block && !ctx->amr_split_fill;
}

namespace {

constexpr int kNumTileShapes = (int)(sizeof(kAmrTileShapes) / sizeof(kAmrTileShapes[0]));

struct StageTiles {
  const lgd_tile *tiles = nullptr;
  int64_t n = 0;
  const std::vector<int64_t> *tile0 = nullptr;
  dim3 blk{kTX, kTY};
};

// The tiles the stage kernels of `level` run over in the context's shape: the plan's
// own, or a retiling of the level's patches kept with the plan until the shape
// changes.
int stage_tiles(lgd_ctx *ctx, const lgd_amr &g, int level, StageTiles &T) {
  const lgd_level_plan &P = g.plans[(size_t)level];
  const int sh = (int)ctx->amr_tile;
  if (sh <= 0 || sh >= kNumTileShapes) {
    T.tiles = P.tiles;
    T.n = P.ntiles;
    T.tile0 = &P.tile0;
    T.blk = dim3(kTX, kTY);
    return LGD_OK;
  }
  const int tx = kAmrTileShapes[sh][0], ty = kAmrTileShapes[sh][1];
  if (P.stx != tx || P.sty != ty) {
    std::vector<lgd_tile> tiles;
    P.stile0.clear();
    for (uint32_t b = g.level_start[level]; b < g.level_start[level + 1]; ++b) {
      P.stile0.push_back((int64_t)tiles.size());
      for (int64_t y0 = 0; y0 < g.box[b].extent(1); y0 += ty)
        for (int64_t x0 = 0; x0 < g.box[b].extent(0); x0 += tx)
          tiles.push_back(lgd_tile{(int32_t)b, (int32_t)x0, (int32_t)y0, 0});
    }
    P.stile0.push_back((int64_t)tiles.size());
    if (P.stiles) (void)hipFreeAsync(P.stiles, ctx->stream);
    P.stiles = nullptr;
    P.nstiles = 0;
    P.stx = P.sty = 0;
    if (!tiles.empty()) {
      hipError_t e = hipMallocAsync((void **)&P.stiles, tiles.size() * sizeof(lgd_tile), ctx->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(P.stiles, tiles.data(), tiles.size() * sizeof(lgd_tile),
                           hipMemcpyHostToDevice, ctx->stream);
      if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);  // `tiles` leaves scope
      if (e != hipSuccess) return hip_fail(ctx, e, "amr stage tiles");
    }
    P.nstiles = (int64_t)tiles.size();
    P.stx = tx;
    P.sty = ty;
  }
  T.tiles = P.stiles;
  T.n = P.nstiles;
  T.tile0 = &P.stile0;
  T.blk = dim3(tx, ty);
  return LGD_OK;
}

struct StageSegs {
  const lgd_sseg *segs = nullptr;
  int64_t n = 0;
  const std::vector<int64_t> *seg0 = nullptr;
};

// The streaming level stage's items of `level` at the context's strip height: the
// plan's own, or a recut kept with the plan until the height changes.
int stage_segs(lgd_ctx *ctx, const lgd_amr &g, int level, StageSegs &T) {
  const lgd_level_plan &P = g.plans[(size_t)level];
  const int rows = ctx->amr_rows > 0 ? (int)ctx->amr_rows : P.auto_rows;
  if (rows == P.seg_rows) {
    T.segs = P.segs;
    T.n = P.nsegs;
    T.seg0 = &P.seg0;
    return LGD_OK;
  }
  if (P.trows != rows) {
    // The previous recut's upload may still read its host copy.
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return hip_fail(ctx, e, "amr stage items");
    build_segs(g, level, rows, P.tsegs_host, P.tseg0);
    if (P.tsegs) (void)hipFreeAsync(P.tsegs, ctx->stream);
    P.tsegs = nullptr;
    P.ntsegs = 0;
    P.trows = 0;
    if (!P.tsegs_host.empty()) {
      const size_t bytes = P.tsegs_host.size() * sizeof(lgd_sseg);
      e = hipMallocAsync((void **)&P.tsegs, bytes, ctx->stream);
      if (e == hipSuccess)
        e = hipMemcpyAsync(P.tsegs, P.tsegs_host.data(), bytes, hipMemcpyHostToDevice, ctx->stream);
      if (e != hipSuccess) return hip_fail(ctx, e, "amr stage items");
    }
    P.ntsegs = (int64_t)P.tsegs_host.size();
    P.trows = rows;
  }
  T.segs = P.tsegs;
  T.n = P.ntsegs;
  T.seg0 = &P.tseg0;
  return LGD_OK;
}

// The next-fill arguments of a stage launch (nops = 0: none).
NextFill next_fill_args(const lgd_level_plan &P, const lgd_amr_next_fill *nf, int nfields) {
  NextFill N{};
  if (!nf) return N;
  N.ops = P.ops;
  N.corners = P.corners;
  N.nops = P.nops;
  for (int f = 0; f < nfields; ++f) {
    N.par[f] = nf->par[f];
    N.old[f] = nf->old[f];
  }
  N.one_minus_alpha = 1.0 - nf->alpha;
  N.alpha = nf->alpha;
  return N;
}

}  // namespace

int lgd_amr_modelc_step(lgd_ctx *ctx, const lgd_amr &g, const lgd_modelc_params &p, int level,
                        double dt, const double *phi, const double *u, double *phi_n, double *u_n,
                        uint64_t seed, uint64_t salt, bool wiener, const lgd_amr_next_fill *nf,
                        const lgd_amr_child_fill *cf, const lgd_amr_parent_step *ps) {
  const lgd_level_plan &P = g.plans[(size_t)level];
  if (cf && (nf || ps || !lgd_amr_child_fill_fuses(ctx, g, level)))
    return fail(ctx, LGD_ERR_UNSUPPORTED, "the child level's fill shares only a streamed stage launch");
  if (ps && (nf || !lgd_amr_restrict_fuses(ctx, g, level)))
    return fail(ctx, LGD_ERR_UNSUPPORTED, "the restriction shares only a streamed ratio-2 stage launch");
  lgd_cart c = g.base;
  c.h[0] = g.spacing[(size_t)level][0];
  c.h[1] = g.spacing[(size_t)level][1];
  const lgd_kr_consts k = lgd_make_kr_consts(p, c);
  const bool noise = wiener && p.noise_amplitude != 0.0;
  const double scale = std::sqrt(dt) * p.noise_amplitude;
  if (level == 0 && lgd_amr_level0_streams(ctx, g)) {
    // The base tiling is a Cartesian grid with mirror faces: the streaming kernel
    // reads neighbour interiors itself (no ghost fill) — the same values and order.
    const int64_t nb0 = g.level_start[1];
    const lgd_launch_shape shape = lgd_modelc_shape(ctx, g.base.n[0], g.base.n[1], nb0, noise,
                                                    lgd_amr_level0_default_shape(g, noise));
    const bool pair = shape.kernel == 3 && lgd_stream_pairable(g.base.n[0], g.base.nb[0]);
    lgd_domain d0;
    d0.k = k;
    d0.noise_scale = noise ? scale : 0.0;
    d0.seed = seed;
    d0.block0 = 0;
    d0.nbx = (int32_t)g.base.nb[0];
    d0.nby = (int32_t)g.base.nb[1];
    d0.per_x = 0;
    d0.per_y = 0;
    const uint64_t wiener0 = (1ull << 32) | 0ull;
    if (ctx->amr_per_block) {
      for (int64_t b = 0; b < nb0; ++b) {
        int rc = lgd_modelc_stream_launch(ctx, g.base.n[0], g.base.n[1], d0, nullptr, nullptr, b, 1,
                                          phi, u, phi_n, u_n, dt, noise, salt, wiener0,
                                          (int)g.base.n[1], shape.prefetch);
        if (rc) return rc;
      }
      return LGD_OK;
    }
    lgd_stream_side_fill side{};
    if (cf) {
      const lgd_level_plan &FP = g.plans[1];
      side.ops = FP.ops;
      side.corners = FP.corners;
      side.nops = FP.nops;
      side.lvl[0] = cf->lvl[0];
      side.lvl[1] = cf->lvl[1];
    }
    return lgd_modelc_stream_launch(ctx, g.base.n[0], g.base.n[1], d0, nullptr, nullptr, 0, nb0, phi,
                                    u, phi_n, u_n, dt, noise, salt, wiener0, shape.strips,
                                    shape.prefetch, pair, cf ? &side : nullptr);
  }
  int rc;
  if (lgd_amr_level_streams(ctx, g, level)) {
    StageSegs T;
    if ((rc = stage_segs(ctx, g, level, T))) return rc;
    // The fused work: 1 the next substep's ghosts, 2 the finer level's first fill,
    // 3 the restriction into the parent (and its next fill); fops: the fill ops.
    int mode = 0;
    int64_t fops = 0;
    if (ps) {
      mode = 3;
      fops = ps->fill ? g.plans[(size_t)level - 1].nops : 0;
    } else if (cf) {
      fops = g.plans[(size_t)level + 1].nops;
      mode = fops ? 2 : 0;
    } else if (nf && P.nops) {
      mode = 1;
      fops = P.nops;
    }
    const bool fill = mode != 0;
    if (T.n == 0 && fops == 0) return LGD_OK;
    if (ctx->amr_per_block && nf)
      return fail(ctx, LGD_ERR_UNSUPPORTED, "per-patch launches fill ghosts in a launch of their own");
    SegArgs A{};
    A.k = k;
    A.dt = dt;
    A.phi = phi;
    A.u = u;
    A.phi_n = phi_n;
    A.u_n = u_n;
    A.noise_scale = noise ? scale : 0.0;
    A.seed = seed;
    A.salt = salt;
    A.salt_src = noise ? ctx->salt_src : nullptr;
    const NextFill N = mode == 1 ? next_fill_args(P, nf, 2) : NextFill{};
    ChildFill C{};
    if (mode == 2) {
      const lgd_level_plan &FP = g.plans[(size_t)level + 1];
      C.ops = FP.ops;
      C.corners = FP.corners;
      C.nops = FP.nops;
      C.lvl[0] = cf->lvl[0];
      C.lvl[1] = cf->lvl[1];
    }
    ParentStep Q{};
    if (mode == 3) {
      const lgd_level_plan &PP = g.plans[(size_t)level - 1];
      Q.coarse[0] = ps->coarse[0];
      Q.coarse[1] = ps->coarse[1];
      Q.items = P.ritems;
      Q.rit0 = P.rit0;
      Q.inv = 1.0 / (double)(2 * 2);
      if (ps->fill) {
        Q.ops = PP.ops;
        Q.corners = PP.corners;
        Q.cchild = PP.cchild;
        Q.nops = PP.nops;
        for (int f = 0; f < 2; ++f) {
          Q.par[f] = ps->par[f];
          Q.old[f] = ps->old[f];
        }
        Q.one_minus_alpha = 1.0 - ps->alpha;
        Q.alpha = ps->alpha;
      }
      Q.b0 = (uint32_t)g.level_start[(size_t)level];
      Q.nb = (uint32_t)(g.level_start[(size_t)level + 1] - g.level_start[(size_t)level]);
    }
    auto go = [&](const lgd_sseg *segs, int64_t n, bool with_fill) -> int {
      const int64_t nwg = (n + kSegWaves - 1) / kSegWaves;
      const int64_t fwg = with_fill ? ((fops + 64 * kSegWaves - 1) / (64 * kSegWaves) + 7) / 8 * 8 : 0;
      const int64_t total = nwg + fwg;
      if (total == 0) return LGD_OK;
      if (total > INT32_MAX) return fail(ctx, LGD_ERR_UNSUPPORTED, "AMR level too large for one launch");
      A.segs = segs;
      A.nsegs = n;
      A.nwg_seg = (int32_t)nwg;
      A.fill_wg = (int32_t)fwg;
      const dim3 grid((unsigned)total), blk(64 * kSegWaves);
      hipStream_t st = ctx->stream;
      const int m = with_fill ? mode : 0;
      if (noise) {
        if (m == 1) amr_stream_kernel<true, 1, 1><<<grid, blk, 0, st>>>(A, g.d_blocks, N, C, Q);
        else if (m == 2) amr_stream_kernel<true, 2, 1><<<grid, blk, 0, st>>>(A, g.d_blocks, N, C, Q);
        else if (m == 3) amr_stream_kernel<true, 3, 1><<<grid, blk, 0, st>>>(A, g.d_blocks, N, C, Q);
        else amr_stream_kernel<true, 0, 1><<<grid, blk, 0, st>>>(A, g.d_blocks, N, C, Q);
      } else {
        if (m == 1) amr_stream_kernel<false, 1, 1><<<grid, blk, 0, st>>>(A, g.d_blocks, N, C, Q);
        else if (m == 2) amr_stream_kernel<false, 2, 1><<<grid, blk, 0, st>>>(A, g.d_blocks, N, C, Q);
        else if (m == 3) amr_stream_kernel<false, 3, 1><<<grid, blk, 0, st>>>(A, g.d_blocks, N, C, Q);
        else amr_stream_kernel<false, 0, 1><<<grid, blk, 0, st>>>(A, g.d_blocks, N, C, Q);
      }
      return LGD_OK;
    };
    if (ctx->amr_per_block) {
      for (size_t b = 0; b + 1 < T.seg0->size(); ++b)
        if ((rc = go(T.segs + (*T.seg0)[b], (*T.seg0)[b + 1] - (*T.seg0)[b], false))) return rc;
    } else if ((rc = go(T.segs, T.n, fill))) {
      return rc;
    }
    return launch_check(ctx, "amr_stream");
  }
  StageTiles T;
  if ((rc = stage_tiles(ctx, g, level, T))) return rc;
  if (T.n == 0 && !(nf && P.nops)) return LGD_OK;
  const NextFill none{};
  auto launch = [&](const lgd_tile *tiles, int64_t n) {
    if (noise)
      amr_modelc_step_kernel<true, false><<<(unsigned)n, T.blk, 0, ctx->stream>>>(
          tiles, n, g.d_blocks, k, dt, phi, u, phi_n, u_n, scale, seed, salt, ctx->salt_src, none);
    else
      amr_modelc_step_kernel<false, false><<<(unsigned)n, T.blk, 0, ctx->stream>>>(
          tiles, n, g.d_blocks, k, dt, phi, u, phi_n, u_n, 0.0, seed, salt, nullptr, none);
  };
  if (ctx->amr_per_block && nf)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "per-patch launches fill ghosts in a launch of their own");
  if (ctx->amr_per_block) {
    for (size_t b = 0; b + 1 < T.tile0->size(); ++b)
      launch(T.tiles + (*T.tile0)[b], (*T.tile0)[b + 1] - (*T.tile0)[b]);
  } else if (nf && P.nops) {
    const NextFill N = next_fill_args(P, nf, 2);
    const int64_t per = (int64_t)T.blk.x * T.blk.y;
    const unsigned nwg = (unsigned)(T.n + (P.nops + per - 1) / per);
    if (noise)
      amr_modelc_step_kernel<true, true><<<nwg, T.blk, 0, ctx->stream>>>(
          T.tiles, T.n, g.d_blocks, k, dt, phi, u, phi_n, u_n, scale, seed, salt, ctx->salt_src, N);
    else
      amr_modelc_step_kernel<false, true><<<nwg, T.blk, 0, ctx->stream>>>(
          T.tiles, T.n, g.d_blocks, k, dt, phi, u, phi_n, u_n, 0.0, seed, salt, nullptr, N);
  } else {
    launch(T.tiles, T.n);
  }
  return launch_check(ctx, "amr_modelc_step");
}

int lgd_amr_heat_step(lgd_ctx *ctx, const lgd_amr &g, double kappa, int level, double dt,
                      const double *u, double *u_n, const lgd_amr_next_fill *nf) {
  const lgd_level_plan &P = g.plans[(size_t)level];
  StageTiles T;
  int rc;
  if ((rc = stage_tiles(ctx, g, level, T))) return rc;
  if (T.n == 0 && !(nf && P.nops)) return LGD_OK;
  const double hx = g.spacing[(size_t)level][0], hy = g.spacing[(size_t)level][1];
  const double ihx = 1.0 / (hx * hx), ihy = 1.0 / (hy * hy);
  const NextFill none{};
  auto launch = [&](const lgd_tile *tiles, int64_t n) {
    amr_heat_step_kernel<false><<<(unsigned)n, T.blk, 0, ctx->stream>>>(
        tiles, n, g.d_blocks, ihx, ihy, kappa, dt, u, u_n, none);
  };
  if (ctx->amr_per_block && nf)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "per-patch launches fill ghosts in a launch of their own");
  if (ctx->amr_per_block) {
    for (size_t b = 0; b + 1 < T.tile0->size(); ++b)
      launch(T.tiles + (*T.tile0)[b], (*T.tile0)[b + 1] - (*T.tile0)[b]);
  } else if (nf && P.nops) {
    const NextFill N = next_fill_args(P, nf, 1);
    const int64_t per = (int64_t)T.blk.x * T.blk.y;
    const unsigned nwg = (unsigned)(T.n + (P.nops + per - 1) / per);
    amr_heat_step_kernel<true><<<nwg, T.blk, 0, ctx->stream>>>(
        T.tiles, T.n, g.d_blocks, ihx, ihy, kappa, dt, u, u_n, N);
  } else {
    launch(T.tiles, T.n);
  }
  return launch_check(ctx, "amr_heat_step");
}

int lgd_amr_copy_level(lgd_ctx *ctx, const lgd_amr &g, int level, const double *src, double *dst) {
  const int64_t lo = g.off[g.level_start[level]], hi = g.off[g.level_start[level + 1]];
  if (hi <= lo) return LGD_OK;
  hipError_t e = hipMemcpyAsync(dst + lo, src + lo, (size_t)(hi - lo) * sizeof(double),
                                hipMemcpyDeviceToDevice, ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "amr level copy");
  return LGD_OK;
}

// ------------------------------------------------------------------- C ABI
extern "C" {

int lgd_amr_create(lgd_ctx *ctx, const lgd_cartesian_desc *base, const uint32_t *ratios,
                   size_t nratios, const int64_t *boxes, size_t nboxes, lgd_amr **out) {
  if (!ctx || !base || !out || (!ratios && nratios) || (!boxes && nboxes))
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  lgd_grid *bg = nullptr;
  int rc = lgd_grid_create_cartesian(ctx, base, &bg);
  if (rc) return rc;
  lgd_cart c = bg->cart;
  lgd_grid_destroy(bg);
  std::vector<uint32_t> r(ratios, ratios + nratios);
  std::vector<std::vector<lgd_box>> levels;
  for (size_t i = 0; i < nboxes; ++i) {
    int64_t lev = boxes[5 * i];
    if (lev < 1) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "patch levels start at 1");
    while ((int64_t)levels.size() < lev) levels.emplace_back();
    levels[(size_t)lev - 1].push_back(
        lgd_box{{boxes[5 * i + 1], boxes[5 * i + 2]}, {boxes[5 * i + 3], boxes[5 * i + 4]}});
  }
  lgd_amr *g = new (std::nothrow) lgd_amr();
  if (!g) return fail(ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  rc = lgd_amr_build(ctx, c, r, levels, *g);
  int32_t bad = 0;
  if (!rc) rc = lgd_amr_plan_error(ctx, *g, &bad);
  if (!rc) {
    hipError_t e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) rc = hip_fail(ctx, e, "amr plan status");
  }
  if (!rc && bad)
    rc = fail(ctx, LGD_ERR_INVALID_ARGUMENT,
              "a ghost cell without a donor (hierarchy not properly nested)");
  if (rc) {
    g->free_device();
    delete g;
    return rc;
  }
  *out = g;
  return LGD_OK;
}

int lgd_amr_destroy(lgd_amr *g) {
  if (!g) return LGD_OK;
  if (g->ctx) (void)hipStreamSynchronize(g->ctx->stream);
  g->free_device();
  delete g;
  return LGD_OK;
}

int64_t lgd_amr_num_blocks(const lgd_amr *g) { return g ? (int64_t)g->box.size() : -1; }
int32_t lgd_amr_num_levels(const lgd_amr *g) { return g ? g->num_levels() : -1; }

int lgd_amr_patch(const lgd_amr *g, int64_t block, int64_t *out5) {
  if (!g || !out5 || block < 0 || block >= (int64_t)g->box.size()) return LGD_ERR_INVALID_ARGUMENT;
  out5[0] = g->level[(size_t)block];
  out5[1] = g->box[(size_t)block].lo[0];
  out5[2] = g->box[(size_t)block].lo[1];
  out5[3] = g->box[(size_t)block].hi[0];
  out5[4] = g->box[(size_t)block].hi[1];
  return LGD_OK;
}

int64_t lgd_amr_field_len(const lgd_amr *g) { return g ? g->off.back() : -1; }

int64_t lgd_amr_block_offset(const lgd_amr *g, int64_t block) {
  if (!g || block < 0 || block > (int64_t)g->box.size()) return -1;
  return g->off[(size_t)block];
}

int64_t lgd_amr_find_patch(const lgd_amr *g, int32_t level, int64_t x, int64_t y) {
  return g ? g->find_patch(level, x, y) : -1;
}

int lgd_amr_restrict_level(lgd_ctx *ctx, const lgd_amr *g, lgd_buf *field, int32_t fine_level) {
  if (!ctx || !g || !field) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (fine_level < 1 || fine_level >= g->num_levels())
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "restriction needs an existing finer level >= 1");
  if ((int64_t)field->len != g->off.back())
    return fail(ctx, LGD_ERR_SIZE_MISMATCH, "field is not congruent with the hierarchy");
  return lgd_amr_restrict(ctx, *g, field->data, fine_level);
}

int lgd_amr_fill_level(lgd_ctx *ctx, const lgd_amr *g, lgd_buf *field, const lgd_buf *old,
                       double alpha, int32_t level) {
  if (!ctx || !g || !field) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  if (level < 0 || level >= g->num_levels())
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "level out of range");
  if ((int64_t)field->len != g->off.back() || (old && old->len != field->len))
    return fail(ctx, LGD_ERR_SIZE_MISMATCH, "field is not congruent with the hierarchy");
  return lgd_amr_fill(ctx, *g, field->data, old ? old->data : nullptr, alpha, level);
}

}  // extern "C"
