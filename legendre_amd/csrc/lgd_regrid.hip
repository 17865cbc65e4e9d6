// Berger–Oliger regrid for the device AMR backend (host work at regrid cadence).
//
// The policy (src/geometry/amr/adapt.rs:152-237): restrict + ghost-fill the
// tagged field on the device, bring it to the host, flag cells whose undivided
// central difference exceeds the threshold (GradientTagger, adapt.rs:78-103),
// dilate by the buffer, add nesting flags from the level being built above,
// clip/sort/dedupe, cluster with Berger–Rigoutsos (src/geometry/amr/cluster.rs),
// keep the old hierarchy if it still serves (hysteresis, adapt.rs:269-290), else
// build the new hierarchy and migrate the state (copy where the old level
// reaches, recursive bilinear sample elsewhere, adapt.rs:323-402).
// All of it is integer work plus exact-order sampling, so regrid decisions are
// the reference's bit for bit.
#include <algorithm>
#include <climits>
#include <cmath>
#include <thread>
#include <vector>

#include "lgd_regrid.h"

namespace {

// A flagged cell in level coordinates (level domains are far below 2^31 cells wide).
using Cell = std::array<int32_t, 2>;

// Grows b (empty: lo > hi) to hold cell c.
inline void grow(lgd_box &b, const Cell &c) {
  for (int d = 0; d < 2; ++d) {
    if (c[d] < b.lo[d]) b.lo[d] = c[d];
    if (c[d] + 1 > b.hi[d]) b.hi[d] = c[d] + 1;
  }
}

inline lgd_box empty_box() { return lgd_box{{INT64_MAX, INT64_MAX}, {INT64_MIN, INT64_MIN}}; }

struct Split {
  int axis;
  int64_t plane;
};

// The two signatures of a box: flagged cells per column (axis 0) and per row (axis 1).
struct Sig {
  const uint32_t *v[2];
  size_t n[2];
};

// Interior zero of a signature nearest the middle (first found wins ties).
bool hole_split(const lgd_box &bx, const Sig &sig, Split &s) {
  bool have = false;
  uint64_t best = 0;
  for (int d = 0; d < 2; ++d) {
    const size_t n = sig.n[d];
    for (size_t i = 1; i + 1 < n; ++i) {
      if (sig.v[d][i] != 0) continue;
      uint64_t c = std::min<uint64_t>(i, n - 1 - i);
      if (!have || c > best) {
        have = true;
        best = c;
        s = Split{d, bx.lo[d] + (int64_t)i};
      }
    }
  }
  return have;
}

// Strongest sign change of the signature's second difference leaving
// min_width cells on each side (stronger, then more central, wins).
bool inflection_split(const lgd_box &bx, const Sig &sig, size_t minw, Split &s) {
  bool have = false;
  int64_t best_s = 0;
  uint64_t best_c = 0;
  for (int d = 0; d < 2; ++d) {
    const size_t n = sig.n[d];
    const uint32_t *v = sig.v[d];
    if (n < 4) continue;
    for (size_t i = 1; i + 2 < n; ++i) {
      int64_t a = (int64_t)v[i + 1] - 2 * (int64_t)v[i] + (int64_t)v[i - 1];
      int64_t b = (int64_t)v[i + 2] - 2 * (int64_t)v[i + 1] + (int64_t)v[i];
      if (!(a * b < 0)) continue;
      size_t plane = i + 1;
      if (plane < minw || n - plane < minw) continue;
      int64_t strength = b - a < 0 ? a - b : b - a;
      uint64_t c = std::min<uint64_t>(plane, n - plane);
      if (!have || strength > best_s || (strength == best_s && c > best_c)) {
        have = true;
        best_s = strength;
        best_c = c;
        s = Split{d, bx.lo[d] + (int64_t)plane};
      }
    }
  }
  return have;
}

// Midpoint of the longest axis (the later axis on a tie).
bool bisect_split(const lgd_box &bx, size_t minw, Split &s) {
  int d = bx.extent(1) >= bx.extent(0) ? 1 : 0;
  if ((uint64_t)bx.extent(d) < 2 * (uint64_t)minw) return false;
  s = Split{d, bx.lo[d] + bx.extent(d) / 2};
  return true;
}

// Flag sets from this size up split their two halves over two host threads while
// `depth` allows (the halves are disjoint ranges of f; their boxes join left first,
// the order of the one-thread recursion).
constexpr size_t kParallelFlags = 4096;

// Berger–Rigoutsos over the n flagged cells f (in any order) whose bounding box is
// bx (cluster.rs:157-281): accept the box at the efficiency, else split at a
// signature hole, an inflection or the midpoint and recurse (left half first).  Each
// half's bounding box is gathered by the partition pass; `sig` is scratch.
void cluster_rec(Cell *f, size_t n, const lgd_box &bx, double eff, size_t minw,
                 std::vector<lgd_box> &out, int depth, std::vector<uint32_t> &sig) {
  if ((double)n / (double)bx.cells() >= eff) {
    out.push_back(bx);
    return;
  }
  const size_t nx = (size_t)bx.extent(0), ny = (size_t)bx.extent(1);
  sig.assign(nx + ny, 0u);
  uint32_t *sx = sig.data(), *sy = sig.data() + nx;
  const int64_t lx = bx.lo[0], ly = bx.lo[1];
  for (size_t i = 0; i < n; ++i) {
    sx[(size_t)(f[i][0] - lx)]++;
    sy[(size_t)(f[i][1] - ly)]++;
  }
  const Sig sg{{sx, sy}, {nx, ny}};
  Split s{0, 0};
  if (!hole_split(bx, sg, s) && !inflection_split(bx, sg, minw, s) && !bisect_split(bx, minw, s)) {
    out.push_back(bx);
    return;
  }
  // Two-pointer partition (cluster.rs:270-281): cells below the plane move to the front.
  lgd_box lb = empty_box(), rb = empty_box();
  size_t lt = 0;
  for (size_t j = 0; j < n; ++j) {
    const Cell c = f[j];
    if (c[s.axis] < s.plane) {
      grow(lb, c);
      f[j] = f[lt];
      f[lt++] = c;
    } else {
      grow(rb, c);
    }
  }
  if (depth > 0 && n >= kParallelFlags) {
    std::vector<lgd_box> left;
    std::thread t;
    bool spawned = false;
    try {
      t = std::thread([&] {
        std::vector<uint32_t> tsig;
        cluster_rec(f, lt, lb, eff, minw, left, depth - 1, tsig);
      });
      spawned = true;
    } catch (...) {
      cluster_rec(f, lt, lb, eff, minw, left, depth - 1, sig);
    }
    std::vector<lgd_box> right;
    cluster_rec(f + lt, n - lt, rb, eff, minw, right, depth - 1, sig);
    if (spawned) t.join();
    out.insert(out.end(), left.begin(), left.end());
    out.insert(out.end(), right.begin(), right.end());
    return;
  }
  cluster_rec(f, lt, lb, eff, minw, out, 0, sig);
  cluster_rec(f + lt, n - lt, rb, eff, minw, out, 0, sig);
}

// Whether every cell of `need` lies in a patch of level l of g (run-skipping walk
// along axis 0 through the level's bucket index).
bool box_covered(const lgd_amr &g, int l, const lgd_box &need) {
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

inline int64_t floor_div(int64_t a, int64_t b) {
  int64_t q = a / b;
  return (a % b < 0) ? q - 1 : q;
}

inline int64_t mirror_coord(int64_t c, int64_t lo, int64_t hi) {
  if (c < lo) return 2 * lo - 1 - c;
  if (c >= hi) return 2 * hi - 1 - c;
  return c;
}

// The old solution at level-`level` cell (x, y) (adapt.rs:352-402).
double sample_old(const lgd_amr &g, const double *f, int level, int64_t x, int64_t y) {
  if (level < g.num_levels()) {
    int64_t b = g.find_patch(level, x, y);
    if (b >= 0) return f[g.index(b, x - g.box[b].lo[0], y - g.box[b].lo[1])];
  }
  const double r = (double)g.ratios[level - 1];
  const lgd_box cd = g.domain(level - 1);
  int64_t c0[2];
  double w[2];
  const int64_t cell[2] = {x, y};
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
    for (int d = 0; d < 2; ++d) cc[d] = mirror_coord(cc[d], cd.lo[d], cd.hi[d]);
    acc = acc + sample_old(g, f, level - 1, cc[0], cc[1]) * weight;
  }
  return acc;
}

// The tagged and buffer-dilated cells of level `src` (GradientTagger::tag_level,
// adapt.rs:78-103, then the dilation of adapt.rs:181-193) as a bitmask over the
// level domain: row-major, 64 cells per word, bit x % 64 of word (y, x / 64).
void host_flag_bits(const lgd_amr &g, const std::vector<double> &tag_field,
                    const lgd_regrid_policy &pol, int src, std::vector<uint64_t> &bits) {
  const lgd_box dom = g.domain(src);
  const int64_t W = dom.hi[0], H = dom.hi[1], wpr = (W + 63) / 64;
  bits.assign((size_t)(H * wpr), 0ull);
  const double *v = tag_field.data();
  for (uint32_t b = g.level_start[src]; b < g.level_start[src + 1]; ++b) {
    const lgd_box &p = g.box[b];
    for (int64_t j = 0; j < p.extent(1); ++j)
      for (int64_t i = 0; i < p.extent(0); ++i) {
        double gx = std::fabs((v[g.index(b, i + 1, j)] - v[g.index(b, i - 1, j)]) * 0.5);
        double gy = std::fabs((v[g.index(b, i, j + 1)] - v[g.index(b, i, j - 1)]) * 0.5);
        double gm = std::fmax(std::fmax(0.0, gx), gy);
        if (!(gm > pol.threshold)) continue;
        const int64_t cx = p.lo[0] + i, cy = p.lo[1] + j, r = pol.buffer > 0 ? pol.buffer : 0;
        for (int64_t y = std::max<int64_t>(cy - r, 0); y <= std::min<int64_t>(cy + r, H - 1); ++y)
          for (int64_t x = std::max<int64_t>(cx - r, 0); x <= std::min<int64_t>(cx + r, W - 1); ++x)
            bits[(size_t)(y * wpr + x / 64)] |= 1ull << (x % 64);
      }
  }
}

}  // namespace

int lgd_regrid_propose_bits(const lgd_amr &g, const std::vector<uint64_t *> &bits,
                            const lgd_regrid_policy &pol, std::vector<std::vector<lgd_box>> &out) {
  const int capacity = g.max_levels() - 1;
  out.assign((size_t)capacity, {});
  for (int target = capacity; target >= 1; --target) {
    const int src = target - 1;
    const lgd_box dom = g.domain(src);
    const int64_t W = dom.hi[0], H = dom.hi[1], wpr = (W + 63) / 64;
    const size_t words = (size_t)(H * wpr);
    std::vector<uint64_t> none;
    uint64_t *m = (size_t)src < bits.size() ? bits[(size_t)src] : nullptr;
    if (!m) {
      none.assign(words, 0ull);
      m = none.data();
    }
    // Nesting flags: the coarse cells under the level being built above, grown by
    // one cell (adapt.rs:194-211).
    if (target < capacity && !out[(size_t)target].empty()) {
      const int64_t rf = g.ratios[(size_t)target], rs = g.ratios[(size_t)target - 1];
      for (const lgd_box &bx : out[(size_t)target]) {
        int64_t nlo[2] = {bx.lo[0] / rf - 1, bx.lo[1] / rf - 1};
        int64_t nhi[2] = {bx.hi[0] / rf + 1, bx.hi[1] / rf + 1};
        const int64_t y0 = std::max<int64_t>(floor_div(nlo[1], rs), 0);
        const int64_t y1 = std::min<int64_t>(floor_div(nhi[1] + rs - 1, rs), H);
        const int64_t x0 = std::max<int64_t>(floor_div(nlo[0], rs), 0);
        const int64_t x1 = std::min<int64_t>(floor_div(nhi[0] + rs - 1, rs), W);
        for (int64_t y = y0; y < y1; ++y)
          for (int64_t x = x0; x < x1; ++x) m[(size_t)(y * wpr + x / 64)] |= 1ull << (x % 64);
      }
    }
    // The flagged cells, each once (the set the reference clusters after its
    // clip / sort / dedupe; clustering does not depend on their order), and their
    // bounding box.
    size_t nflags = 0;
    for (size_t i = 0; i < words; ++i) nflags += (size_t)__builtin_popcountll(m[i]);
    if (nflags == 0) continue;
    std::vector<Cell> flags(nflags);
    lgd_box fb = empty_box();
    size_t k = 0;
    for (int64_t y = 0; y < H; ++y) {
      const uint64_t *row = m + (size_t)(y * wpr);
      bool any = false;
      for (int64_t w = 0; w < wpr; ++w) {
        uint64_t word = row[w];
        if (!word) continue;
        any = true;
        fb.lo[0] = std::min<int64_t>(fb.lo[0], w * 64 + __builtin_ctzll(word));
        fb.hi[0] = std::max<int64_t>(fb.hi[0], w * 64 + 64 - __builtin_clzll(word));
        while (word) {
          flags[k++] = Cell{(int32_t)(w * 64 + __builtin_ctzll(word)), (int32_t)y};
          word &= word - 1;
        }
      }
      if (any) {
        fb.lo[1] = std::min<int64_t>(fb.lo[1], y);
        fb.hi[1] = y + 1;
      }
    }
    std::vector<lgd_box> boxes;
    std::vector<uint32_t> sig;
    cluster_rec(flags.data(), flags.size(), fb, pol.efficiency, (size_t)pol.min_width, boxes, 3, sig);
    const int64_t r = g.ratios[(size_t)target - 1];
    for (lgd_box &b : boxes)
      for (int d = 0; d < 2; ++d) {
        b.lo[d] *= r;
        b.hi[d] *= r;
      }
    out[(size_t)target - 1] = std::move(boxes);
  }
  while (!out.empty() && out.back().empty()) out.pop_back();
  // Hysteresis: keep the current hierarchy while it still serves.
  if ((int)out.size() != g.num_levels() - 1) return 1;
  for (size_t l = 0; l < out.size(); ++l) {
    uint64_t pc = 0, cc = 0;
    for (const lgd_box &b : out[l]) pc += b.cells();
    for (uint32_t p = g.level_start[l + 1]; p < g.level_start[l + 2]; ++p) cc += g.box[p].cells();
    if (cc > pc + pc / 2) return 1;
    for (const lgd_box &b : out[l])
      if (!box_covered(g, (int)l + 1, b)) return 1;
  }
  return 0;
}

void lgd_regrid_bits_from_cells(const lgd_amr &g, int src, const std::vector<int64_t> &cells,
                                int64_t buffer, std::vector<uint64_t> &bits) {
  const lgd_box dom = g.domain(src);
  const int64_t W = dom.hi[0], H = dom.hi[1], wpr = (W + 63) / 64;
  bits.assign((size_t)(H * wpr), 0ull);
  const int64_t r = buffer > 0 ? buffer : 0;
  for (size_t k = 0; k + 1 < cells.size(); k += 2) {
    const int64_t cx = cells[k], cy = cells[k + 1];
    // The ball of the buffer dilation (adapt.rs:181-193), clipped to the domain.
    for (int64_t y = std::max<int64_t>(cy - r, 0); y <= std::min<int64_t>(cy + r, H - 1); ++y)
      for (int64_t x = std::max<int64_t>(cx - r, 0); x <= std::min<int64_t>(cx + r, W - 1); ++x)
        bits[(size_t)(y * wpr + x / 64)] |= 1ull << (x % 64);
  }
}

int lgd_regrid_propose(const lgd_amr &g, const std::vector<double> &tag_field,
                       const lgd_regrid_policy &pol, std::vector<std::vector<lgd_box>> &out) {
  std::vector<std::vector<uint64_t>> bits((size_t)g.num_levels());
  std::vector<uint64_t *> ptr((size_t)g.num_levels(), nullptr);
  for (int src = 0; src < g.num_levels() && src < g.max_levels() - 1; ++src) {
    host_flag_bits(g, tag_field, pol, src, bits[(size_t)src]);
    ptr[(size_t)src] = bits[(size_t)src].data();
  }
  return lgd_regrid_propose_bits(g, ptr, pol, out);
}

void lgd_regrid_migrate(const lgd_amr &old_g, const std::vector<double> &old_f, const lgd_amr &new_g,
                        std::vector<double> &new_f) {
  // Level 0 is identical in both hierarchies (the base tiling never changes) and
  // is copied on the device by the caller; new_f holds levels >= 1 only, from
  // new_g.off[level_start[1]] on.
  const int64_t fine0 = new_g.off[new_g.level_start[1]];
  new_f.assign((size_t)(new_g.off.back() - fine0), 0.0);
  for (int l = 1; l < new_g.num_levels(); ++l)
    for (uint32_t b = new_g.level_start[l]; b < new_g.level_start[l + 1]; ++b) {
      const lgd_box &p = new_g.box[b];
      for (int64_t j = 0; j < p.extent(1); ++j)
        for (int64_t i = 0; i < p.extent(0); ++i)
          new_f[(size_t)(new_g.index(b, i, j) - fine0)] =
              sample_old(old_g, old_f.data(), l, p.lo[0] + i, p.lo[1] + j);
    }
}

// ------------------------------------------------------------------ device path
// The same regrid with the field-sized work on the device: the tagger and the
// buffer dilation write a bitmask per level (so the host reads bits, not the
// field), and the migration samples the old hierarchy where the state lives.
namespace {

constexpr int kRTX = 64, kRTY = 4;  // the level plans' tile shape
constexpr int kMaxSampleDepth = 7;  // levels below the sampled one (8-level hierarchies)

using DevIndex = lgd_dev_index;

struct DevRatios {
  int64_t cells[2];  // level-0 domain
  uint32_t r[kMaxSampleDepth + 1];
};

__device__ int64_t dev_mirror(int64_t c, int64_t lo, int64_t hi) {
  if (c < lo) return 2 * lo - 1 - c;
  if (c >= hi) return 2 * hi - 1 - c;
  return c;
}

// sample_old above, on the device: the same lookups, corner order and operation
// order.  K bounds the recursion (level 0 always has a patch).
template <int K>
__device__ __attribute__((noinline)) double dev_sample(const DevIndex &ix, const DevRatios &R,
                                                       const double *f, int level, int64_t x,
                                                       int64_t y) {
  const int64_t b = lgd_dev_find_patch(ix, level, x, y);
  if (b >= 0) {
    const lgd_block_desc bd = ix.blocks[b];
    return f[bd.off + (y - ix.lo[2 * b + 1] + 1) * (int64_t)(bd.nx + 2) + (x - ix.lo[2 * b] + 1)];
  }
  if constexpr (K == 0) {
    return 0.0;
  } else {
    const double r = (double)R.r[level - 1];
    int64_t s = 1;
    for (int i = 0; i < level - 1; ++i) s *= (int64_t)R.r[i];
    const int64_t chi[2] = {R.cells[0] * s, R.cells[1] * s};
    int64_t c0[2];
    double w[2];
    const int64_t cell[2] = {x, y};
    for (int d = 0; d < 2; ++d) {
      double pc = ((double)cell[d] + 0.5) / r - 0.5;
      double fl = floor(pc);
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
      for (int d = 0; d < 2; ++d) cc[d] = dev_mirror(cc[d], 0, chi[d]);
      acc = acc + dev_sample<K - 1>(ix, R, f, level - 1, cc[0], cc[1]) * weight;
    }
    return acc;
  }
}

// New patch interiors of one level sampled from the old hierarchy.
__global__ __launch_bounds__(kRTX *kRTY) void migrate_kernel(const lgd_tile *tiles,
                                                             const lgd_block_desc *nblocks,
                                                             const int64_t *nlo, DevIndex old,
                                                             DevRatios R, const double *oldf,
                                                             double *newf, int level) {
  const lgd_tile tl = tiles[blockIdx.x];
  const lgd_block_desc bd = nblocks[tl.block];
  const int i = tl.x0 + threadIdx.x, j = tl.y0 + threadIdx.y;
  if (i >= bd.nx || j >= bd.ny) return;
  const int64_t x = nlo[2 * tl.block] + i, y = nlo[2 * tl.block + 1] + j;
  newf[bd.off + (int64_t)(j + 1) * (bd.nx + 2) + (i + 1)] =
      dev_sample<kMaxSampleDepth>(old, R, oldf, level, x, y);
}

// GradientTagger on one level's patch interiors: the flagged cells' bits in a
// row-major bitmask of the level domain (wpr words of 64 cells per row).
__global__ __launch_bounds__(kRTX *kRTY) void tag_kernel(const lgd_tile *tiles,
                                                         const lgd_block_desc *blocks,
                                                         const int64_t *lo, const double *v,
                                                         double threshold,
                                                         unsigned long long *T, int64_t wpr) {
  const lgd_tile tl = tiles[blockIdx.x];
  const lgd_block_desc bd = blocks[tl.block];
  const int i = tl.x0 + threadIdx.x, j = tl.y0 + threadIdx.y;
  if (i >= bd.nx || j >= bd.ny) return;
  const int64_t sx = bd.nx + 2;
  const double *c = v + bd.off + (int64_t)(j + 1) * sx + (i + 1);
  const double gx = fabs((c[1] - c[-1]) * 0.5);
  const double gy = fabs((c[sx] - c[-sx]) * 0.5);
  const double gm = fmax(fmax(0.0, gx), gy);
  if (gm > threshold) {
    const int64_t x = lo[2 * tl.block] + i, y = lo[2 * tl.block + 1] + j;
    atomicOr(T + y * wpr + x / 64, 1ull << (x % 64));
  }
}

// Buffer dilation, one bitmask word (64 cells) per thread: the OR over rows
// y−b … y+b of each row's word shifted by −b … b cells (bits carried in from the
// neighbouring words), clipped to the domain (width W).
__global__ __launch_bounds__(256) void dilate_kernel(const unsigned long long *T, int64_t W,
                                                     int64_t H, int64_t wpr, int buffer,
                                                     unsigned long long *bits) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= H * wpr) return;
  const int64_t y = t / wpr, w = t - y * wpr;
  unsigned long long out = 0;
  for (int64_t yy = y - buffer; yy <= y + buffer; ++yy) {
    if (yy < 0 || yy >= H) continue;
    const unsigned long long *row = T + yy * wpr;
    const unsigned long long c = row[w], L = w > 0 ? row[w - 1] : 0ull,
                             R = w + 1 < wpr ? row[w + 1] : 0ull;
    out |= c;
    for (int d = 1; d <= buffer; ++d) {
      out |= (c << d) | (L >> (64 - d));  // cell x takes the flag of x − d
      out |= (c >> d) | (R << (64 - d));  // and of x + d
    }
  }
  const int64_t valid = W - w * 64;  // cells of this word inside the domain
  if (valid < 64) out &= (1ull << valid) - 1ull;
  bits[t] = out;
}

DevIndex dev_index(const lgd_amr &g) {
  return DevIndex{g.d_levels, g.d_bstart, g.d_bids, g.d_lo, g.d_blocks, g.num_levels()};
}

}  // namespace

using namespace lgd;

bool lgd_regrid_device_supported(const lgd_amr &g) { return g.max_levels() <= kMaxSampleDepth + 1; }

int lgd_regrid_flag_bits(lgd_ctx *ctx, const lgd_amr &g, const double *tag_field,
                         const lgd_regrid_policy &pol, lgd_regrid_scratch &sc,
                         std::vector<uint64_t *> &bits) {
  const int levels = std::min(g.num_levels(), g.max_levels() - 1);
  bits.assign((size_t)g.num_levels(), nullptr);
  if (levels <= 0) return LGD_OK;
  if (pol.buffer > 63)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "device regrid dilates by at most 63 cells");
  size_t wtotal = 0;
  std::vector<size_t> wofs;
  for (int src = 0; src < levels; ++src) {
    const lgd_box d = g.domain(src);
    wofs.push_back(wtotal);
    wtotal += (size_t)(d.hi[1] * ((d.hi[0] + 63) / 64));
  }
  // Tags (T) and dilated flags (bits) for every level, one region each.
  hipError_t e = hipSuccess;
  if (sc.wcap < wtotal) {
    (void)hipFree(sc.T);
    (void)hipFree(sc.bits);
    if (sc.host) (void)hipHostFree(sc.host);
    sc.T = nullptr;
    sc.bits = nullptr;
    sc.host = nullptr;
    sc.wcap = 0;
    e = hipMalloc((void **)&sc.T, wtotal * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&sc.bits, wtotal * sizeof(uint64_t));
    if (e == hipSuccess) e = hipHostMalloc((void **)&sc.host, wtotal * sizeof(uint64_t), hipHostMallocDefault);
    if (e != hipSuccess) {
      sc.host = nullptr;
      return hip_fail(ctx, e, "regrid tag scratch");
    }
    sc.wcap = wtotal;
  }
  e = hipMemsetAsync(sc.T, 0, wtotal * sizeof(uint64_t), ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "regrid tag clear");
  int rc;
  const int buffer = (int)std::max<int64_t>(pol.buffer, 0);
  for (int src = 0; src < levels; ++src) {
    const lgd_box d = g.domain(src);
    const int64_t W = d.hi[0], H = d.hi[1], wpr = (W + 63) / 64;
    unsigned long long *T = (unsigned long long *)(sc.T + wofs[(size_t)src]);
    const lgd_level_plan &P = g.plans[(size_t)src];
    if (P.ntiles > 0) {
      tag_kernel<<<(unsigned)P.ntiles, dim3(kRTX, kRTY), 0, ctx->stream>>>(
          P.tiles, g.d_blocks, g.d_lo, tag_field, pol.threshold, T, wpr);
      if ((rc = launch_check(ctx, "regrid tag"))) return rc;
    }
    dilate_kernel<<<blocks_for(H * wpr, 256), 256, 0, ctx->stream>>>(
        T, W, H, wpr, buffer, (unsigned long long *)(sc.bits + wofs[(size_t)src]));
    if ((rc = launch_check(ctx, "regrid dilate"))) return rc;
  }
  e = hipMemcpyAsync(sc.host, sc.bits, wtotal * sizeof(uint64_t), hipMemcpyDeviceToHost,
                     ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "regrid bitmask download");
  for (int src = 0; src < levels; ++src) bits[(size_t)src] = sc.host + wofs[(size_t)src];
  return LGD_OK;
}

int lgd_regrid_migrate_device(lgd_ctx *ctx, const lgd_amr &old_g, const double *old_f,
                              const lgd_amr &new_g, double *new_f) {
  if (!lgd_regrid_device_supported(new_g))
    return fail(ctx, LGD_ERR_UNSUPPORTED, "device migration supports at most 8 levels");
  DevRatios R{};
  R.cells[0] = old_g.cells[0];
  R.cells[1] = old_g.cells[1];
  for (size_t i = 0; i < old_g.ratios.size() && i <= (size_t)kMaxSampleDepth; ++i) R.r[i] = old_g.ratios[i];
  const DevIndex old = dev_index(old_g);
  int rc;
  for (int l = 1; l < new_g.num_levels(); ++l) {
    const lgd_level_plan &P = new_g.plans[(size_t)l];
    if (P.ntiles == 0) continue;
    migrate_kernel<<<(unsigned)P.ntiles, dim3(kRTX, kRTY), 0, ctx->stream>>>(
        P.tiles, new_g.d_blocks, new_g.d_lo, old, R, old_f, new_f, l);
    if ((rc = launch_check(ctx, "regrid migrate"))) return rc;
  }
  return LGD_OK;
}

void lgd_regrid_scratch_free(lgd_regrid_scratch &sc) {
  (void)hipFree(sc.T);
  (void)hipFree(sc.bits);
  if (sc.host) (void)hipHostFree(sc.host);
  sc = lgd_regrid_scratch{};
}
