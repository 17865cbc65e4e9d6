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
#include <cmath>
#include <vector>

#include "lgd_regrid.h"

namespace {

using Cell = std::array<int64_t, 2>;

lgd_box bounding_box(const Cell *c, size_t n) {
  lgd_box b{{c[0][0], c[0][1]}, {c[0][0] + 1, c[0][1] + 1}};
  for (size_t i = 1; i < n; ++i)
    for (int d = 0; d < 2; ++d) {
      if (c[i][d] < b.lo[d]) b.lo[d] = c[i][d];
      if (c[i][d] + 1 > b.hi[d]) b.hi[d] = c[i][d] + 1;
    }
  return b;
}

struct Split {
  int axis;
  int64_t plane;
};

// Interior zero of a signature nearest the middle (first found wins ties).
bool hole_split(const lgd_box &bx, const std::vector<uint32_t> (&sig)[2], Split &s) {
  bool have = false;
  uint64_t best = 0;
  for (int d = 0; d < 2; ++d) {
    const size_t n = sig[d].size();
    for (size_t i = 1; i + 1 < n; ++i) {
      if (sig[d][i] != 0) continue;
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
bool inflection_split(const lgd_box &bx, const std::vector<uint32_t> (&sig)[2], size_t minw,
                      Split &s) {
  bool have = false;
  int64_t best_s = 0;
  uint64_t best_c = 0;
  for (int d = 0; d < 2; ++d) {
    const size_t n = sig[d].size();
    if (n < 4) continue;
    for (size_t i = 1; i + 2 < n; ++i) {
      int64_t a = (int64_t)sig[d][i + 1] - 2 * (int64_t)sig[d][i] + (int64_t)sig[d][i - 1];
      int64_t b = (int64_t)sig[d][i + 2] - 2 * (int64_t)sig[d][i + 1] + (int64_t)sig[d][i];
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

void cluster_rec(Cell *f, size_t n, double eff, size_t minw, std::vector<lgd_box> &out) {
  const lgd_box bx = bounding_box(f, n);
  if ((double)n / (double)bx.cells() >= eff) {
    out.push_back(bx);
    return;
  }
  std::vector<uint32_t> sig[2];
  sig[0].assign((size_t)bx.extent(0), 0u);
  sig[1].assign((size_t)bx.extent(1), 0u);
  for (size_t i = 0; i < n; ++i) {
    sig[0][(size_t)(f[i][0] - bx.lo[0])]++;
    sig[1][(size_t)(f[i][1] - bx.lo[1])]++;
  }
  Split s{0, 0};
  if (!hole_split(bx, sig, s) && !inflection_split(bx, sig, minw, s) && !bisect_split(bx, minw, s)) {
    out.push_back(bx);
    return;
  }
  // Two-pointer partition (cluster.rs:270-281): cells below the plane move to the front.
  size_t lt = 0;
  for (size_t j = 0; j < n; ++j)
    if (f[j][s.axis] < s.plane) std::swap(f[lt++], f[j]);
  cluster_rec(f, lt, eff, minw, out);
  cluster_rec(f + lt, n - lt, eff, minw, out);
}

bool box_covered(const lgd_box &need, const std::vector<lgd_box> &cover) {
  for (int64_t y = need.lo[1]; y < need.hi[1]; ++y) {
    int64_t x = need.lo[0];
    while (x < need.hi[0]) {
      const lgd_box *hit = nullptr;
      for (const lgd_box &b : cover)
        if (b.contains(x, y)) {
          hit = &b;
          break;
        }
      if (!hit) return false;
      x = hit->hi[0];
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

}  // namespace

int lgd_regrid_propose(const lgd_amr &g, const std::vector<double> &tag_field,
                       const lgd_regrid_policy &pol, std::vector<std::vector<lgd_box>> &out) {
  const int capacity = g.max_levels() - 1;
  out.assign((size_t)capacity, {});
  for (int target = capacity; target >= 1; --target) {
    const int src = target - 1;
    std::vector<Cell> flags;
    if (src < g.num_levels()) {
      for (uint32_t b = g.level_start[src]; b < g.level_start[src + 1]; ++b) {
        const lgd_box &p = g.box[b];
        for (int64_t j = 0; j < p.extent(1); ++j)
          for (int64_t i = 0; i < p.extent(0); ++i) {
            const double *v = tag_field.data();
            double gx = std::fabs((v[g.index(b, i + 1, j)] - v[g.index(b, i - 1, j)]) * 0.5);
            double gy = std::fabs((v[g.index(b, i, j + 1)] - v[g.index(b, i, j - 1)]) * 0.5);
            double gm = std::fmax(std::fmax(0.0, gx), gy);
            if (gm > pol.threshold) flags.push_back({p.lo[0] + i, p.lo[1] + j});
          }
      }
    }
    if (pol.buffer > 0) {
      std::vector<Cell> seeds;
      seeds.swap(flags);
      for (const Cell &c : seeds)
        for (int64_t y = c[1] - pol.buffer; y <= c[1] + pol.buffer; ++y)
          for (int64_t x = c[0] - pol.buffer; x <= c[0] + pol.buffer; ++x) flags.push_back({x, y});
    }
    if (target < capacity && !out[(size_t)target].empty()) {
      const int64_t rf = g.ratios[(size_t)target], rs = g.ratios[(size_t)target - 1];
      for (const lgd_box &bx : out[(size_t)target]) {
        int64_t nlo[2] = {bx.lo[0] / rf - 1, bx.lo[1] / rf - 1};
        int64_t nhi[2] = {bx.hi[0] / rf + 1, bx.hi[1] / rf + 1};
        for (int64_t y = floor_div(nlo[1], rs); y < floor_div(nhi[1] + rs - 1, rs); ++y)
          for (int64_t x = floor_div(nlo[0], rs); x < floor_div(nhi[0] + rs - 1, rs); ++x)
            flags.push_back({x, y});
      }
    }
    const lgd_box dom = g.domain(src);
    flags.erase(std::remove_if(flags.begin(), flags.end(),
                               [&](const Cell &c) { return !dom.contains(c[0], c[1]); }),
                flags.end());
    std::sort(flags.begin(), flags.end());
    flags.erase(std::unique(flags.begin(), flags.end()), flags.end());
    if (flags.empty()) continue;
    std::vector<lgd_box> boxes;
    cluster_rec(flags.data(), flags.size(), pol.efficiency, (size_t)pol.min_width, boxes);
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
    std::vector<lgd_box> cur(g.box.begin() + g.level_start[l + 1], g.box.begin() + g.level_start[l + 2]);
    uint64_t pc = 0, cc = 0;
    for (const lgd_box &b : out[l]) pc += b.cells();
    for (const lgd_box &b : cur) cc += b.cells();
    if (cc > pc + pc / 2) return 1;
    for (const lgd_box &b : out[l])
      if (!box_covered(b, cur)) return 1;
  }
  return 0;
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
