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
#include <vector>

#include "lgd_regrid.h"

namespace {

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
// min_width cells on each side (stronger, then more central, wins; the first
// found, axis 0 before axis 1 and lower planes first, on a full tie).  Planes
// closer than min_width to either end are never candidates, so only the second
// differences around the others are formed.
bool inflection_split(const lgd_box &bx, const Sig &sig, size_t minw, Split &s) {
  bool have = false;
  int64_t best_s = 0;
  uint64_t best_c = 0;
  for (int d = 0; d < 2; ++d) {
    const size_t n = sig.n[d];
    const uint32_t *v = sig.v[d];
    if (n < 4) continue;
    // plane = i + 1 with 1 <= i <= n - 3, minw <= plane <= n - minw.
    const size_t i0 = std::max<size_t>(1, minw > 0 ? minw - 1 : 0);
    if (n < minw + 1) continue;
    const size_t i1 = std::min<size_t>(n - 3, n - minw - 1);  // inclusive
    if (i0 > i1) continue;
    int64_t a = (int64_t)v[i0 + 1] - 2 * (int64_t)v[i0] + (int64_t)v[i0 - 1];
    for (size_t i = i0; i <= i1; ++i) {
      const int64_t b = (int64_t)v[i + 2] - 2 * (int64_t)v[i + 1] + (int64_t)v[i];
      if ((a < 0 && b > 0) || (a > 0 && b < 0)) {
        const size_t plane = i + 1;
        const int64_t strength = b - a < 0 ? a - b : b - a;
        const uint64_t c = std::min<uint64_t>(plane, n - plane);
        if (!have || strength > best_s || (strength == best_s && c > best_c)) {
          have = true;
          best_s = strength;
          best_c = c;
          s = Split{d, bx.lo[d] + (int64_t)plane};
        }
      }
      a = b;
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

// A run of flagged cells: the set bits of `bits` are cells (64·w + bit, y) of the
// level.  A node's runs are in increasing (y, w) order, with no bits outside its box.
struct Run {
  int32_t y, w;
  uint64_t bits;
};

// The bits of each byte value as eight 0/1 counts (column signatures add a byte of
// a run at a time).
struct ByteBits {
  uint32_t v[256][8];
  ByteBits() {
    for (int x = 0; x < 256; ++x)
      for (int j = 0; j < 8; ++j) v[x][j] = (uint32_t)((x >> j) & 1);
  }
};
const ByteBits kByteBits;

// Column (axis 0) or row (axis 1) counts of the runs' bits into v, indexed from lo.
void count_runs(const Run *r, size_t m, int axis, int64_t lo, size_t len, uint32_t *v) {
  if (axis == 1) {
    for (size_t i = 0; i < m; ++i) v[(size_t)(r[i].y - lo)] += (uint32_t)__builtin_popcountll(r[i].bits);
    return;
  }
  for (size_t i = 0; i < m; ++i) {
    // A byte of the run at a time (the run's bits lie inside [lo, lo + len)).
    uint64_t b = r[i].bits;
    const int64_t base = (int64_t)r[i].w * 64 - lo;
    while (b) {
      const int k = __builtin_ctzll(b) >> 3;
      const uint32_t *lut = kByteBits.v[(b >> (8 * k)) & 0xffu];
      const int64_t c0 = base + 8 * k;
      if (c0 >= 0 && c0 + 8 <= (int64_t)len) {
        for (int j = 0; j < 8; ++j) v[c0 + j] += lut[j];
      } else {
        for (int j = 0; j < 8; ++j)
          if (lut[j]) v[c0 + j] += 1;
      }
      b &= ~(0xffull << (8 * k));
    }
  }
}

// Berger–Rigoutsos over the n flagged cells of the m runs r whose bounding box is bx
// (cluster.rs:157-281): accept the box at the efficiency, else split at a signature
// hole, an inflection or the midpoint and recurse (left half first).  sig[0] and
// sig[1] are the box's column and row counts.  A split across rows divides the run
// list in two, a split across columns masks each run at the plane.  The halves'
// counts along the split axis are slices of the box's; across it, the half with
// fewer runs counts its own and the other's are the box's less those.  Each half's
// bounding box comes from the counts along the split axis and from its runs across
// it.  One host thread: per regrid the clustering is a few hundred microseconds, less
// than starting threads for it costs.
void cluster_rec(const Run *r, size_t m, size_t n, const lgd_box &bx, std::vector<uint32_t> sig[2],
                 double eff, size_t minw, std::vector<lgd_box> &out) {
  if ((double)n / (double)bx.cells() >= eff) {
    out.push_back(bx);
    return;
  }
  const size_t nx = (size_t)bx.extent(0), ny = (size_t)bx.extent(1);
  const Sig sg{{sig[0].data(), sig[1].data()}, {nx, ny}};
  Split s{0, 0};
  if (!hole_split(bx, sg, s) && !inflection_split(bx, sg, minw, s) && !bisect_split(bx, minw, s)) {
    out.push_back(bx);
    return;
  }
  // The halves: flag counts and extents along the split axis from its counts.
  const int ax = s.axis, ox = 1 - ax;
  const uint32_t *sa = sig[ax].data();
  const size_t na = ax == 0 ? nx : ny, cut = (size_t)(s.plane - bx.lo[ax]);
  size_t nl = 0;
  for (size_t i = 0; i < cut; ++i) nl += sa[i];
  lgd_box lb = empty_box(), rb = empty_box();
  for (size_t i = 0; i < cut; ++i)
    if (sa[i]) {
      lb.lo[ax] = std::min<int64_t>(lb.lo[ax], bx.lo[ax] + (int64_t)i);
      lb.hi[ax] = bx.lo[ax] + (int64_t)i + 1;
    }
  for (size_t i = cut; i < na; ++i)
    if (sa[i]) {
      rb.lo[ax] = std::min<int64_t>(rb.lo[ax], bx.lo[ax] + (int64_t)i);
      rb.hi[ax] = bx.lo[ax] + (int64_t)i + 1;
    }
  // Across the split axis, from the runs of each half.
  auto extent_of = [&](const Run *q, size_t k, lgd_box &b) {
    if (ox == 1) {
      if (k) {
        b.lo[1] = q[0].y;
        b.hi[1] = (int64_t)q[k - 1].y + 1;
      }
      return;
    }
    for (size_t i = 0; i < k; ++i) {
      const int64_t x0 = (int64_t)q[i].w * 64;
      b.lo[0] = std::min<int64_t>(b.lo[0], x0 + __builtin_ctzll(q[i].bits));
      b.hi[0] = std::max<int64_t>(b.hi[0], x0 + 64 - __builtin_clzll(q[i].bits));
    }
  };
  const Run *left, *right;
  size_t ml, mr;
  std::vector<Run> lv, rv;
  if (ax == 1) {
    // Rows below the plane: a prefix of the runs.
    size_t k = 0;
    while (k < m && r[k].y < s.plane) ++k;
    left = r;
    ml = k;
    right = r + k;
    mr = m - k;
  } else {
    lv.reserve(m);
    rv.reserve(m);
    for (size_t i = 0; i < m; ++i) {
      const int64_t c = s.plane - (int64_t)r[i].w * 64;  // bits below c go left
      const uint64_t lm = c >= 64 ? ~0ull : c <= 0 ? 0ull : (1ull << c) - 1ull;
      const uint64_t lbits = r[i].bits & lm, rbits = r[i].bits & ~lm;
      if (lbits) lv.push_back(Run{r[i].y, r[i].w, lbits});
      if (rbits) rv.push_back(Run{r[i].y, r[i].w, rbits});
    }
    left = lv.data();
    ml = lv.size();
    right = rv.data();
    mr = rv.size();
  }
  extent_of(left, ml, lb);
  extent_of(right, mr, rb);
  const size_t nr = n - nl;
  // The halves' counts, each over its own box.
  const size_t no = ox == 0 ? nx : ny;
  std::vector<uint32_t> cross((size_t)no, 0u);
  const bool left_counts = ml <= mr;
  count_runs(left_counts ? left : right, left_counts ? ml : mr, ox, bx.lo[ox], no, cross.data());
  std::vector<uint32_t> ls[2], rs[2];
  auto slice = [](const uint32_t *v, int64_t from, int64_t to, std::vector<uint32_t> &o) {
    o.assign(v + from, v + to);
  };
  slice(sa, lb.lo[ax] - bx.lo[ax], lb.hi[ax] - bx.lo[ax], ls[ax]);
  slice(sa, rb.lo[ax] - bx.lo[ax], rb.hi[ax] - bx.lo[ax], rs[ax]);
  {
    const uint32_t *so = sig[ox].data();
    std::vector<uint32_t> &mine = left_counts ? ls[ox] : rs[ox], &other = left_counts ? rs[ox] : ls[ox];
    const lgd_box &mb = left_counts ? lb : rb, &obx = left_counts ? rb : lb;
    slice(cross.data(), mb.lo[ox] - bx.lo[ox], mb.hi[ox] - bx.lo[ox], mine);
    other.resize((size_t)(obx.hi[ox] - obx.lo[ox]));
    for (size_t i = 0; i < other.size(); ++i) {
      const size_t j = (size_t)(obx.lo[ox] - bx.lo[ox]) + i;
      other[i] = so[j] - cross[j];
    }
  }
  cluster_rec(left, ml, nl, lb, ls, eff, minw, out);
  cluster_rec(right, mr, nr, rb, rs, eff, minw, out);
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

int lgd_regrid_propose_words(const lgd_amr &g, const std::vector<lgd_flag_words> &flags,
                             const lgd_regrid_policy &pol, std::vector<std::vector<lgd_box>> &out) {
  const int capacity = g.max_levels() - 1;
  out.assign((size_t)capacity, {});
  std::vector<std::pair<int64_t, uint64_t>> nest;
  std::vector<int64_t> midx;
  std::vector<uint64_t> mbits;
  for (int target = capacity; target >= 1; --target) {
    const int src = target - 1;
    const lgd_box dom = g.domain(src);
    const int64_t W = dom.hi[0], H = dom.hi[1], wpr = (W + 63) / 64;
    lgd_flag_words fw = (size_t)src < flags.size() ? flags[(size_t)src] : lgd_flag_words{};
    // Nesting flags: the coarse cells under the level being built above, grown by
    // one cell (adapt.rs:194-211), as words in row-major order merged into the
    // level's own (OR where both have a word).
    if (target < capacity && !out[(size_t)target].empty()) {
      const int64_t rf = g.ratios[(size_t)target], rs = g.ratios[(size_t)target - 1];
      nest.clear();
      for (const lgd_box &bx : out[(size_t)target]) {
        int64_t nlo[2] = {bx.lo[0] / rf - 1, bx.lo[1] / rf - 1};
        int64_t nhi[2] = {bx.hi[0] / rf + 1, bx.hi[1] / rf + 1};
        const int64_t y0 = std::max<int64_t>(floor_div(nlo[1], rs), 0);
        const int64_t y1 = std::min<int64_t>(floor_div(nhi[1] + rs - 1, rs), H);
        const int64_t x0 = std::max<int64_t>(floor_div(nlo[0], rs), 0);
        const int64_t x1 = std::min<int64_t>(floor_div(nhi[0] + rs - 1, rs), W);
        if (x0 >= x1) continue;
        for (int64_t y = y0; y < y1; ++y)
          for (int64_t w = x0 / 64; w <= (x1 - 1) / 64; ++w) {
            const int64_t a = std::max<int64_t>(x0 - w * 64, 0), b = std::min<int64_t>(x1 - w * 64, 64);
            const uint64_t m = (b == 64 ? ~0ull : (1ull << b) - 1ull) & ~((1ull << a) - 1ull);
            nest.emplace_back(y * wpr + w, m);
          }
      }
      std::sort(nest.begin(), nest.end(),
                [](const std::pair<int64_t, uint64_t> &p, const std::pair<int64_t, uint64_t> &q) {
                  return p.first < q.first;
                });
      midx.clear();
      mbits.clear();
      size_t i = 0, j = 0;
      while (i < fw.n || j < nest.size()) {
        int64_t k;
        uint64_t m = 0;
        if (j >= nest.size() || (i < fw.n && fw.idx[i] <= nest[j].first)) {
          k = fw.idx[i];
        } else {
          k = nest[j].first;
        }
        while (i < fw.n && fw.idx[i] == k) m |= fw.bits[i++];
        while (j < nest.size() && nest[j].first == k) m |= nest[j++].second;
        midx.push_back(k);
        mbits.push_back(m);
      }
      fw = lgd_flag_words{midx.data(), mbits.data(), midx.size()};
    }
    // The flagged cells, each once (the set the reference clusters after its
    // clip / sort / dedupe; clustering does not depend on their order), as runs, and
    // their bounding box.
    size_t nflags = 0;
    std::vector<Run> runs;
    runs.reserve(fw.n);
    lgd_box fb = empty_box();
    for (size_t i = 0; i < fw.n; ++i) {
      const uint64_t word = fw.bits[i];
      if (!word) continue;
      const int64_t y = fw.idx[i] / wpr, w = fw.idx[i] - y * wpr;
      nflags += (size_t)__builtin_popcountll(word);
      fb.lo[0] = std::min<int64_t>(fb.lo[0], w * 64 + __builtin_ctzll(word));
      fb.hi[0] = std::max<int64_t>(fb.hi[0], w * 64 + 64 - __builtin_clzll(word));
      fb.lo[1] = std::min<int64_t>(fb.lo[1], y);
      fb.hi[1] = std::max<int64_t>(fb.hi[1], y + 1);
      runs.push_back(Run{(int32_t)y, (int32_t)w, word});
    }
    if (nflags == 0) continue;
    std::vector<lgd_box> boxes;
    std::vector<uint32_t> sig[2];
    for (int d = 0; d < 2; ++d) {
      sig[d].assign((size_t)fb.extent(d), 0u);
      count_runs(runs.data(), runs.size(), d, fb.lo[d], sig[d].size(), sig[d].data());
    }
    cluster_rec(runs.data(), runs.size(), nflags, fb, sig, pol.efficiency, (size_t)pol.min_width, boxes);
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

int lgd_regrid_propose_bits(const lgd_amr &g, const std::vector<uint64_t *> &bits,
                            const lgd_regrid_policy &pol, std::vector<std::vector<lgd_box>> &out) {
  std::vector<std::vector<int64_t>> idx(bits.size());
  std::vector<std::vector<uint64_t>> words(bits.size());
  std::vector<lgd_flag_words> flags(bits.size());
  for (size_t l = 0; l < bits.size() && (int)l < g.max_levels() - 1; ++l) {
    if (!bits[l]) continue;
    const lgd_box dom = g.domain((int)l);
    const int64_t n = dom.hi[1] * ((dom.hi[0] + 63) / 64);
    for (int64_t k = 0; k < n; ++k)
      if (bits[l][k]) {
        idx[l].push_back(k);
        words[l].push_back(bits[l][k]);
      }
    flags[l] = lgd_flag_words{idx[l].data(), words[l].data(), idx[l].size()};
  }
  return lgd_regrid_propose_words(g, flags, pol, out);
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

// The fields a migration moves: old hierarchy's values in, new patch interiors out.
struct MigrateFields {
  const double *old[2];
  double *nw[2];
  int nf;
};

// New patch interiors of one level sampled from the old hierarchy, one 64×4 tile of
// cells per workgroup.  A tile inside one patch of the same level in the old
// hierarchy (where dev_sample's first lookup ends for each of its cells) copies from
// that patch after one lookup; any other tile samples cell by cell.
__global__ __launch_bounds__(kRTX *kRTY) void migrate_kernel(const lgd_tile *tiles,
                                                             const lgd_block_desc *nblocks,
                                                             const int64_t *nlo, DevIndex old,
                                                             DevRatios R, MigrateFields F, int level) {
  __shared__ int64_t src_patch;
  const lgd_tile tl = tiles[blockIdx.x];
  const lgd_block_desc bd = nblocks[tl.block];
  const int64_t X0 = nlo[2 * tl.block] + tl.x0, Y0 = nlo[2 * tl.block + 1] + tl.y0;
  if (threadIdx.x == 0 && threadIdx.y == 0) {
    const int64_t tw = min(kRTX, bd.nx - tl.x0), th = min(kRTY, bd.ny - tl.y0);
    int64_t p = lgd_dev_find_patch(old, level, X0, Y0);
    if (p >= 0) {
      const lgd_block_desc ob = old.blocks[p];
      if (X0 + tw > old.lo[2 * p] + ob.nx || Y0 + th > old.lo[2 * p + 1] + ob.ny) p = -1;
    }
    src_patch = p;
  }
  __syncthreads();
  const int i = tl.x0 + threadIdx.x, j = tl.y0 + threadIdx.y;
  if (i >= bd.nx || j >= bd.ny) return;
  const int64_t x = X0 + threadIdx.x, y = Y0 + threadIdx.y;
  const int64_t dst = bd.off + (int64_t)(j + 1) * (bd.nx + 2) + (i + 1);
  const int64_t p = src_patch;
  if (p >= 0) {
    const lgd_block_desc ob = old.blocks[p];
    const int64_t o = ob.off + (y - old.lo[2 * p + 1] + 1) * (int64_t)(ob.nx + 2) + (x - old.lo[2 * p] + 1);
    for (int f = 0; f < F.nf; ++f) F.nw[f][dst] = F.old[f][o];
    return;
  }
  for (int f = 0; f < F.nf; ++f) F.nw[f][dst] = dev_sample<kMaxSampleDepth>(old, R, F.old[f], level, x, y);
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

// The bitmask words of every refinable level, one after another: level l's
// H[l]·wpr[l] words start at word wofs[l] and at workgroup first[l] of the launches
// below (kWordWG words per workgroup).
constexpr int kWordWG = 256;
constexpr int kMaxWordLevels = kMaxSampleDepth + 1;
struct LevelWords {
  int64_t first[kMaxWordLevels + 1];
  int64_t wofs[kMaxWordLevels];
  int64_t W[kMaxWordLevels], H[kMaxWordLevels], wpr[kMaxWordLevels];
  int n;
};

// The level of workgroup b and the level-local word of its thread.
__device__ __forceinline__ int word_level(const LevelWords &L, int64_t b, int64_t &t) {
  int l = 0;
  while (l + 1 < L.n && b >= L.first[l + 1]) ++l;
  t = (b - L.first[l]) * kWordWG + threadIdx.x;
  return l;
}

// The nonzero flags of this workgroup's words, summed over its wavefronts.
__device__ __forceinline__ int block_count(bool nz, int *lds) {
  const unsigned long long m = __ballot(nz);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  int c = 0;
  for (int w = 0; w < kWordWG / 64; ++w) c += lds[w];
  return c;
}

// Buffer dilation, one bitmask word (64 cells) per thread: the OR over rows
// y−b … y+b of each row's word shifted by −b … b cells (bits carried in from the
// neighbouring words), clipped to the domain (width W); each workgroup also counts
// its nonzero words.
__global__ __launch_bounds__(kWordWG) void dilate_kernel(LevelWords L, const unsigned long long *T0,
                                                         int buffer, unsigned long long *bits0,
                                                         int32_t *counts) {
  __shared__ int lds[kWordWG / 64];
  int64_t t;
  const int l = word_level(L, blockIdx.x, t);
  const int64_t W = L.W[l], H = L.H[l], wpr = L.wpr[l];
  const unsigned long long *T = T0 + L.wofs[l];
  unsigned long long out = 0;
  const bool valid = t < H * wpr;
  if (valid) {
    const int64_t y = t / wpr, w = t - y * wpr;
    for (int64_t yy = y - buffer; yy <= y + buffer; ++yy) {
      if (yy < 0 || yy >= H) continue;
      const unsigned long long *row = T + yy * wpr;
      const unsigned long long c = row[w], Lw = w > 0 ? row[w - 1] : 0ull,
                               Rw = w + 1 < wpr ? row[w + 1] : 0ull;
      out |= c;
      for (int d = 1; d <= buffer; ++d) {
        out |= (c << d) | (Lw >> (64 - d));  // cell x takes the flag of x − d
        out |= (c >> d) | (Rw << (64 - d));  // and of x + d
      }
    }
    const int64_t vcells = W - w * 64;  // cells of this word inside the domain
    if (vcells < 64) out &= (1ull << vcells) - 1ull;
    bits0[L.wofs[l] + t] = out;
  }
  const int c = block_count(out != 0ull, lds);
  if (threadIdx.x == 0) counts[blockIdx.x] = c;
}

// Exclusive prefix sums of the workgroups' counts (one workgroup of kScanWG
// threads), and where each level's nonzero words start (starts[n]: their total).
constexpr int kScanWG = 1024;
__global__ __launch_bounds__(kScanWG) void count_scan_kernel(const int32_t *counts, int64_t nblocks,
                                                             LevelWords L, int64_t *offsets,
                                                             int64_t *starts) {
  __shared__ int64_t part[kScanWG];
  int64_t carry = 0;
  for (int64_t base = 0; base < nblocks; base += kScanWG) {
    const int64_t i = base + threadIdx.x;
    const int64_t v = i < nblocks ? counts[i] : 0;
    part[threadIdx.x] = v;
    __syncthreads();
    for (int d = 1; d < kScanWG; d <<= 1) {  // Hillis–Steele inclusive scan
      const int64_t add = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
      __syncthreads();
      part[threadIdx.x] += add;
      __syncthreads();
    }
    if (i < nblocks) offsets[i] = carry + part[threadIdx.x] - v;
    carry += part[kScanWG - 1];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    for (int l = 0; l < L.n; ++l) starts[l] = L.first[l] < nblocks ? offsets[L.first[l]] : carry;
    starts[L.n] = carry;
  }
}

// Each nonzero word with its level-local index, in increasing word order per level.
__global__ __launch_bounds__(kWordWG) void compact_kernel(LevelWords L, const unsigned long long *bits0,
                                                          const int64_t *offsets, int64_t *idx,
                                                          unsigned long long *words) {
  __shared__ int lds[kWordWG / 64];
  int64_t t;
  const int l = word_level(L, blockIdx.x, t);
  const unsigned long long v = t < L.H[l] * L.wpr[l] ? bits0[L.wofs[l] + t] : 0ull;
  const unsigned long long m = __ballot(v != 0ull);
  if ((threadIdx.x & 63) == 0) lds[threadIdx.x >> 6] = __popcll(m);
  __syncthreads();
  if (v == 0ull) return;
  int64_t o = offsets[blockIdx.x];
  for (int w = 0; w < (int)(threadIdx.x >> 6); ++w) o += lds[w];
  const unsigned lane = threadIdx.x & 63;
  o += __popcll(m & ((1ull << lane) - 1ull));
  idx[o] = t;
  words[o] = v;
}

DevIndex dev_index(const lgd_amr &g) {
  return DevIndex{g.d_levels, g.d_bstart, g.d_bids, g.d_lo, g.d_blocks, g.num_levels()};
}

}  // namespace

using namespace lgd;

bool lgd_regrid_device_supported(const lgd_amr &g) { return g.max_levels() <= kMaxSampleDepth + 1; }

int lgd_regrid_flag_words(lgd_ctx *ctx, const lgd_amr &g, const double *tag_field,
                          const lgd_regrid_policy &pol, lgd_regrid_scratch &sc,
                          std::vector<lgd_flag_words> &flags) {
  const int levels = std::min(g.num_levels(), g.max_levels() - 1);
  flags.assign((size_t)g.num_levels(), lgd_flag_words{});
  if (levels <= 0) return LGD_OK;
  if (levels > kMaxWordLevels) return fail(ctx, LGD_ERR_UNSUPPORTED, "device regrid tags at most 8 levels");
  if (pol.buffer > 63)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "device regrid dilates by at most 63 cells");
  LevelWords L{};
  L.n = levels;
  size_t wtotal = 0;
  int64_t nblocks = 0;
  for (int src = 0; src < levels; ++src) {
    const lgd_box d = g.domain(src);
    L.W[src] = d.hi[0];
    L.H[src] = d.hi[1];
    L.wpr[src] = (d.hi[0] + 63) / 64;
    L.wofs[src] = (int64_t)wtotal;
    L.first[src] = nblocks;
    const int64_t nw = L.H[src] * L.wpr[src];
    wtotal += (size_t)nw;
    nblocks += (nw + kWordWG - 1) / kWordWG;
  }
  L.first[levels] = nblocks;
  if (nblocks > INT32_MAX) return fail(ctx, LGD_ERR_UNSUPPORTED, "regrid bitmask too large for one launch");
  // Tags (T), dilated flags (bits), per-workgroup counts and offsets, the compacted
  // words (idx, words) and where each level's start; a pinned host copy of the
  // starts and the compacted words.
  hipError_t e = hipSuccess;
  if (sc.wcap < wtotal || sc.bcap < (size_t)nblocks) {
    lgd_regrid_scratch_free(sc);
    const size_t wc = wtotal, bc = (size_t)nblocks;
    e = hipMalloc((void **)&sc.T, wc * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&sc.bits, wc * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&sc.idx, wc * sizeof(int64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&sc.words, wc * sizeof(uint64_t));
    if (e == hipSuccess) e = hipMalloc((void **)&sc.counts, bc * sizeof(int32_t));
    if (e == hipSuccess) e = hipMalloc((void **)&sc.offsets, (bc + kMaxWordLevels + 1) * sizeof(int64_t));
    if (e == hipSuccess)
      e = hipHostMalloc((void **)&sc.host, (2 * wc + kMaxWordLevels + 1) * sizeof(uint64_t),
                        hipHostMallocDefault);
    if (e != hipSuccess) {
      lgd_regrid_scratch_free(sc);
      return hip_fail(ctx, e, "regrid tag scratch");
    }
    sc.wcap = wc;
    sc.bcap = bc;
  }
  int64_t *starts = sc.offsets + sc.bcap;
  e = hipMemsetAsync(sc.T, 0, wtotal * sizeof(uint64_t), ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "regrid tag clear");
  int rc;
  for (int src = 0; src < levels; ++src) {
    unsigned long long *T = (unsigned long long *)(sc.T + L.wofs[src]);
    const lgd_level_plan &P = g.plans[(size_t)src];
    if (P.ntiles > 0) {
      tag_kernel<<<(unsigned)P.ntiles, dim3(kRTX, kRTY), 0, ctx->stream>>>(
          P.tiles, g.d_blocks, g.d_lo, tag_field, pol.threshold, T, L.wpr[src]);
      if ((rc = launch_check(ctx, "regrid tag"))) return rc;
    }
  }
  const int buffer = (int)std::max<int64_t>(pol.buffer, 0);
  dilate_kernel<<<(unsigned)nblocks, kWordWG, 0, ctx->stream>>>(
      L, (const unsigned long long *)sc.T, buffer, (unsigned long long *)sc.bits, sc.counts);
  if ((rc = launch_check(ctx, "regrid dilate"))) return rc;
  count_scan_kernel<<<1, kScanWG, 0, ctx->stream>>>(sc.counts, nblocks, L, sc.offsets, starts);
  if ((rc = launch_check(ctx, "regrid count scan"))) return rc;
  compact_kernel<<<(unsigned)nblocks, kWordWG, 0, ctx->stream>>>(
      L, (const unsigned long long *)sc.bits, sc.offsets, sc.idx, (unsigned long long *)sc.words);
  if ((rc = launch_check(ctx, "regrid compact"))) return rc;
  // Download the starts and, in the same round trip, as many words as the last
  // regrid had (with room to grow); the rest, if any, after the total is known.
  int64_t *h_starts = (int64_t *)sc.host;
  int64_t *h_idx = h_starts + kMaxWordLevels + 1;
  uint64_t *h_words = (uint64_t *)(h_idx + sc.wcap);
  const size_t guess = std::min(wtotal, std::max<size_t>(4096, 2 * sc.last_total));
  e = hipMemcpyAsync(h_starts, starts, (size_t)(levels + 1) * sizeof(int64_t), hipMemcpyDeviceToHost,
                     ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h_idx, sc.idx, guess * sizeof(int64_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(h_words, sc.words, guess * sizeof(uint64_t), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  const size_t total = e == hipSuccess ? (size_t)h_starts[levels] : 0;
  if (e == hipSuccess && total > guess) {
    e = hipMemcpyAsync(h_idx + guess, sc.idx + guess, (total - guess) * sizeof(int64_t),
                       hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess)
      e = hipMemcpyAsync(h_words + guess, sc.words + guess, (total - guess) * sizeof(uint64_t),
                         hipMemcpyDeviceToHost, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  }
  if (e != hipSuccess) return hip_fail(ctx, e, "regrid flag download");
  sc.last_total = total;
  for (int src = 0; src < levels; ++src) {
    const int64_t a = h_starts[src], b = h_starts[src + 1];
    flags[(size_t)src] = lgd_flag_words{h_idx + a, h_words + a, (size_t)(b - a)};
  }
  return LGD_OK;
}

int lgd_regrid_migrate_device(lgd_ctx *ctx, const lgd_amr &old_g, const double *const *old_f,
                              const lgd_amr &new_g, double *const *new_f, int nf) {
  if (!lgd_regrid_device_supported(new_g))
    return fail(ctx, LGD_ERR_UNSUPPORTED, "device migration supports at most 8 levels");
  DevRatios R{};
  R.cells[0] = old_g.cells[0];
  R.cells[1] = old_g.cells[1];
  for (size_t i = 0; i < old_g.ratios.size() && i <= (size_t)kMaxSampleDepth; ++i) R.r[i] = old_g.ratios[i];
  const DevIndex old = dev_index(old_g);
  int rc;
  // Two fields per launch.
  for (int f0 = 0; f0 < nf; f0 += 2) {
    MigrateFields F{};
    F.nf = std::min(2, nf - f0);
    for (int f = 0; f < F.nf; ++f) {
      F.old[f] = old_f[f0 + f];
      F.nw[f] = new_f[f0 + f];
    }
    for (int l = 1; l < new_g.num_levels(); ++l) {
      const lgd_level_plan &P = new_g.plans[(size_t)l];
      if (P.ntiles == 0) continue;
      migrate_kernel<<<(unsigned)P.ntiles, dim3(kRTX, kRTY), 0, ctx->stream>>>(
          P.tiles, new_g.d_blocks, new_g.d_lo, old, R, F, l);
      if ((rc = launch_check(ctx, "regrid migrate"))) return rc;
    }
  }
  return LGD_OK;
}

void lgd_regrid_scratch_free(lgd_regrid_scratch &sc) {
  (void)hipFree(sc.T);
  (void)hipFree(sc.bits);
  (void)hipFree(sc.idx);
  (void)hipFree(sc.words);
  (void)hipFree(sc.counts);
  (void)hipFree(sc.offsets);
  if (sc.host) (void)hipHostFree(sc.host);
  sc = lgd_regrid_scratch{};
}
