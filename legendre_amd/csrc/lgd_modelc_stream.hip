// Model C fused step, streaming form (the hot path for blocks up to 126 cells wide).
//
// Work decomposition: one wavefront advances one strip of rows of one block.  The
// 64 lanes cover the ghost-inclusive slab row (up to 128 columns) with two columns
// per lane (two coalesced 8-byte loads per lane per row and field, the halo
// cells folded into lane 0 and lane (nx+1)/2 by their addresses).  The wave
// marches up the strip keeping rows j-1, j, j+1 in registers while the loads of
// the next one or two rows are in flight.
//
// Face sharing: the Karma–Rappel face flux at (i+1/2, j) is the same expression,
// operand for operand, as the left-face flux of cell i+1, and the flux at
// (i, j+1/2) the same as the bottom-face flux of cell (i, j+1)
// (src/discretization/finite_volume.rs:261-280), so each face is evaluated once:
// y-faces carry over in registers from one row to the next, x-faces come from the
// neighbouring lane by a DPP wavefront shift.  Results are bitwise those of the
// per-cell evaluation; divisions per cell drop from 10 to 6.
//
// Halos: the ghost ring of the slab is never read.  The two halo columns come
// from the left/right neighbour block's interior (or this block's own edge column
// at a mirror face, or the wrapped block at a periodic face) and the halo rows from
// the block below/above, resolved per dimension in global coordinates — the values
// fill_ghosts_mirror would have written (src/geometry/cartesian.rs:439-526).
//
// Block pairs: a block at most 64 cells wide leaves half the wave idle, so such
// blocks can also run two to a wave: blocks b and b + 1, x-neighbours in one
// block row, advanced over the same rows.  Lane L holds interior columns 2L and
// 2L + 1 of the 2·nx-wide pair; the face between the two blocks comes from the
// lane shift like any other, and the pair's two outer halo columns are loaded by
// lanes 0 and nx - 1 into one extra register per row and field.  The x-face left
// of the pair's first column is the one face no lane shift supplies: every lane
// evaluates one more face per row, and lane 0 keeps it.
//
// Scheduling: workgroups of four independent waves; workgroup ids are remapped so
// that each of the eight XCDs sweeps one contiguous range of blocks in order, so
// neighbouring blocks (whose edge columns a wave reads) are in flight on the same
// XCD and share its L2.
#include <cmath>

#include "lgd_amr.h"
#include "lgd_internal.h"
#include "lgd_stream_ops.h"

namespace {

using namespace lgd_stream;

constexpr int kWavesPerWG = 4;
constexpr int kXcds = 8;

// A RungeKutta4 stage on the streaming kernel (kStage 1..4; src/integrators/rk4.rs):
// the stencil reads the stage input (phi, u of StreamArgs; stage 1: y itself);
// with k_s = f(input), stages 1-3 write the next input y + c_s·k_s (to phi_n, u_n)
// and the running sum A of y + w1·k1 + w2·k2 + w3·k3 (stage 1 starts it from y,
// stages 2-3 add to it in place), and stage 4 writes y = A + w4·k4.
using Rk4Stage = lgd_rk4_stage;

struct StreamArgs {
  // The domains (independent Cartesian grids sharing the block shape) and,
  // per block, which domain it belongs to; a single grid passes its domain in
  // d0 with domains == nullptr.
  lgd_domain d0;
  const lgd_domain *domains;
  const int32_t *block_domain;
  const double *phi;
  const double *u;
  double *phi_n;
  double *u_n;
  double dt;
  uint64_t salt;
  const double *salt_src;  // time to take the salt from (graph capture), or null
  uint64_t stream;
  int64_t nblocks;      // blocks (block pairs) in this launch
  int64_t block_begin;  // first block of this launch
  int32_t nx, ny;      // block interior extent
  int32_t strips;      // strips per block
  int32_t rows;        // rows per strip
  int32_t nwg;         // workgroups of the step in the launch
  Rk4Stage rk;         // kStage >= 1 only
  lgd_stream_side_fill side;  // kSide only: written by the first side_wg workgroups
  int32_t side_wg;
};

__device__ __forceinline__ int64_t wrap_or_reflect(int64_t g, int64_t n, int periodic) {
  if (g < 0) return periodic ? g + n : -1 - g;
  if (g >= n) return periodic ? g - n : 2 * n - 1 - g;
  return g;
}

// Element offset, from a field's base, of the first column of the slab row that
// holds row r (-1 <= r <= ny) of block row by, in block column 0 of that block row.
// Rows -1 and ny are halo rows: the block below/above (exchange or periodic wrap)
// or this block's own edge row (mirror, reflect_coord in src/geometry/boundary.rs:82-90).
__device__ __forceinline__ int64_t row_base(const StreamArgs &a, const lgd_domain &D, int64_t by,
                                            int r, int64_t slab, int sx) {
  int64_t byr = by, jr = r;
  if (r < 0 || r >= a.ny) {
    int64_t gj = by * a.ny + r;
    int64_t rj = wrap_or_reflect(gj, (int64_t)D.nby * a.ny, D.per_y);
    byr = rj / a.ny;
    jr = rj - byr * a.ny;
  }
  return (D.block0 + byr * D.nbx) * slab + (jr + 1) * (int64_t)sx;
}

// Two consecutive f64 columns as one 16-byte vector (for nontemporal stores).
typedef double F64x2 __attribute__((ext_vector_type(2)));

// One row of a field as this lane holds it: two consecutive columns and, for a
// block pair, the halo column value lanes 0 and nx - 1 need.
struct RowVal {
  double v0, v1, h;
};

template <bool kNoise, bool kVec, bool kTable, int kDepth, bool kPair, int kStage, bool kSide = false>
__global__ __launch_bounds__(64 * kWavesPerWG) void modelc_stream_kernel(StreamArgs a) {
  static_assert(kStage == 0 || (!kNoise && !kPair && !kVec), "RK4 stages: no noise, one block");
  static_assert(!kSide || (kStage == 0 && !kTable), "a side fill rides on a one-domain step");
  int w = blockIdx.x;
  if constexpr (kSide) {
    // The side fill first (α = 0: the corner sum of this launch's inputs, which it
    // only reads; copies within the filled level).
    if (w < a.side_wg) {
      const int64_t t = (int64_t)w * blockDim.x + threadIdx.x;
      if (t >= a.side.nops) return;
      const lgd_fill_op op = a.side.ops[t];
      if (op.clen == 0) {
        a.side.lvl[0][op.target] = a.side.lvl[0][op.src];
        a.side.lvl[1][op.target] = a.side.lvl[1][op.src];
        return;
      }
      lgd_corner k[4];
#pragma unroll
      for (int c = 0; c < 4; ++c) k[c] = a.side.corners[4 * t + c];
      double vp = 0.0, vu = 0.0;
#pragma unroll
      for (int c = 0; c < 4; ++c)
        if (c < op.clen) {
          vp = vp + a.phi[k[c].src] * k[c].weight;
          vu = vu + a.u[k[c].src] * k[c].weight;
        }
      a.side.lvl[0][op.target] = vp;
      a.side.lvl[1][op.target] = vu;
      return;
    }
    w -= a.side_wg;
  }
  // XCD-aware workgroup order: XCD x (= blockIdx.x % 8) takes one contiguous
  // range of logical workgroups (the side fill's count is a multiple of 8).
  const int xcd = w % kXcds, idx = w / kXcds;
  const int per = a.nwg / kXcds, rem = a.nwg % kXcds;
  const int lw = xcd * per + min(xcd, rem) + idx;
  // Wave-uniform by construction; readfirstlane lets the compiler keep the
  // block/strip arithmetic on the scalar unit.
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int64_t item = (int64_t)lw * kWavesPerWG + wave;
  if (item >= a.nblocks * a.strips) return;  // whole wave exits together
  // One block, or (kPair) blocks b and b + 1 of one block row.
  const int64_t b = a.block_begin + (kPair ? 2 : 1) * (item / a.strips);
  const int strip = (int)(item % a.strips);
  const int j0 = strip * a.rows;
  const int j1 = min(j0 + a.rows, a.ny);
  // The block's domain (wave-uniform): its topology, constants and noise key.
  const lgd_domain D =
      kTable ? load_uniform(a.domains + __builtin_amdgcn_readfirstlane(a.block_domain[b])) : a.d0;
  const int64_t rb = b - D.block0;  // BlockId within the domain
  const int64_t bx = rb % D.nbx, by = rb / D.nbx;
  const int nx = a.nx, sx = nx + 2;
  const int64_t slab = (int64_t)sx * (a.ny + 2);
  const lgd_kr_consts &k = D.k;

  // Halo columns: global column just left / right of this block (pair), resolved.
  const int64_t NX = (int64_t)D.nbx * nx;
  int64_t gl = wrap_or_reflect(bx * nx - 1, NX, D.per_x);
  int64_t gr = wrap_or_reflect((bx + (kPair ? 2 : 1)) * nx, NX, D.per_x);
  const int64_t bxl = gl / nx, bxr = gr / nx;
  const int il = (int)(gl - bxl * nx), ir = (int)(gr - bxr * nx);
  const int64_t lcol = bxl * slab + (il + 1), rcol = bxr * slab + (ir + 1);

  // Per-lane column offsets (from the start of a slab row in block column 0),
  // valid for every row, the lane's output offset within a slab row (obase) and
  // the interior column of its first cell (kc0, for the cell key).
  int64_t col0, col1, hcol = 0, obase;
  bool in0, in1;    // the lane's columns are interior cells
  bool st0, st1;    // the lane's columns lie in the slab row (stored by the step)
  int kc0, seg = 0;
  if constexpr (kPair) {
    // Lanes 0..nx-1 hold the pair's 2·nx interior columns (nx even, so a lane's
    // two columns are in one block); lanes past them re-read block b's first
    // columns and store nothing.  Lane 0 also reads the left halo column, lane
    // nx - 1 the right one.
    const bool act = lane < nx;
    seg = 2 * lane >= nx ? 1 : 0;
    const int ic = act ? 2 * lane - seg * nx : 0;
    col0 = (bx + seg) * slab + (ic + 1);
    col1 = col0 + 1;
    hcol = lane == 0 ? lcol : (lane == nx - 1 ? rcol : col0);
    obase = (b + seg) * slab + (ic + 1);
    in0 = act;
    in1 = act;
    st0 = act;
    st1 = act;
    kc0 = ic;
  } else {
    // Lane 0 reads the left halo cell in its first column, lane (nx+1)/2 the
    // right halo cell in its first (nx odd) or second (nx even) column, lanes
    // past the slab row re-read its last column (their values are never stored
    // and never an interior column's neighbour).
    const int hr_lane = (nx + 1) >> 1, hr_slot = (nx + 1) & 1;
    const int c0 = 2 * lane, c1 = c0 + 1;
    col0 = lane == 0 ? lcol
           : (lane == hr_lane && hr_slot == 0) ? rcol
                                               : bx * slab + min(c0, sx - 1);
    col1 = (lane == hr_lane && hr_slot == 1) ? rcol : bx * slab + min(c1, sx - 1);
    obase = b * slab + c0;
    in0 = c0 >= 1 && c0 <= nx;
    in1 = c1 <= nx;
    st0 = c0 < sx;
    st1 = c1 < sx;
    kc0 = c0 - 1;
  }

  // Wiener(0) amplitude b on interior cells; zero amplitude draws nothing
  // (src/core/driver.rs:144-146).
  const bool noisy = kNoise && D.noise_scale != 0.0;
  uint64_t block_key = 0;
  if (noisy) {
    const uint64_t salt = lgd_salt(a.salt, a.salt_src);
    block_key = lgd_mix_key3(D.seed, salt, a.stream, (uint64_t)rb);
    if constexpr (kPair) {
      const uint64_t key1 = lgd_mix_key3(D.seed, salt, a.stream, (uint64_t)(rb + 1));
      block_key = seg ? key1 : block_key;
    }
  }

  auto fetch = [&](const double *f, int r, bool edge) {
    const double *row = f + row_base(a, D, by, r, slab, sx);
    RowVal v{row[col0], row[col1], 0.0};
    if constexpr (kPair) {
      if (edge) v.h = row[hcol];
    }
    return v;
  };
  // The neighbour column's value: from the lane to the left / right, or (block
  // pair) the halo column at the pair's edges.  The shift runs on every lane.
  auto lft = [&](double v, double edge) {
    const double s = from_left(v);
    if constexpr (kPair) return lane == 0 ? edge : s;
    return s;
  };
  auto rgt = [&](double v, double edge) {
    const double s = from_right(v);
    if constexpr (kPair) return lane == nx - 1 ? edge : s;
    return s;
  };

  // Rows j0-1, j0, j0+1 of phi and u (h*: the pair's halo column, rows m, c, p).
  double pm0, pm1, pc0, pc1, pp0, pp1, um0, um1, uc0, uc1, up0, up1;
  double hm = 0.0, hc = 0.0, hp = 0.0, huc = 0.0, hup = 0.0;
  {
    RowVal r0 = fetch(a.phi, j0 - 1, true), r1 = fetch(a.phi, j0, true),
           r2 = fetch(a.phi, j0 + 1, true);
    RowVal r3 = fetch(a.u, j0 - 1, false), r4 = fetch(a.u, j0, true), r5 = fetch(a.u, j0 + 1, true);
    pm0 = r0.v0; pm1 = r0.v1;
    pc0 = r1.v0; pc1 = r1.v1;
    pp0 = r2.v0; pp1 = r2.v1;
    um0 = r3.v0; um1 = r3.v1;
    uc0 = r4.v0; uc1 = r4.v1;
    up0 = r5.v0; up1 = r5.v1;
    if constexpr (kPair) {
      hm = r0.h; hc = r1.h; hp = r2.h;
      huc = r4.h; hup = r5.h;
    }
  }
  const double q = 0.25;
  double pmR = rgt(pm0, hm);
  double pcL = lft(pc1, hc), pcR = rgt(pc0, hc);
  double ppL = lft(pp1, hp), ppR = rgt(pp0, hp);
  double ucL = lft(uc1, huc), ucR = rgt(uc0, huc);
  // Flux through the bottom face of row j0 (the top face of row j0-1).
  double jb0, jb1;
  {
    double pmL = lft(pm1, hm);
    jb0 = flux_y(k, q * (((pc1 - pcL) + pm1) - pmL), pc0 - pm0);
    jb1 = flux_y(k, q * (((pcR - pc0) + pmR) - pm0), pc1 - pm1);
  }

  // Row pipeline: at row j the loads of row j+1+kDepth are issued; the row that
  // becomes the top neighbour of row j+1 (issued kDepth-1 rows earlier) is moved
  // into place only after row j is computed, and row j is stored last, so the
  // wave waits on loads issued one (kDepth 1) or two (kDepth 2) rows of compute
  // ago.  Two buffers alternate (the loop is unrolled by two) so no register
  // with a load in flight is copied.
  auto body = [&](int j, RowVal &ip, RowVal &iu, RowVal &cp, RowVal &cu) {
    // RK4 stage: the pointwise operands at row j (interior columns only), used
    // once the row's vector field is known.  They are issued before the rows
    // loaded ahead: loads complete in order, so waiting for them then never waits
    // for the look-ahead.
    const int64_t o = obase + (int64_t)(j + 1) * sx;
    double ry[4] = {0.0, 0.0, 0.0, 0.0}, ra[4] = {0.0, 0.0, 0.0, 0.0};
    if constexpr (kStage == 2 || kStage == 3) {
      if (in0) {
        ry[0] = a.rk.y_phi[o];
        ry[2] = a.rk.y_u[o];
      }
      if (in1) {
        ry[1] = a.rk.y_phi[o + 1];
        ry[3] = a.rk.y_u[o + 1];
      }
    }
    if constexpr (kStage >= 2) {
      if (in0) {
        ra[0] = a.rk.acc_phi[o];
        ra[2] = a.rk.acc_u[o];
      }
      if (in1) {
        ra[1] = a.rk.acc_phi[o + 1];
        ra[3] = a.rk.acc_u[o + 1];
      }
    }
    // Issued on every row (past the strip's last needed row it re-reads row j1,
    // unused) so the wait for the consumed row never depends on a branch.
    const int rn = min(j + 1 + kDepth, j1);
    ip = fetch(a.phi, rn, true);
    iu = fetch(a.u, rn, true);
    // x-faces right of c0 and c1; the face left of c0 is the left lane's right-of-c1.
    double jx0 = flux_x(k, pc1 - pc0, q * (((pp1 - pm1) + pp0) - pm0));
    double jx1 = flux_x(k, pcR - pc1, q * (((ppR - pmR) + pp1) - pm1));
    double jxl0 = from_left(jx1);
    if constexpr (kPair) {
      // The face between the left halo column and the pair's first column
      // (derx_l, dery_l of src/discretization/finite_volume.rs:261-280).
      const double je = flux_x(k, pc0 - hc, q * (((pp0 - pm0) + hp) - hm));
      jxl0 = lane == 0 ? je : jxl0;
    }
    // y-faces above c0 and c1.
    double jt0 = flux_y(k, q * (((pp1 - ppL) + pc1) - pcL), pp0 - pc0);
    double jt1 = flux_y(k, q * (((ppR - pp0) + pcR) - pc0), pp1 - pc1);
    double f0 = (((jx0 - jxl0) + jt0) - jb0) * k.inv_h2;
    double f1 = (((jx1 - jx0) + jt1) - jb1) * k.inv_h2;
    double n0, n1, v0, v1, d0, d1, e0, e1;
    cell_update(k, f0, pc0, pc1, pcL, pp0, pm0, uc0, uc1, ucL, up0, um0, a.dt, n0, v0, d0, e0);
    cell_update(k, f1, pc1, pcR, pc0, pp1, pm1, uc1, ucR, uc0, up1, um1, a.dt, n1, v1, d1, e1);
    if constexpr (kStage == 1) {  // the input is y: row j's own values (before the rotation)
      ry[0] = pc0;
      ry[1] = pc1;
      ry[2] = uc0;
      ry[3] = uc1;
    }
    if (noisy) {
      // Wiener(0) on phi after the drift (src/core/driver.rs:139-158); cell key =
      // interior linear index (src/geometry/cartesian.rs:599-619).
      uint64_t rowkey = (uint64_t)j * (uint64_t)nx;
      if (in0) n0 = n0 + D.noise_scale * lgd_standard_normal(lgd_splitmix64(block_key ^ (rowkey + (uint64_t)kc0)));
      if (in1) n1 = n1 + D.noise_scale * lgd_standard_normal(lgd_splitmix64(block_key ^ (rowkey + (uint64_t)(kc0 + 1))));
    }
    jb0 = jt0;
    jb1 = jt1;
    // The output's ghost columns receive the halo values read for this step, so
    // every store covers whole cache lines (no read-modify-write of the line
    // shared with the ghost cell).  Ghost cells carry no state between steps:
    // they are refilled before any stencil reads them (src/core/driver.rs:128-130).
    if (!in0) {
      n0 = pc0;
      v0 = uc0;
    }
    if (!in1) {
      n1 = pc1;
      v1 = uc1;
    }
    // Block pair: the ghost columns of both output slabs take this row's halo
    // values, as the single-block form stores them (lane 0: left halo; lane
    // nx/2 - 1: block b's last column, block b + 1's left ghost; lane nx/2: block
    // b + 1's first column, block b's right ghost; lane nx - 1: right halo).
    int64_t go = -1;
    double gp = 0.0, gu = 0.0;
    if constexpr (kPair) {
      const int64_t ro = (int64_t)(j + 1) * sx;
      if (lane == 0) {
        go = b * slab + ro;
        gp = hc;
        gu = huc;
      } else if (lane == nx / 2 - 1) {
        go = (b + 1) * slab + ro;
        gp = pc1;
        gu = uc1;
      } else if (lane == nx / 2) {
        go = b * slab + ro + (nx + 1);
        gp = pc0;
        gu = uc0;
      } else if (lane == nx - 1) {
        go = (b + 1) * slab + ro + (nx + 1);
        gp = hc;
        gu = huc;
      }
    }
    if (j + 1 < j1) {
      pm0 = pc0; pm1 = pc1; pmR = pcR;
      pc0 = pp0; pc1 = pp1; pcL = ppL; pcR = ppR;
      pp0 = cp.v0; pp1 = cp.v1;
      if constexpr (kPair) {
        hm = hc; hc = hp; hp = cp.h;
        huc = hup; hup = cu.h;
      }
      ppL = lft(pp1, hp);
      ppR = rgt(pp0, hp);
      um0 = uc0; um1 = uc1;
      uc0 = up0; uc1 = up1;
      up0 = cu.v0; up1 = cu.v1;
      ucL = lft(uc1, huc);
      ucR = rgt(uc0, huc);
    }
    if constexpr (kStage >= 1 && kStage <= 3) {
      // The running sum (started from y at stage 1) and the next stage's input
      // y + c_s·k_s, interior cells only (the ghost ring of these buffers is
      // never read).
      const double c = a.rk.coef, w = a.rk.w;
      const double s00 = kStage == 1 ? ry[0] : ra[0], s01 = kStage == 1 ? ry[1] : ra[1];
      const double s10 = kStage == 1 ? ry[2] : ra[2], s11 = kStage == 1 ? ry[3] : ra[3];
      if (in0) {
        a.rk.acc_phi[o] = s00 + w * d0;
        a.rk.acc_u[o] = s10 + w * e0;
        a.phi_n[o] = ry[0] + c * d0;
        a.u_n[o] = ry[2] + c * e0;
      }
      if (in1) {
        a.rk.acc_phi[o + 1] = s01 + w * d1;
        a.rk.acc_u[o + 1] = s11 + w * e1;
        a.phi_n[o + 1] = ry[1] + c * d1;
        a.u_n[o + 1] = ry[3] + c * e1;
      }
      return;
    }
    if constexpr (kStage == 4) {
      // y = A + (dt/6)·k4, A = y + (dt/6)·k1 + (dt/3)·k2 + (dt/3)·k3 summed in that
      // order (src/integrators/rk4.rs), in place: y is written only at its own cell.
      const double w = a.rk.w;
      if (in0) {
        a.rk.y_phi[o] = ra[0] + w * d0;
        a.rk.y_u[o] = ra[2] + w * e0;
      }
      if (in1) {
        a.rk.y_phi[o + 1] = ra[1] + w * d1;
        a.rk.y_u[o + 1] = ra[3] + w * e1;
      }
      return;
    }
    // Streaming (nontemporal) stores: the new state is not read again in this
    // launch, so it need not displace the halo lines neighbouring waves re-read.
    if (kVec) {
      if (st1) {
        __builtin_nontemporal_store(F64x2{n0, n1}, reinterpret_cast<F64x2 *>(a.phi_n + o));
        __builtin_nontemporal_store(F64x2{v0, v1}, reinterpret_cast<F64x2 *>(a.u_n + o));
      }
    } else {
      if (st0) {
        __builtin_nontemporal_store(n0, a.phi_n + o);
        __builtin_nontemporal_store(v0, a.u_n + o);
      }
      if (st1) {
        __builtin_nontemporal_store(n1, a.phi_n + o + 1);
        __builtin_nontemporal_store(v1, a.u_n + o + 1);
      }
    }
    if (kPair && go >= 0) {
      __builtin_nontemporal_store(gp, a.phi_n + go);
      __builtin_nontemporal_store(gu, a.u_n + go);
    }
  };
  RowVal ap{}, au{}, bp{}, bu{};
  if (kDepth == 2) {
    ap = fetch(a.phi, min(j0 + 2, j1), true);
    au = fetch(a.u, min(j0 + 2, j1), true);
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

template <int kDepth>
void launch_variant(dim3 grid, dim3 block, hipStream_t st, const StreamArgs &a, bool noise,
                    bool vec, bool table, bool pair) {
  if (pair) {  // block pairs store column by column (odd slab columns start a lane's pair)
#define LGD_PAIR_LAUNCH(N, T) modelc_stream_kernel<N, false, T, kDepth, true, 0><<<grid, block, 0, st>>>(a)
    if (noise) {
      if (table) LGD_PAIR_LAUNCH(true, true); else LGD_PAIR_LAUNCH(true, false);
    } else {
      if (table) LGD_PAIR_LAUNCH(false, true); else LGD_PAIR_LAUNCH(false, false);
    }
#undef LGD_PAIR_LAUNCH
    return;
  }
#define LGD_STREAM_LAUNCH(N, V, T) modelc_stream_kernel<N, V, T, kDepth, false, 0><<<grid, block, 0, st>>>(a)
  if (noise) {
    if (vec) {
      if (table) LGD_STREAM_LAUNCH(true, true, true); else LGD_STREAM_LAUNCH(true, true, false);
    } else {
      if (table) LGD_STREAM_LAUNCH(true, false, true); else LGD_STREAM_LAUNCH(true, false, false);
    }
  } else {
    if (vec) {
      if (table) LGD_STREAM_LAUNCH(false, true, true); else LGD_STREAM_LAUNCH(false, true, false);
    } else {
      if (table) LGD_STREAM_LAUNCH(false, false, true); else LGD_STREAM_LAUNCH(false, false, false);
    }
  }
#undef LGD_STREAM_LAUNCH
}

// The step with a side fill (one domain, no table).
template <int kDepth>
void launch_side(dim3 grid, dim3 block, hipStream_t st, const StreamArgs &a, bool noise, bool vec,
                 bool pair) {
  if (pair) {
    if (noise) modelc_stream_kernel<true, false, false, kDepth, true, 0, true><<<grid, block, 0, st>>>(a);
    else modelc_stream_kernel<false, false, false, kDepth, true, 0, true><<<grid, block, 0, st>>>(a);
  } else if (vec) {
    if (noise) modelc_stream_kernel<true, true, false, kDepth, false, 0, true><<<grid, block, 0, st>>>(a);
    else modelc_stream_kernel<false, true, false, kDepth, false, 0, true><<<grid, block, 0, st>>>(a);
  } else {
    if (noise) modelc_stream_kernel<true, false, false, kDepth, false, 0, true><<<grid, block, 0, st>>>(a);
    else modelc_stream_kernel<false, false, false, kDepth, false, 0, true><<<grid, block, 0, st>>>(a);
  }
}

}  // namespace

bool lgd_stream_pairable(int64_t nx, int64_t nbx) {
  return nx % 2 == 0 && nx >= 4 && nx <= 64 && nbx % 2 == 0;
}

// Streaming launch over blocks [block_begin, block_begin + nblocks) of a set of
// domains sharing the block shape nx × ny.  Returns LGD_ERR_UNSUPPORTED when the
// block is wider than one wavefront row (126 interior cells); the caller then
// uses the per-cell kernel.  With `pair`, two x-neighbouring blocks share a wave:
// the caller guarantees lgd_stream_pairable for every domain, an even block_begin
// and nblocks, and even first blocks of every domain, so that blocks 2i and 2i + 1
// of the launch are neighbours in one block row.
int lgd_modelc_stream_launch(lgd_ctx *ctx, int64_t nx, int64_t ny, const lgd_domain &d0,
                             const lgd_domain *domains, const int32_t *block_domain,
                             int64_t block_begin, int64_t nblocks, const double *phi,
                             const double *u, double *phi_n, double *u_n, double dt, bool noise,
                             uint64_t salt, uint64_t stream, int strips, int prefetch, bool pair,
                             const lgd_stream_side_fill *side) {
  if (nx + 2 > 128 || nx < 1 || ny < 1) return LGD_ERR_UNSUPPORTED;
  if (side && domains) return lgd::fail(ctx, LGD_ERR_INVALID_ARGUMENT, "a side fill needs one domain");
  if (nblocks <= 0 && !(side && side->nops > 0)) return LGD_OK;
  if (pair && (!lgd_stream_pairable(nx, 2) || block_begin % 2 != 0 || nblocks % 2 != 0))
    return lgd::fail(ctx, LGD_ERR_INVALID_ARGUMENT, "block pairs need an even block count and width");
  StreamArgs a;
  a.d0 = d0;
  a.domains = domains;
  a.block_domain = block_domain;
  a.phi = phi;
  a.u = u;
  a.phi_n = phi_n;
  a.u_n = u_n;
  a.dt = dt;
  a.salt = salt;
  a.salt_src = ctx->salt_src;
  a.stream = stream;
  a.nblocks = pair ? nblocks / 2 : nblocks;
  a.block_begin = block_begin;
  a.nx = (int32_t)nx;
  a.ny = (int32_t)ny;
  if (strips < 1) strips = 1;
  if (strips > a.ny) strips = a.ny;
  a.rows = (a.ny + strips - 1) / strips;
  a.strips = (a.ny + a.rows - 1) / a.rows;
  int64_t items = a.nblocks * a.strips;
  int64_t nwg = (items + kWavesPerWG - 1) / kWavesPerWG;
  const int64_t swg =
      side && side->nops > 0 ? ((side->nops + 64 * kWavesPerWG - 1) / (64 * kWavesPerWG) + 7) / 8 * 8 : 0;
  if (nwg + swg > INT32_MAX) return LGD_ERR_UNSUPPORTED;
  a.nwg = (int32_t)nwg;
  a.side = side ? *side : lgd_stream_side_fill{};
  a.side_wg = (int32_t)swg;
  const bool vec = ((nx + 2) % 2) == 0;
  const bool table = domains != nullptr;
  const int depth = prefetch == 1 ? 1 : 2;
  dim3 grid((unsigned)(nwg + swg)), block(64 * kWavesPerWG);
  if (swg > 0) {
    if (depth == 2) launch_side<2>(grid, block, ctx->stream, a, noise, vec, pair);
    else launch_side<1>(grid, block, ctx->stream, a, noise, vec, pair);
  } else if (depth == 2) {
    launch_variant<2>(grid, block, ctx->stream, a, noise, vec, table, pair);
  } else {
    launch_variant<1>(grid, block, ctx->stream, a, noise, vec, table, pair);
  }
  return lgd::launch_check(ctx, "modelc_stream");
}

int lgd_modelc_stream_rk4_launch(lgd_ctx *ctx, int64_t nx, int64_t ny, const lgd_domain &d0,
                                 int64_t nblocks, const double *pin, const double *uin,
                                 double *pout, double *uout, int stage, const lgd_rk4_stage &rk,
                                 int strips, int prefetch) {
  if (nx + 2 > 128 || nx < 1 || ny < 1) return LGD_ERR_UNSUPPORTED;
  if (stage < 1 || stage > 4) return lgd::fail(ctx, LGD_ERR_INVALID_ARGUMENT, "RK4 stage is 1..4");
  if (nblocks <= 0) return LGD_OK;
  StreamArgs a{};
  a.d0 = d0;
  a.domains = nullptr;
  a.block_domain = nullptr;
  a.phi = pin;
  a.u = uin;
  a.phi_n = pout;
  a.u_n = uout;
  a.dt = 0.0;  // the stage update is the caller's (rk), not x + dt·f
  a.salt = 0;
  a.salt_src = nullptr;
  a.stream = 0;
  a.nblocks = nblocks;
  a.block_begin = 0;
  a.nx = (int32_t)nx;
  a.ny = (int32_t)ny;
  a.rk = rk;
  if (strips < 1) strips = 1;
  if (strips > a.ny) strips = a.ny;
  a.rows = (a.ny + strips - 1) / strips;
  a.strips = (a.ny + a.rows - 1) / a.rows;
  const int64_t nwg = (nblocks * a.strips + kWavesPerWG - 1) / kWavesPerWG;
  if (nwg > INT32_MAX) return LGD_ERR_UNSUPPORTED;
  a.nwg = (int32_t)nwg;
  const dim3 grid((unsigned)nwg), block(64 * kWavesPerWG);
  hipStream_t st = ctx->stream;
#define LGD_RK4_LAUNCH(D, S) modelc_stream_kernel<false, false, false, D, false, S><<<grid, block, 0, st>>>(a)
  if (prefetch == 1) {
    if (stage == 1) LGD_RK4_LAUNCH(1, 1); else if (stage == 2) LGD_RK4_LAUNCH(1, 2);
    else if (stage == 3) LGD_RK4_LAUNCH(1, 3); else LGD_RK4_LAUNCH(1, 4);
  } else {
    if (stage == 1) LGD_RK4_LAUNCH(2, 1); else if (stage == 2) LGD_RK4_LAUNCH(2, 2);
    else if (stage == 3) LGD_RK4_LAUNCH(2, 3); else LGD_RK4_LAUNCH(2, 4);
  }
#undef LGD_RK4_LAUNCH
  return lgd::launch_check(ctx, "modelc_stream_rk4");
}
