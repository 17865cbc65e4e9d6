// Reverse Model C step in one streaming pass.
//
// The same reverse step as the three-pass form in lgd_adjoint.hip (cells →
// faces → gather), restructured like the forward streaming kernel: one wavefront
// advances one strip of rows of one column tile, one column per lane, rows held in
// registers.  Tiles are laid over the global grid, not over blocks: every state
// address is resolved per lane (block, ghost image), so a wave's 64 lanes cover
// global columns x0−2 … x0+61 for up to 60 owned columns x0 … x0+W−1 (the owned
// columns, one halo column each side whose R and S the owned cells' faces need, and
// one more each side for that column's stencil).
//
// At row r the wave computes, once each: the x-faces of row r and the y-face
// between rows r and r+1 (values for F(r) and the derivative factors), F, A and the
// cell sensitivities R = ∂L/∂(numerator of F), S = ∂L/∂A at row r, and then adds
// every contribution that the faces of row r, the y-face between r−1 and r (whose
// factors were kept from the previous row) and the centres of row r make to rows
// r−1, r, r+1 — column neighbours by DPP lane shifts.  Row r−1 has then received
// everything and is written.  Rows r from y0−1 to y1 are processed for owned rows
// [y0, y1) (halo rows' R and faces are recomputed by each strip).
//
// Ownership: a wave writes only its own cells; contributions that land on halo
// positions owned by another wave are dropped (that wave computes them itself).
// At a mirror face of the domain the ghost positions −1 and N are images of the
// edge cells: their accumulated contributions are folded into the edge cells, and
// their own R, S (no equation lives there) are zero.  The θ partials of cells and
// faces are summed per wave in a fixed order.
//
// Arithmetic: the face and centre functions share one reciprocal each (1/|∇φ|²)
// and F, R, S one reciprocal of A, and products and sums may contract to fused
// multiply-adds here (this file only): the derivative needs no bitwise match with
// the forward, and agrees with the three-pass form to rounding.
#include <cmath>
#include <type_traits>

#include "lgd_internal.h"

namespace {

__device__ __forceinline__ int resolve_i(int i, int n, int per) {
  if (i < 0) return per ? i + n : -1 - i;
  if (i >= n) return per ? i - n : 2 * n - 1 - i;
  return i;
}

__device__ __forceinline__ double from_left(double v) {  // value of lane − 1 (0 at lane 0)
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xf, 0xf, true);
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}
__device__ __forceinline__ double from_right(double v) {  // value of lane + 1 (0 at lane 63)
  int lo = __double2loint(v), hi = __double2hiint(v);
  lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xf, 0xf, true);
  hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xf, 0xf, true);
  return __hiloint2double(hi, lo);
}

// face(dx, dy) (finite_volume.rs:129-142) and its derivatives w.r.t. dx, dy and ε₄
// (through ā = 1 − 3ε₄, ε′ = 3ε₄/ā, a₁₂ = 4āε′: dā/dε₄ = −3, dε′/dε₄ = 3/ā²,
// da₁₂/dε₄ = 12), with s = dx² + dy², m = s², Q = dx⁴ + dy⁴, H = dx·dy·(dx² − dy²):
//   a = ā(1 + ε′Q/m),  a′ = −a₁₂H/m,
//   ∂a/∂dx = 4āε′·dx·dy²(dx² − dy²)/s³,  ∂a/∂dy = 4āε′·dy·dx²(dy² − dx²)/s³,
//   ∂a′/∂dx = −a₁₂(H_x·s − 4dx·H)/s³,    ∂a′/∂dy = −a₁₂(H_y·s − 4dy·H)/s³,
//   ∂a/∂ε₄ = −3(1 + ε′Q/m) + (3/ā)Q/m,   ∂a′/∂ε₄ = −12H/m;
// the flat branch (m <= tol) has a = ā, a′ = 0, ∂a/∂ε₄ = −3 and no other
// derivative.
struct FD {
  double a, da, a_x, a_y, a_e, da_x, da_y, da_e;
};

// 1/x for normal positive x: the hardware estimate and one Newton step.
__device__ __forceinline__ double recip(double x) {
  double r = __builtin_amdgcn_rcp(x);
  const double e = fma(-x, r, 1.0);
  return fma(r, e, r);
}

// With is = 1/s every term above carries a factor is, except a = ā(1 + ε′Q·is²) and
// ∂a/∂ε₄; is = 0 therefore gives exactly the flat branch, which needs no other
// select.
template <bool kAngular>
__device__ __forceinline__ FD fderiv(const lgd_kr_consts &k, double tol, double dx, double dy) {
#pragma clang fp contract(fast)
  FD f;
  const double dx2 = dx * dx, dy2 = dy * dy;
  const double s = dx2 + dy2;
  const double m = s * s;
  const double is = m <= tol ? 0.0 : recip(s);
  const double im = is * is;
  const double is3 = im * is;
  const double Q = dx2 * dx2 + dy2 * dy2;
  const double d2 = dx2 - dy2;
  const double qm = Q * im;
  f.a = k.a_bar * (1.0 + k.eps_prime * qm);
  // dx·dy²·(dx² − dy²) = dy·H and dy·dx²·(dx² − dy²) = dx·H.
  const double H = dx * dy * d2;
  const double cH = ((4.0 * k.a_bar * k.eps_prime) * is3) * H;
  f.a_x = cH * dy;
  f.a_y = -cH * dx;
  f.a_e = -3.0 * (1.0 + k.eps_prime * qm) + (3.0 / k.a_bar) * qm;
  if (kAngular) {
    const double Him = H * im;
    f.da = -k.a12 * Him;
    const double H_x = dy * (3.0 * dx2 - dy2), H_y = dx * (dx2 - 3.0 * dy2);
    const double K = -k.a12 * is3;
    f.da_x = K * (H_x * s - 4.0 * dx * H);
    f.da_y = K * (H_y * s - 4.0 * dy * H);
    f.da_e = -12.0 * Him;
  } else {
    f.da = f.da_x = f.da_y = f.da_e = 0.0;
  }
  return f;
}

struct RevArgs {
  lgd_kr_consts k;
  double dt;
  const double *phi, *u;    // state (slabs)
  const double *APi, *AUi;  // incoming adjoint (global arrays, N0 × N1)
  double *APo, *AUo;        // outgoing adjoint
  double *acc;              // 4 θ partials per wave
  int N0, N1, n0, n1, nbx, sx;
  int64_t slab;
  int per0, per1;
  int tile, ntx, rows, nstrips, nwaves;
};

constexpr int kWaves = 4;  // waves per workgroup

// Row-iteration phases of the reverse kernel.
constexpr int kGeneral = 0, kHead = 1, kFirst = 2, kMid = 3, kLast = 4;

__global__ __launch_bounds__(64 * kWaves) void adj_stream_kernel(RevArgs a) {
#pragma clang fp contract(fast)
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int lane = threadIdx.x & 63;
  const int w = blockIdx.x * kWaves + wave;
  if (w >= a.nwaves) return;  // whole wave
  const int ty = w / a.ntx, tx = w - ty * a.ntx;
  const int x0 = tx * a.tile, x1 = min(x0 + a.tile, a.N0);
  const int y0 = ty * a.rows, y1 = min(y0 + a.rows, a.N1);
  const int N0 = a.N0, N1 = a.N1;
  const lgd_kr_consts &k = a.k;
  const double q = 0.25;

  // This lane's global column; lanes past x1 + 1 repeat column x1 + 1.
  const int gx = min(x0 - 2 + lane, x1 + 1);
  const int rx = resolve_i(gx, N0, a.per0);
  const bool colreal = a.per0 || (gx >= 0 && gx < N0);
  const bool colown = gx >= x0 && gx < x1 && lane >= 2;
  const bool left_edge = !a.per0 && gx == 0;        // folds the ghost column −1
  const bool right_edge = !a.per0 && gx == N0 - 1;  // folds the ghost column N0
  const bool face_left_ghost = !a.per0 && gx == -1 && x0 == 0;  // owns face (−1|0)
  int64_t colofs;
  {
    const int bxi = rx / a.n0, i = rx - bxi * a.n0;
    colofs = (int64_t)bxi * a.slab + (i + 1);
  }
  // Slab offset of the start of global row g's resolved image (block column 0), and
  // the resolved row's position within its block.
  const int64_t bstride = (int64_t)a.nbx * a.slab;
  auto slab_row = [&](int g, int *jout) {
    const unsigned ry = (unsigned)resolve_i(g, N1, a.per1);
    const unsigned byi = ry / (unsigned)a.n1, j = ry - byi * (unsigned)a.n1;
    *jout = (int)j;
    return (int64_t)byi * bstride + (int64_t)(j + 1) * a.sx;
  };
  auto load = [&](const double *f, int g) {
    int j;
    return f[slab_row(g, &j) + colofs];
  };

  // Rows y0−2 .. y0 of φ and u (m: r−1, c: r, p: r+1 for r = y0−1).
  double Pm = load(a.phi, y0 - 2), Pc = load(a.phi, y0 - 1), Pp = load(a.phi, y0);
  double Um = load(a.u, y0 - 2), Uc = load(a.u, y0 - 1), Up = load(a.u, y0);
  double PmR = from_right(Pm), PcL = from_left(Pc), PcR = from_right(Pc);
  double UcL = from_left(Uc), UcR = from_right(Uc);
  // Flux through the y-face between rows y0−2 and y0−1 (J_B of row y0−1).
  double JB;
  {
    const double PmL = from_left(Pm);
    const double derx = q * (((PcR - PcL) + PmR) - PmL);
    const double dery = Pc - Pm;
    double fa, fda;
    lgd_kr_face(k, derx, dery, fa, fda);
    JB = fa * ((fa * dery) + (fda * derx));
  }
  // Incoming adjoint at row r (loaded one row ahead).
  auto adj_row = [&](int g) { return (int64_t)N0 * resolve_i(g, N1, a.per1); };
  double ap = a.APi[adj_row(y0 - 1) + rx], au = a.AUi[adj_row(y0 - 1) + rx];
  // Row offsets carried across iterations: the state row gS loaded next (offS, its
  // position jS within the block) and the adjoint row gA loaded next (offA).  Rows
  // inside the domain advance by one row (and one block row at a block edge); the
  // two rows each side of the domain take the general form.
  int gS = y0 + 1, jS;
  int64_t offS = slab_row(gS, &jS);
  int gA = y0;
  int64_t offA = adj_row(gA);
  double Rprev = 0.0, FYx = 0.0, FYy = 0.0, FYe = 0.0;  // y-face (r−1|r)
  double accm = 0.0, accc = 0.0, accp = 0.0;
  double Vmm = 0.0, Vm = 0.0, PUm = 0.0;
  double part_l = 0.0, part_d = 0.0, part_ec = 0.0, part_ef = 0.0;

  // One row iteration.  kGeneral evaluates every ownership and mirror condition at
  // run time.  The other phases serve interior strips, whose rows y0−2 .. y1+1 all
  // lie inside the domain (no ghost row resolves and no mirror fold applies): the
  // halo row y0−1 (kHead), the first owned row (kFirst, row y0−1 is not written),
  // the owned rows after it (kMid) and the halo row y1 (kLast, which loads nothing).
  const int64_t bjump = bstride - (int64_t)a.n1 * a.sx;  // slab row n1 → next block row
  // kXEdge: the tile reaches within two columns of a mirror edge in x, so some lane
  // is a ghost column (colreal false) or folds one (left_edge, right_edge,
  // face_left_ghost); without it those flags are false on every lane.
  auto row_step = [&](int r, auto phase, auto xedge) {
    constexpr int kPhase = decltype(phase)::value;
    constexpr bool kGen = kPhase == kGeneral;
    constexpr bool kXEdge = decltype(xedge)::value;
    // Row r+2 of the state and the adjoint at row r+1 (in the general form, rows
    // y1+1 and y1 again at the last row, where they go unused), then the offsets of
    // the next ones.
    double Pn = 0.0, Un = 0.0, apn = 0.0, aun = 0.0;
    if (kPhase != kLast) {
      Pn = a.phi[offS + colofs];
      Un = a.u[offS + colofs];
      apn = a.APi[offA + rx];
      aun = a.AUi[offA + rx];
    }
    if (!kGen) {
      offS += a.sx;
      if (++jS == a.n1) {
        jS = 0;
        offS += bjump;
      }
      offA += N0;
    } else {
      const int g = min(r + 3, y1 + 1);
      if (g != gS) {
        if (g >= 1 && g < N1) {
          offS += a.sx;
          if (++jS == a.n1) {
            jS = 0;
            offS += bjump;
          }
        } else {
          offS = slab_row(g, &jS);
        }
        gS = g;
      }
      const int ga = min(r + 2, y1);
      if (ga != gA) {
        offA = (ga >= 1 && ga < N1) ? offA + N0 : adj_row(ga);
        gA = ga;
      }
    }
    const bool row_real = !kGen || a.per1 || (r >= 0 && r < N1);
    const bool real = kXEdge ? colreal && row_real : row_real;
    const bool row_own = kGen ? r >= y0 && r < y1 : kPhase == kFirst || kPhase == kMid;
    // Row r−1 is owned (complete after this row, and its y-face (r−1|r) is owned).
    const bool prev_own = kGen ? r - 1 >= y0 && r - 1 < y1 : kPhase == kMid || kPhase == kLast;
    // The mirror folds (never in an interior strip).
    const bool fold_lo = kGen && !a.per1 && r == 0;   // ghost row −1 onto row 0
    const bool fold_hi = kGen && !a.per1 && r == N1;  // ghost row N1 onto row N1−1
    const double PpL = from_left(Pp), PpR = from_right(Pp);
    const double UpL = from_left(Up), UpR = from_right(Up);

    // x-face right of this column at row r.
    double JX, FXx, FXy, FXe;
    {
      const double derx = PcR - Pc;
      const double dery = q * (((PpR - PmR) + Pp) - Pm);
      const FD h = fderiv<true>(k, k.fv_tol, derx, dery);
      // With P = inner + a·derx: ∂J/∂derx = a_x·inner + a·(a_x·derx + a − a′_x·dery)
      // = a_x·P + a·(a − a′_x·dery), and likewise for dery and ε₄.
      const double inner = h.a * derx - h.da * dery;
      const double P = inner + h.a * derx;
      JX = h.a * ((h.a * derx) - (h.da * dery));
      FXx = h.a_x * P + h.a * (h.a - h.da_x * dery);
      FXy = h.a_y * P - h.a * (h.da_y * dery + h.da);
      FXe = h.a_e * P - h.a * (h.da_e * dery);
    }
    // y-face between rows r and r+1.
    double JT, NYx, NYy, NYe;
    {
      const double derx = q * (((PpR - PpL) + PcR) - PcL);
      const double dery = Pp - Pc;
      const FD g = fderiv<true>(k, k.fv_tol, derx, dery);
      // With Q = inner + a·dery, as for the x-face.
      const double inner = g.a * dery + g.da * derx;
      const double Q = inner + g.a * dery;
      JT = g.a * ((g.a * dery) + (g.da * derx));
      NYx = g.a_x * Q + g.a * (g.da_x * derx + g.da);
      NYy = g.a_y * Q + g.a * (g.a + g.da_y * derx);
      NYe = g.a_e * Q + g.a * (g.da_e * derx);
    }
    // Cell quantities at row r.
    const double JL = from_left(JX);
    const double flux = (((JX - JL) + JT) - JB) * k.inv_h2;
    const double g_prime = fma(Pc * Pc, Pc, -Pc);
    const double one_m_p2 = fma(Pc, -Pc, 1.0);
    const double p_prime = one_m_p2 * one_m_p2;
    const FD ca = fderiv<false>(k, k.model_tol, 0.5 * (PcR - PcL), 0.5 * (Pp - Pm));
    const double iA = recip(ca.a);
    const double iA2 = iA * iA;
    const double F = fma(k.lambda * Uc, -p_prime, flux - g_prime) * iA2;
    const double W = real ? a.dt * (ap + 0.5 * au) : 0.0;
    const double R = W * iA2;
    const double S = -2.0 * (W * F) * iA;
    const double SX = S * ca.a_x * 0.5;
    const double SY = S * ca.a_y * 0.5;
    const double PW =
        real ? ap + R * (-(3.0 * Pc * Pc - 1.0) - k.lambda * Uc * (2.0 * one_m_p2 * (-2.0 * Pc)))
             : 0.0;
    const double PU = au + R * (-k.lambda * p_prime);
    const double V = real ? a.dt * k.d_thermal * au : 0.0;
    if (row_own && colown) {
      double lap = 0.0;
      lap = lap + (((UcR - (Uc + Uc)) + UcL) * k.inv_hx2);
      lap = lap + (((Up - (Uc + Uc)) + Um) * k.inv_hy2);
      part_l += R * (-Uc * p_prime);
      part_d += a.dt * au * lap;
      part_ec += S * ca.a_e;
    }
    // x-face weights (R(c) − R(c+1))/(hx·hy); y-face (r−1|r) weights (R(r−1) − R(r))/(hx·hy).
    const double G = (R - from_right(R)) * k.inv_h2;
    const double gxs = G * FXx, gys = q * (G * FXy);
    if (row_own && (colown || (kXEdge && face_left_ghost))) part_ef += G * FXe;
    const double Hy = (Rprev - R) * k.inv_h2;
    const double hys = Hy * FYy, hxs = q * (Hy * FYx);
    if (colown && (prev_own || (fold_lo && r == y0))) part_ef += Hy * FYe;
    // x-faces: −gx to the left cell, +gx to the right cell, ±q·gy to rows r±1;
    // y-face (r−1|r): +gy to row r, −gy to row r−1, ±q·gx to columns c±1;
    // centres: ±SX to columns c±1, ±SY to rows r±1.
    const double gx_in = from_left(gxs) - gxs;
    const double gy_in = from_left(gys) + gys;
    const double hx_in = from_left(hxs) - from_right(hxs);
    const double sx_in = from_left(SX) - from_right(SX);
    accc += ((PW + gx_in) + (hys + hx_in)) + sx_in;
    accp += gy_in + SY;
    accm += (hx_in - hys) - (gy_in + SY);
    // Row r−1 is complete.
    if (fold_lo) {
      accc += accm;  // the mirror ghost row −1 is an image of row 0
    } else if (prev_own) {
      if (fold_hi) accm += accc;  // the mirror ghost row N1 is an image of row N1−1
      const double vl = from_left(Vm), vr = from_right(Vm);
      double lv = 0.0, rv = 0.0;
      if (kXEdge) {
        lv = from_left(accm);
        rv = from_right(accm);
      }
      // Outside the general form every lane stores, the unowned ones into the pad past
      // the N0 × N1 outputs: no branch around the stores, so the wait for a row's
      // loads never has to count in a store that a skipped branch might not have
      // issued, and so never waits for this row's stores.
      if (!kGen || colown) {
        double out = accm;
        double lx = vl + vr;
        if (kXEdge && left_edge) {
          out += lv;
          lx += Vm;
        }
        if (kXEdge && right_edge) {
          out += rv;
          lx += Vm;
        }
        double ly = Vmm + V;
        if (kGen && !a.per1 && r - 1 == 0) ly += Vm;
        if (kGen && !a.per1 && r - 1 == N1 - 1) ly += Vm;
        const double lap = (lx - 2.0 * Vm) * k.inv_hx2 + (ly - 2.0 * Vm) * k.inv_hy2;
        const int64_t o = kGen || colown ? (int64_t)gx + (int64_t)N0 * (r - 1)
                                         : (int64_t)N0 * N1 + lane;
        a.APo[o] = out;
        a.AUo[o] = PUm + lap;
      }
    }
    // Advance one row.
    accm = accc;
    accc = accp;
    accp = 0.0;
    Vmm = Vm;
    Vm = V;
    PUm = PU;
    Rprev = R;
    FYx = NYx;
    FYy = NYy;
    FYe = NYe;
    JB = JT;
    Pm = Pc;
    Pc = Pp;
    Pp = Pn;
    PmR = PcR;
    PcL = PpL;
    PcR = PpR;
    Um = Uc;
    Uc = Up;
    Up = Un;
    UcL = UpL;
    UcR = UpR;
    ap = apn;
    au = aun;
  };
  auto run = [&](auto xedge) {
    if (y0 >= 2 && y1 + 1 <= N1 - 1) {
      row_step(y0 - 1, std::integral_constant<int, kHead>(), xedge);
      row_step(y0, std::integral_constant<int, kFirst>(), xedge);
      for (int r = y0 + 1; r < y1; ++r) row_step(r, std::integral_constant<int, kMid>(), xedge);
      row_step(y1, std::integral_constant<int, kLast>(), xedge);
    } else {
      for (int r = y0 - 1; r <= y1; ++r) row_step(r, std::integral_constant<int, kGeneral>(), xedge);
    }
  };
  if (!a.per0 && (x0 <= 2 || x1 + 2 >= N0))
    run(std::true_type());
  else
    run(std::false_type());
  // Fixed-order wave sums of the θ partials.
  double p4[4] = {part_l, part_d, part_ec, part_ef};
  for (int n = 0; n < 4; ++n)
    for (int off = 32; off > 0; off >>= 1) p4[n] += __shfl_xor(p4[n], off, 64);
  if (lane == 0)
    for (int n = 0; n < 4; ++n) a.acc[4 * w + n] += p4[n];
}

// Column tile width: at most 60 owned columns, split evenly.
int tile_width(int64_t N0) {
  const int64_t ntx = (N0 + 59) / 60;
  return (int)((N0 + ntx - 1) / ntx);
}

// Strip height for `ntx` column tiles over N1 rows on a device of `simds` SIMDs:
// the least time when every wave runs rows + 2 row iterations and the busiest SIMD
// holds ⌈waves / simds⌉ of them, among the heights that give at least 1.5 waves per
// SIMD where any does (a lone wave per SIMD leaves its load and DPP latency
// exposed: measured slower at 1024²); ties go to more waves.
int default_rows(int64_t ntx, int64_t N1, int64_t simds) {
  int best = 2;
  double best_cost = 1e300;
  const bool can_fill = ntx * ((N1 + 1) / 2) * 2 >= 3 * simds;
  for (int rows = 2; rows <= 64 && rows <= N1; ++rows) {
    const int64_t waves = ntx * ((N1 + rows - 1) / rows);
    if (can_fill && 2 * waves < 3 * simds) continue;
    const double per_simd = (double)((waves + simds - 1) / simds);
    const double cost = per_simd * (rows + 2);
    if (cost < best_cost) {
      best_cost = cost;
      best = rows;
    }
  }
  return best;
}

}  // namespace

int lgd_adjoint_stream_launch(lgd_ctx *ctx, const lgd_cart &c, const lgd_kr_consts &k, double dt,
                              const double *phi, const double *u, const double *APi,
                              const double *AUi, double *APo, double *AUo, double *acc,
                              int rows, int64_t *nwaves_out) {
  if (c.dim != 2) return LGD_ERR_UNSUPPORTED;
  const int64_t N0 = c.n[0] * c.nb[0], N1 = c.n[1] * c.nb[1];
  // Ghost images two cells out must resolve with one reflection or wrap.
  if (N0 < 2 || N1 < 2) return LGD_ERR_UNSUPPORTED;
  const int tile = tile_width(N0);
  const int64_t ntx = (N0 + tile - 1) / tile;
  if (rows == 0) rows = default_rows(ntx, N1, 4 * (int64_t)(ctx->num_cus > 0 ? ctx->num_cus : 256));
  if (rows < 2) rows = 2;
  if (rows > N1) rows = (int)N1;
  RevArgs a;
  a.k = k;
  a.dt = dt;
  a.phi = phi;
  a.u = u;
  a.APi = APi;
  a.AUi = AUi;
  a.APo = APo;
  a.AUo = AUo;
  a.acc = acc;
  a.N0 = (int)N0;
  a.N1 = (int)N1;
  a.n0 = (int)c.n[0];
  a.n1 = (int)c.n[1];
  a.nbx = (int)c.nb[0];
  a.sx = (int)(c.n[0] + 2);
  a.slab = (int64_t)(c.n[0] + 2) * (c.n[1] + 2);
  a.per0 = c.periodic[0];
  a.per1 = c.periodic[1];
  a.tile = tile;
  a.ntx = (int)ntx;
  a.rows = rows;
  a.nstrips = (int)((N1 + rows - 1) / rows);
  const int64_t nwaves = (int64_t)a.ntx * a.nstrips;
  a.nwaves = (int)nwaves;
  if (nwaves_out) *nwaves_out = nwaves;
  if (!acc) return LGD_OK;  // size query
  const unsigned wgs = (unsigned)((nwaves + kWaves - 1) / kWaves);
  adj_stream_kernel<<<wgs, 64 * kWaves, 0, ctx->stream>>>(a);
  return lgd::launch_check(ctx, "adjoint stream");
}
