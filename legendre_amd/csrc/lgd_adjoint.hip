// Reverse-mode derivative of the explicit Model C step and the checkpointed
// calibration gradient.
//
// The forward step being differentiated is the fused Model C step with b = 0
// (Euler–Maruyama ≡ forward Euler, src/integrators/euler_maruyama.rs:34-75):
//   φ' = φ + dt·F,   u' = u + dt·(D·∇²u + ½F),
//   F = fma(λu, −(1−φ²)², ∇·J(φ) − (φ³−φ)) / A(∇φ)²
// (src/physics/phasefield.rs:394-421, src/discretization/finite_volume.rs:248-284).
//
// A reverse step runs three deterministic passes over global (unblocked) scratch
// arrays; state and adjoint fields stay in the slab layout of the forward:
//   1. cells:  recompute F and A at each interior cell and form the per-cell
//              sensitivities R = ∂L/∂(numerator of F), S = ∂L/∂A (split into the
//              central-difference directions), the pointwise terms, the Laplacian
//              weight V = dt·D·∂L/∂u', and the cell's share of ∂L/∂λ, ∂L/∂D, ∂L/∂ε₄;
//   2. faces:  every x- and y-face any interior cell reads, weighted by
//              (R_left − R_right)/(hx·hy), times the derivative of the face flux
//              J = A(A·d_n ∓ A′·d_t) w.r.t. its normal and transverse differences;
//   3. gather: each interior cell sums, in a fixed order, every contribution whose
//              stencil reads it — at its own position and at each ghost position
//              that resolves to it (mirror reflection or periodic wrap), so the
//              transpose of the ghost fill needs no scatter and no atomics.
// Where the block shape allows, one streaming kernel (lgd_adjoint_stream.hip)
// replaces the three passes: it reads the state slabs and the incoming global
// adjoint and writes the outgoing one into a second pair of global arrays.
// θ partials accumulate per cell / per face (per wave in the streaming kernel)
// across steps and are reduced once, in a fixed order, so the gradient is bitwise
// reproducible.
//
// Checkpointing: binomial (Griewank–Walther) reversal with a fixed number of state
// slots; the schedule's first chain of checkpoints is laid down by the forward sweep.
#include <algorithm>
#include <cmath>
#include <mutex>
#include <vector>

#include "lgd_internal.h"

namespace {

struct AdjGrid {
  int N0, N1;    // global interior cells
  int n0, n1;    // block interior cells
  int nb0;       // blocks per row
  int sx;        // slab row length (ghost 1)
  int64_t slab;  // slab length
  int per0, per1;
  int P0;        // padded row length N0 + 2
};

__device__ __forceinline__ int resolve(int i, int n, int per) {
  if (i < 0) return per ? i + n : -1 - i;
  if (i >= n) return per ? i - n : 2 * n - 1 - i;
  return i;
}

// Slab offset of global interior cell (gi, gj) (both in range).
__device__ __forceinline__ int64_t slab_off(const AdjGrid &G, int gi, int gj) {
  const int bx = gi / G.n0, by = gj / G.n1;
  const int i = gi - bx * G.n0, j = gj - by * G.n1;
  return (int64_t)(bx + G.nb0 * by) * G.slab + (int64_t)((j + 1) * G.sx + (i + 1));
}

// face(dx, dy) (finite_volume.rs:129-142) with its partial derivatives w.r.t. dx, dy
// and ε₄ (through ā = 1 − 3ε₄, ε′ = 3ε₄/ā, a₁₂ = 4āε′: dā/dε₄ = −3,
// dε′/dε₄ = 3/ā², da₁₂/dε₄ = 12).  The same function with the model tolerance is the
// centre anisotropy (phasefield.rs:159-175).
struct FaceD {
  double a, da, a_x, a_y, a_e, da_x, da_y, da_e;
};

__device__ __forceinline__ FaceD face_deriv(const lgd_kr_consts &k, double tol, double dx,
                                            double dy) {
  FaceD f;
  double dx2 = dx * dx, dy2 = dy * dy;
  double s = dx2 + dy2;
  double m = s * s;
  if (m <= tol) {
    f.a = k.a_bar;
    f.da = 0.0;
    f.a_x = f.a_y = f.da_x = f.da_y = f.da_e = 0.0;
    f.a_e = -3.0;
    return f;
  }
  double Q = dx2 * dx2 + dy2 * dy2;
  double H = dx * dy * (dx2 - dy2);
  f.a = k.a_bar * (1.0 + k.eps_prime * Q / m);
  f.da = -(k.a12) * dx * dy * (dx2 - dy2) / m;
  double s3 = s * s * s;
  double c4 = 4.0 * k.a_bar * k.eps_prime;
  f.a_x = c4 * dx * dy2 * (dx2 - dy2) / s3;
  f.a_y = c4 * dy * dx2 * (dy2 - dx2) / s3;
  double H_x = dy * (3.0 * dx2 - dy2), H_y = dx * (dx2 - 3.0 * dy2);
  f.da_x = -k.a12 * (H_x * s - 4.0 * dx * H) / s3;
  f.da_y = -k.a12 * (H_y * s - 4.0 * dy * H) / s3;
  f.a_e = -3.0 * (1.0 + k.eps_prime * Q / m) + k.a_bar * (3.0 / (k.a_bar * k.a_bar)) * Q / m;
  f.da_e = -12.0 * H / m;
  return f;
}

constexpr int kWG = 256;

// Fixed-order workgroup sum of kN values per thread; thread 0 adds the totals to
// this workgroup's accumulator slots (acc[kN*blockIdx.x + n]), so across launches
// each slot sums in step order: bitwise reproducible.
template <int kN>
__device__ __forceinline__ void wg_accumulate(const double (&v)[kN], double *acc) {
  __shared__ double part[kN][kWG];
  for (int n = 0; n < kN; ++n) part[n][threadIdx.x] = v[n];
  __syncthreads();
  for (int w = kWG / 2; w > 0; w >>= 1) {
    if ((int)threadIdx.x < w)
      for (int n = 0; n < kN; ++n) part[n][threadIdx.x] += part[n][threadIdx.x + w];
    __syncthreads();
  }
  if (threadIdx.x == 0)
    for (int n = 0; n < kN; ++n) acc[kN * blockIdx.x + n] += part[n][0];
}

// Stage: φ and u on the padded global grid ((N0+2) × (N1+2); the ring holds the
// resolved neighbour values, i.e. the ghost fill), and (kAdj) the incoming adjoint
// from slabs onto the unpadded global grid.
template <bool kAdj>
__global__ __launch_bounds__(kWG) void adj_stage_kernel(AdjGrid G, const double *phi,
                                                        const double *u, const double *pphi,
                                                        const double *pu, double *Pp, double *Up,
                                                        double *AP, double *AU) {
  const int t = blockIdx.x * kWG + threadIdx.x;
  if (t >= G.P0 * (G.N1 + 2)) return;
  const int x = t % G.P0 - 1, y = t / G.P0 - 1;
  const int rx = resolve(x, G.N0, G.per0), ry = resolve(y, G.N1, G.per1);
  const int64_t o = slab_off(G, rx, ry);
  Pp[t] = phi[o];
  Up[t] = u[o];
  if (kAdj && x == rx && y == ry) {
    AP[x + G.N0 * y] = pphi[o];
    AU[x + G.N0 * y] = pu[o];
  }
}

// Pass 1: per interior cell.  acc gets (∂L/∂λ, ∂L/∂D, ∂L/∂ε₄ centre part) per workgroup.
__global__ __launch_bounds__(kWG) void adj_cell_kernel(AdjGrid G, lgd_kr_consts k, double dt,
                                                       const double *Pp, const double *Up,
                                                       const double *AP, const double *AU,
                                                       double *R, double *SX, double *SY,
                                                       double *PW, double *PU, double *V,
                                                       double *acc) {
  const int t = blockIdx.x * kWG + threadIdx.x;
  double part[3] = {0.0, 0.0, 0.0};
  if (t < G.N0 * G.N1) {
    const int i = t % G.N0, j = t / G.N0;
    const int c = (i + 1) + G.P0 * (j + 1);
    double p[3][3];
    for (int dj = -1; dj <= 1; ++dj)
      for (int di = -1; di <= 1; ++di) p[dj + 1][di + 1] = Pp[c + di + G.P0 * dj];
    const double uc = Up[c], ue = Up[c + 1], uw = Up[c - 1];
    const double un = Up[c + G.P0], us = Up[c - G.P0];
    // The forward quantities, in the forward's operation order.
    double F, du_unused;
    lgd_modelc_cell(k, p, uc, ue, uw, un, us, F, du_unused);
    (void)du_unused;
    const double pv = p[1][1];
    const double one_m_p2 = fma(pv, -pv, 1.0);
    const double p_prime = one_m_p2 * one_m_p2;
    const FaceD ca =
        face_deriv(k, k.model_tol, 0.5 * (p[1][2] - p[1][0]), 0.5 * (p[2][1] - p[0][1]));
    const double A = ca.a;
    double lap = 0.0;
    lap = lap + (((ue - (uc + uc)) + uw) * k.inv_hx2);
    lap = lap + (((un - (uc + uc)) + us) * k.inv_hy2);
    // Reverse: ∂L/∂F = dt·(P_φ' + ½P_u'), ∂L/∂lap = dt·D·P_u'.
    const double ap = AP[t], au = AU[t];
    const double W = dt * (ap + 0.5 * au);
    const double Rr = W / (A * A);
    const double S = -2.0 * W * F / A;
    R[t] = Rr;
    SX[t] = S * ca.a_x * 0.5;
    SY[t] = S * ca.a_y * 0.5;
    PW[t] = ap + Rr * (-(3.0 * pv * pv - 1.0) - k.lambda * uc * (2.0 * one_m_p2 * (-2.0 * pv)));
    PU[t] = au + Rr * (-k.lambda * p_prime);
    V[t] = dt * k.d_thermal * au;
    part[0] = Rr * (-uc * p_prime);
    part[1] = dt * au * lap;
    part[2] = S * ca.a_e;
  }
  wg_accumulate<3>(part, acc);
}

// Pass 2: faces.  Thread (xp, yp) over [0, N0] × [0, N1] owns the x-face between
// global columns xp−1 and xp at row yp (yp < N1) and the y-face between rows yp−1
// and yp at column xp (xp < N0).  acc gets the faces' ∂L/∂ε₄ per workgroup.
__global__ __launch_bounds__(kWG) void adj_face_kernel(AdjGrid G, lgd_kr_consts k,
                                                       const double *Pp, const double *R,
                                                       double *GXx, double *GYx, double *GXy,
                                                       double *GYy, double *acc) {
  const int t = blockIdx.x * kWG + threadIdx.x;
  const int W0 = G.N0 + 1;
  double part[1] = {0.0};
  if (t < W0 * (G.N1 + 1)) {
    const int xp = t % W0, yp = t / W0;
    const double q = 0.25;
    auto Rin = [&](int i, int j) {
      return (i >= 0 && i < G.N0 && j >= 0 && j < G.N1) ? R[i + G.N0 * j] : 0.0;
    };
    // padded value at extended position (x, y) ∈ [−1, N0] × [−1, N1]
    auto P = [&](int x, int y) { return Pp[(x + 1) + G.P0 * (y + 1)]; };
    if (yp < G.N1) {
      const int i = xp - 1, j = yp;
      const double Gf = (Rin(i, j) - Rin(i + 1, j)) * k.inv_h2;
      double gx = 0.0, gy = 0.0;
      if (Gf != 0.0) {
        const double derx = P(i + 1, j) - P(i, j);
        const double dery = q * (((P(i + 1, j + 1) - P(i + 1, j - 1)) + P(i, j + 1)) - P(i, j - 1));
        const FaceD h = face_deriv(k, k.fv_tol, derx, dery);
        const double inner = h.a * derx - h.da * dery;  // J = a·inner
        const double j_x = h.a_x * inner + h.a * (h.a_x * derx + h.a - h.da_x * dery);
        const double j_y = h.a_y * inner + h.a * (h.a_y * derx - h.da_y * dery - h.da);
        const double j_e = h.a_e * inner + h.a * (h.a_e * derx - h.da_e * dery);
        gx = Gf * j_x;
        gy = Gf * j_y;
        part[0] += Gf * j_e;
      }
      const int fx = xp + W0 * yp;
      GXx[fx] = gx;
      GYx[fx] = gy;
    }
    if (xp < G.N0) {
      const int i = xp, j = yp - 1;
      const double Gf = (Rin(i, j) - Rin(i, j + 1)) * k.inv_h2;
      double gx = 0.0, gy = 0.0;
      if (Gf != 0.0) {
        const double derx = q * (((P(i + 1, j + 1) - P(i - 1, j + 1)) + P(i + 1, j)) - P(i - 1, j));
        const double dery = P(i, j + 1) - P(i, j);
        const FaceD h = face_deriv(k, k.fv_tol, derx, dery);
        const double inner = h.a * dery + h.da * derx;  // J = a·inner
        const double j_x = h.a_x * inner + h.a * (h.a_x * dery + h.da_x * derx + h.da);
        const double j_y = h.a_y * inner + h.a * (h.a_y * dery + h.a + h.da_y * derx);
        const double j_e = h.a_e * inner + h.a * (h.a_e * dery + h.da_e * derx);
        gx = Gf * j_x;
        gy = Gf * j_y;
        part[0] += Gf * j_e;
      }
      const int fy = xp + G.N0 * yp;
      GXy[fy] = gx;
      GYy[fy] = gy;
    }
  }
  wg_accumulate<1>(part, acc);
}

// Pass 3: gather at each interior cell; the outgoing adjoint goes to slabs (kSlab)
// or, in place, to the global arrays the next reverse step reads.
template <bool kSlab>
__global__ __launch_bounds__(kWG) void adj_gather_kernel(AdjGrid G, lgd_kr_consts k,
                                                         const double *GXx, const double *GYx,
                                                         const double *GXy, const double *GYy,
                                                         const double *SX, const double *SY,
                                                         const double *PW, const double *PU,
                                                         const double *V, double *pphi,
                                                         double *pu) {
  const int t = blockIdx.x * kWG + threadIdx.x;
  if (t >= G.N0 * G.N1) return;
  const int a = t % G.N0, b = t / G.N0;
  const int N0 = G.N0, N1 = G.N1, W0 = N0 + 1;
  const double q = 0.25;
  // x-face (x|x+1) at row y, x ∈ [−1, N0−1]; y-face (y|y+1) at column x, y ∈ [−1, N1−1].
  auto gxx = [&](int x, int y) {
    return (x >= -1 && x < N0 && y >= 0 && y < N1) ? GXx[(x + 1) + W0 * y] : 0.0;
  };
  auto gyx = [&](int x, int y) {
    return (x >= -1 && x < N0 && y >= 0 && y < N1) ? GYx[(x + 1) + W0 * y] : 0.0;
  };
  auto gxy = [&](int x, int y) {
    return (x >= 0 && x < N0 && y >= -1 && y < N1) ? GXy[x + N0 * (y + 1)] : 0.0;
  };
  auto gyy = [&](int x, int y) {
    return (x >= 0 && x < N0 && y >= -1 && y < N1) ? GYy[x + N0 * (y + 1)] : 0.0;
  };
  auto inner = [&](const double *f, int x, int y) {
    return (x >= 0 && x < N0 && y >= 0 && y < N1) ? f[x + N0 * y] : 0.0;
  };
  // Everything that reads the value at extended position (x, y).
  auto contrib = [&](int x, int y) {
    double xf = (gxx(x - 1, y) - gxx(x, y)) +
                q * ((gyx(x, y - 1) + gyx(x - 1, y - 1)) - (gyx(x, y + 1) + gyx(x - 1, y + 1)));
    double yf = (gyy(x, y - 1) - gyy(x, y)) +
                q * ((gxy(x - 1, y) + gxy(x - 1, y - 1)) - (gxy(x + 1, y) + gxy(x + 1, y - 1)));
    double cf = (inner(SX, x - 1, y) - inner(SX, x + 1, y)) +
                (inner(SY, x, y - 1) - inner(SY, x, y + 1));
    return (xf + yf) + cf;
  };
  // Positions along each dimension that resolve to this cell.
  int xs[3], ys[3];
  int nx = 0, ny = 0;
  xs[nx++] = a;
  if (a == 0) xs[nx++] = G.per0 ? N0 : -1;
  if (a == N0 - 1) xs[nx++] = G.per0 ? -1 : N0;
  ys[ny++] = b;
  if (b == 0) ys[ny++] = G.per1 ? N1 : -1;
  if (b == N1 - 1) ys[ny++] = G.per1 ? -1 : N1;
  double acc = PW[t];
  for (int iy = 0; iy < ny; ++iy)
    for (int ix = 0; ix < nx; ++ix) acc += contrib(xs[ix], ys[iy]);
  // Transpose of the u Laplacian: Σ_d ih_d²·(Σ_images V(p∓e_d) − 2V).
  double lx = 0.0, ly = 0.0;
  for (int ix = 0; ix < nx; ++ix) lx += inner(V, xs[ix] - 1, b) + inner(V, xs[ix] + 1, b);
  for (int iy = 0; iy < ny; ++iy) ly += inner(V, a, ys[iy] - 1) + inner(V, a, ys[iy] + 1);
  const double v = V[t];
  const double lap = (lx - 2.0 * v) * k.inv_hx2 + (ly - 2.0 * v) * k.inv_hy2;
  if (kSlab) {
    const int64_t o = slab_off(G, a, b);
    pphi[o] = acc;
    pu[o] = PU[t] + lap;
  } else {
    pphi[t] = acc;
    pu[t] = PU[t] + lap;
  }
}

// Slab interiors → global adjoint arrays.
__global__ __launch_bounds__(kWG) void adj_flatten_kernel(AdjGrid G, const double *pphi,
                                                          const double *pu, double *AP,
                                                          double *AU) {
  const int t = blockIdx.x * kWG + threadIdx.x;
  if (t >= G.N0 * G.N1) return;
  const int64_t o = slab_off(G, t % G.N0, t / G.N0);
  AP[t] = pphi[o];
  AU[t] = pu[o];
}

// Global adjoint arrays → slab interiors.
__global__ __launch_bounds__(kWG) void adj_unflatten_kernel(AdjGrid G, const double *AP,
                                                            const double *AU, double *pphi,
                                                            double *pu) {
  const int t = blockIdx.x * kWG + threadIdx.x;
  if (t >= G.N0 * G.N1) return;
  const int64_t o = slab_off(G, t % G.N0, t / G.N0);
  if (pphi) pphi[o] = AP[t];
  if (pu) pu[o] = AU[t];
}

__global__ void adj_zero_kernel(int64_t n, double *x) {
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t < n) x[t] = 0.0;
}

// Terminal adjoint of L = ½‖φ − target‖²: P_φ = φ − target, P_u = 0, r² per cell.
__global__ void adj_loss_kernel(AdjGrid G, const double *phi, const double *target, double *AP,
                                double *AU, double *r2) {
  const int t = blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= G.N0 * G.N1) return;
  const int64_t o = slab_off(G, t % G.N0, t / G.N0);
  const double r = phi[o] - target[o];
  AP[t] = r;
  AU[t] = 0.0;
  r2[t] = r * r;
}

// Fixed-order sum of n values into out[0]: one workgroup, strided partial sums,
// then a fixed tree.
// Fixed-order sums of n records of `stride` values (record i at x[stride*i]):
// out[c] = Σ_i x[stride*i + c] for c < stride.
constexpr int kSumThreads = 1024;
__global__ __launch_bounds__(kSumThreads) void adj_sum_kernel(int64_t n, const double *x,
                                                              int stride, double *out) {
  __shared__ double part[kSumThreads];
  for (int c = 0; c < stride; ++c) {
    double s = 0.0;
    for (int64_t i = threadIdx.x; i < n; i += kSumThreads) s += x[stride * i + c];
    part[threadIdx.x] = s;
    __syncthreads();
    for (int w = kSumThreads / 2; w > 0; w >>= 1) {
      if ((int)threadIdx.x < w) part[threadIdx.x] += part[threadIdx.x + w];
      __syncthreads();
    }
    if (threadIdx.x == 0) out[c] = part[0];
    __syncthreads();
  }
}

// Binomial checkpointing cost: forward steps to reverse n steps from a stored
// state with s further free slots (T(n,0) = n(n−1)/2; otherwise r·n − β(s+2, r−1)
// for the least r with β(s+1, r) ≥ n, β(a, b) = C(a+b, a)).
int64_t binom(int64_t a, int64_t b) {  // C(a + b, a), saturating
  const int64_t cap = (int64_t)1 << 60;
  int64_t r = 1;
  for (int64_t i = 1; i <= b; ++i) {
    // r = r * (a + i) / i, exact at every step
    __int128 x = (__int128)r * (a + i);
    x /= i;
    if (x > cap) return cap;
    r = (int64_t)x;
  }
  return r;
}

int64_t reverse_cost(int64_t n, int64_t s) {
  if (n <= 1) return 0;
  if (s <= 0) return n * (n - 1) / 2;
  int64_t r = 1;
  while (binom(s + 1, r) < n) ++r;
  return r * n - binom(s + 2, r - 1);
}

// Optimal first checkpoint offset for reversing n >= 2 steps with s >= 1 free slots.
int64_t best_split(int64_t n, int64_t s) {
  int64_t best = 1, cost = INT64_MAX;
  for (int64_t m = 1; m < n; ++m) {
    int64_t c = m + reverse_cost(n - m, s - 1) + reverse_cost(m, s);
    if (c < cost) {
      cost = c;
      best = m;
    }
  }
  return best;
}

}  // namespace

using namespace lgd;

struct lgd_adjoint {
  lgd_ctx *ctx = nullptr;
  const lgd_grid *grid = nullptr;
  AdjGrid G{};
  int64_t ncell = 0, nfx = 0, nfy = 0, npad = 0;
  int64_t wg_cell = 0, wg_face = 0;  // workgroups of the cell and face passes
  // Scratch (global layout): padded state, incoming adjoint, per-cell and per-face terms.
  double *Pp = nullptr, *Up = nullptr, *AP = nullptr, *AU = nullptr;
  double *R = nullptr, *SX = nullptr, *SY = nullptr, *PW = nullptr, *PU = nullptr, *V = nullptr;
  double *GXx = nullptr, *GYx = nullptr, *GXy = nullptr, *GYy = nullptr;
  // θ accumulators, per workgroup: cell pass (λ, D, ε₄ centre) and face pass (ε₄ faces).
  double *acc_cell = nullptr, *acc_face = nullptr;
  // Streaming reverse step: the outgoing adjoint arrays (swapped with AP, AU after
  // each step) and 4 θ partials per wave, sized for the most waves any strip
  // height gives (two rows per strip).
  double *AP2 = nullptr, *AU2 = nullptr, *acc_stream = nullptr;
  int64_t nacc_stream = 0;
  bool stream_ok = false;
  double *sums = nullptr;  // reduction outputs
  std::vector<double *> allocs;
};

namespace {

// The reverse step's view of a 2-D Cartesian grid; false when the flat indices of
// the global arrays would not fit 32 bits.
bool adj_grid_of(const lgd_cart &c, AdjGrid &G) {
  const int64_t N0 = c.n[0] * c.nb[0], N1 = c.n[1] * c.nb[1];
  if ((N0 + 2) * (N1 + 2) >= ((int64_t)1 << 31)) return false;
  G.n0 = (int)c.n[0];
  G.n1 = (int)c.n[1];
  G.nb0 = (int)c.nb[0];
  G.N0 = (int)N0;
  G.N1 = (int)N1;
  G.sx = (int)(c.n[0] + 2);
  G.slab = (int64_t)G.sx * (c.n[1] + 2);
  G.per0 = c.periodic[0];
  G.per1 = c.periodic[1];
  G.P0 = G.N0 + 2;
  return true;
}

int adj_alloc(lgd_adjoint *a, int64_t n, double **p) {
  hipError_t e = hipMalloc((void **)p, (size_t)std::max<int64_t>(n, 1) * sizeof(double));
  if (e != hipSuccess) {
    *p = nullptr;
    return hip_fail(a->ctx, e, "adjoint workspace");
  }
  a->allocs.push_back(*p);
  return LGD_OK;
}

int zero(lgd_ctx *ctx, double *x, int64_t n) {
  if (n <= 0) return LGD_OK;
  adj_zero_kernel<<<blocks_for(n, kWG), kWG, 0, ctx->stream>>>(n, x);
  return launch_check(ctx, "adjoint zero");
}

int sum_strided(lgd_ctx *ctx, int64_t n, const double *x, int stride, double *out) {
  adj_sum_kernel<<<1, kSumThreads, 0, ctx->stream>>>(n, x, stride, out);
  return launch_check(ctx, "adjoint sum");
}

int sum_into(lgd_ctx *ctx, int64_t n, const double *x, double *out) {
  return sum_strided(ctx, n, x, 1, out);
}

}  // namespace

namespace {

// One reverse step.  slab = true: (pphi, pu) are slab fields, read and written in
// place; slab = false: the adjoint lives in the workspace's global arrays AP, AU
// (the calibration's reverse sweep) and pphi, pu are unused.
int adjoint_passes(lgd_adjoint *a, const lgd_modelc_params &params, const double *phi,
                   const double *u, double dt, double *pphi, double *pu, bool slab) {
  lgd_ctx *ctx = a->ctx;
  int rc;
  const lgd_kr_consts k = lgd_make_kr_consts(params, a->grid->cart);
  const AdjGrid &G = a->G;
  if (ctx->adjoint_kernel == 2 && !a->stream_ok)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "streaming adjoint step: unsupported block shape");
  if (a->stream_ok && ctx->adjoint_kernel != 1) {
    const int rows = (int)std::min<int64_t>(ctx->adjoint_rows, 1 << 20);  // 0: by occupancy
    if (slab) {
      adj_flatten_kernel<<<blocks_for(a->ncell, kWG), kWG, 0, ctx->stream>>>(G, pphi, pu, a->AP,
                                                                            a->AU);
      if ((rc = launch_check(ctx, "adjoint flatten"))) return rc;
    }
    rc = lgd_adjoint_stream_launch(ctx, a->grid->cart, k, dt, phi, u, a->AP, a->AU, a->AP2, a->AU2,
                                   a->acc_stream, rows, nullptr);
    if (rc) return rc;
    std::swap(a->AP, a->AP2);
    std::swap(a->AU, a->AU2);
    if (slab) {
      adj_unflatten_kernel<<<blocks_for(a->ncell, kWG), kWG, 0, ctx->stream>>>(G, a->AP, a->AU,
                                                                              pphi, pu);
      return launch_check(ctx, "adjoint unflatten");
    }
    return LGD_OK;
  }
  if (slab)
    adj_stage_kernel<true><<<blocks_for(a->npad, kWG), kWG, 0, ctx->stream>>>(
        G, phi, u, pphi, pu, a->Pp, a->Up, a->AP, a->AU);
  else
    adj_stage_kernel<false><<<blocks_for(a->npad, kWG), kWG, 0, ctx->stream>>>(
        G, phi, u, nullptr, nullptr, a->Pp, a->Up, a->AP, a->AU);
  if ((rc = launch_check(ctx, "adjoint stage"))) return rc;
  adj_cell_kernel<<<(unsigned)a->wg_cell, kWG, 0, ctx->stream>>>(
      G, k, dt, a->Pp, a->Up, a->AP, a->AU, a->R, a->SX, a->SY, a->PW, a->PU, a->V, a->acc_cell);
  if ((rc = launch_check(ctx, "adjoint cells"))) return rc;
  adj_face_kernel<<<(unsigned)a->wg_face, kWG, 0, ctx->stream>>>(G, k, a->Pp, a->R, a->GXx, a->GYx,
                                                                 a->GXy, a->GYy, a->acc_face);
  if ((rc = launch_check(ctx, "adjoint faces"))) return rc;
  if (slab)
    adj_gather_kernel<true><<<blocks_for(a->ncell, kWG), kWG, 0, ctx->stream>>>(
        G, k, a->GXx, a->GYx, a->GXy, a->GYy, a->SX, a->SY, a->PW, a->PU, a->V, pphi, pu);
  else
    adj_gather_kernel<false><<<blocks_for(a->ncell, kWG), kWG, 0, ctx->stream>>>(
        G, k, a->GXx, a->GYx, a->GXy, a->GYy, a->SX, a->SY, a->PW, a->PU, a->V, a->AP, a->AU);
  return launch_check(ctx, "adjoint gather");
}

}  // namespace

extern "C" {

int lgd_adjoint_create(lgd_ctx *ctx, const lgd_grid *grid, lgd_adjoint **out) {
  if (!ctx || !grid || !out) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  const lgd_cart &c = grid->cart;
  if (c.dim != 2) return fail(ctx, LGD_ERR_UNSUPPORTED, "the Model C adjoint is 2-D");
  lgd_adjoint *a = new (std::nothrow) lgd_adjoint();
  if (!a) return fail(ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  a->ctx = ctx;
  a->grid = grid;
  AdjGrid &G = a->G;
  if (!adj_grid_of(c, G)) {
    delete a;
    return fail(ctx, LGD_ERR_UNSUPPORTED, "adjoint domain above 2^31 cells");
  }
  a->ncell = (int64_t)G.N0 * G.N1;
  a->nfx = (int64_t)(G.N0 + 1) * G.N1;
  a->nfy = (int64_t)G.N0 * (G.N1 + 1);
  a->npad = (int64_t)G.P0 * (G.N1 + 2);
  a->wg_cell = (a->ncell + kWG - 1) / kWG;
  a->wg_face = ((int64_t)(G.N0 + 1) * (G.N1 + 1) + kWG - 1) / kWG;
  int rc = LGD_OK;
  double **cells[] = {&a->R, &a->SX, &a->SY, &a->PW, &a->PU, &a->V};
  for (double **p : cells)
    if (!rc) rc = adj_alloc(a, a->ncell, p);
  // The adjoint pairs swap roles each streaming step: both carry its output pad.
  if (!rc) rc = adj_alloc(a, a->ncell + kAdjStreamPad, &a->AP);
  if (!rc) rc = adj_alloc(a, a->ncell + kAdjStreamPad, &a->AU);
  if (!rc) rc = adj_alloc(a, a->npad, &a->Pp);
  if (!rc) rc = adj_alloc(a, a->npad, &a->Up);
  if (!rc) rc = adj_alloc(a, a->nfx, &a->GXx);
  if (!rc) rc = adj_alloc(a, a->nfx, &a->GYx);
  if (!rc) rc = adj_alloc(a, a->nfy, &a->GXy);
  if (!rc) rc = adj_alloc(a, a->nfy, &a->GYy);
  if (!rc) rc = adj_alloc(a, 3 * a->wg_cell, &a->acc_cell);
  if (!rc) rc = adj_alloc(a, a->wg_face, &a->acc_face);
  if (!rc) rc = adj_alloc(a, 8, &a->sums);
  {
    // The streaming step's wave count is largest at two rows per strip.
    int64_t nw = 0;
    const lgd_kr_consts k0{};
    a->stream_ok = lgd_adjoint_stream_launch(ctx, c, k0, 0.0, nullptr, nullptr, nullptr, nullptr,
                                             nullptr, nullptr, nullptr, 2, &nw) == LGD_OK;
    if (a->stream_ok) {
      a->nacc_stream = 4 * nw;
      if (!rc) rc = adj_alloc(a, a->ncell + kAdjStreamPad, &a->AP2);
      if (!rc) rc = adj_alloc(a, a->ncell + kAdjStreamPad, &a->AU2);
      if (!rc) rc = adj_alloc(a, a->nacc_stream, &a->acc_stream);
    }
  }
  if (!rc) rc = lgd_adjoint_reset(a);
  if (rc) {
    lgd_adjoint_destroy(a);
    return rc;
  }
  *out = a;
  return LGD_OK;
}

int lgd_adjoint_destroy(lgd_adjoint *a) {
  if (!a) return LGD_OK;
  for (double *p : a->allocs) (void)hipFree(p);
  delete a;
  return LGD_OK;
}

int lgd_adjoint_reset(lgd_adjoint *a) {
  if (!a) return LGD_ERR_INVALID_ARGUMENT;
  int rc;
  if ((rc = zero(a->ctx, a->acc_cell, 3 * a->wg_cell))) return rc;
  if (a->acc_stream && (rc = zero(a->ctx, a->acc_stream, a->nacc_stream))) return rc;
  return zero(a->ctx, a->acc_face, a->wg_face);
}

int lgd_modelc_adjoint_step(lgd_adjoint *a, const lgd_modelc_params *params, const lgd_buf *phi,
                            const lgd_buf *u, double dt, lgd_buf *pphi, lgd_buf *pu) {
  if (!a || !params) return LGD_ERR_INVALID_ARGUMENT;
  lgd_ctx *ctx = a->ctx;
  int rc;
  if ((rc = check_field(ctx, a->grid, phi, 1, "phi"))) return rc;
  if ((rc = check_field(ctx, a->grid, u, 1, "u"))) return rc;
  if ((rc = check_field(ctx, a->grid, pphi, 1, "pphi"))) return rc;
  if ((rc = check_field(ctx, a->grid, pu, 1, "pu"))) return rc;
  if (pphi == pu || pphi == phi || pphi == u || pu == phi || pu == u)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "adjoint buffers must be distinct from the state");
  return adjoint_passes(a, *params, phi->data, u->data, dt, pphi->data, pu->data, true);
}

int lgd_adjoint_grad(lgd_adjoint *a, double *grad) {
  if (!a || !grad) return LGD_ERR_INVALID_ARGUMENT;
  lgd_ctx *ctx = a->ctx;
  int rc;
  if ((rc = sum_strided(ctx, a->wg_cell, a->acc_cell, 3, a->sums))) return rc;
  if ((rc = sum_strided(ctx, a->wg_face, a->acc_face, 1, a->sums + 3))) return rc;
  double h[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
  const size_t nh = a->acc_stream ? 8 : 4;
  if (a->acc_stream && (rc = sum_strided(ctx, a->nacc_stream / 4, a->acc_stream, 4, a->sums + 4)))
    return rc;
  hipError_t e = hipMemcpyAsync(h, a->sums, nh * sizeof(double), hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "adjoint gradient download");
  grad[0] = h[0] + h[4];
  grad[1] = h[1] + h[5];
  grad[2] = (h[2] + h[3]) + (h[6] + h[7]);
  return LGD_OK;
}

}  // extern "C"

// ------------------------------------------------- the reverse step without allocation
// lgd_modelc_adjoint_step_noalloc keeps its workspace in one static device array per
// GPU instead of an lgd_adjoint: the flattened incoming and outgoing adjoint (four
// arrays of kNoAllocCells, each followed by the streaming step's output pad), the
// per-wave θ partials of the streaming reverse step and the reduced sums.  Host code never uses the array's name as an address: each
// device's copy is looked up once with hipGetSymbolAddress and kept in
// g_scratch_ptr.
constexpr int64_t kNoAllocCells = (int64_t)1 << 20;  // LGD_NOALLOC_ADJOINT_CELLS
constexpr int64_t kNoAllocAcc = (int64_t)1 << 16;    // θ partials: 4 per wave
constexpr int64_t kNoAllocArray = kNoAllocCells + kAdjStreamPad;
constexpr int64_t kNoAllocDoubles = 4 * kNoAllocArray + kNoAllocAcc + 8;

static __device__ __align__(256) double scratch[kNoAllocDoubles];

namespace {

constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
double *g_scratch_ptr[kMaxDevices] = {};

int noalloc_scratch(lgd_ctx *ctx, double **out) {
  const int dev = ctx->device;
  if (dev < 0 || dev >= kMaxDevices) return fail(ctx, LGD_ERR_UNSUPPORTED, "device index");
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    int prev = -1;
    hipError_t e = hipGetDevice(&prev);
    if (e == hipSuccess && prev != dev) e = hipSetDevice(dev);
    double *p = nullptr;
    if (e == hipSuccess) e = hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch));
    if (prev >= 0 && prev != dev) (void)hipSetDevice(prev);
    if (e != hipSuccess || !p) return hip_fail(ctx, e, "adjoint scratch symbol address");
    g_scratch_ptr[dev] = p;
  }
  *out = g_scratch_ptr[dev];
  return LGD_OK;
}

// grad = (0 + Σλ, 0 + ΣD, (0 + 0) + (Σε₄ centre + Σε₄ faces)): lgd_adjoint_grad's
// operations with the three-pass accumulators at zero, so both report the same bits.
__global__ void adj_noalloc_grad_kernel(const double *sums, double *grad) {
  if (threadIdx.x == 0) {
    grad[0] = 0.0 + sums[0];
    grad[1] = 0.0 + sums[1];
    grad[2] = (0.0 + 0.0) + (sums[2] + sums[3]);
  }
}

}  // namespace

extern "C" int lgd_modelc_adjoint_step_noalloc(lgd_ctx *ctx, const lgd_grid *grid,
                                               const lgd_modelc_params *params, const lgd_buf *phi,
                                               const lgd_buf *u, double dt, lgd_buf *pphi,
                                               lgd_buf *pu, lgd_buf *grad) {
  if (!ctx || !grid || !params || !grad) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  const lgd_cart &c = grid->cart;
  if (c.dim != 2) return fail(ctx, LGD_ERR_UNSUPPORTED, "the Model C adjoint is 2-D");
  int rc;
  if ((rc = check_field(ctx, grid, phi, 1, "phi"))) return rc;
  if ((rc = check_field(ctx, grid, u, 1, "u"))) return rc;
  if ((rc = check_field(ctx, grid, pphi, 1, "pphi"))) return rc;
  if ((rc = check_field(ctx, grid, pu, 1, "pu"))) return rc;
  if (pphi == pu || pphi == phi || pphi == u || pu == phi || pu == u)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "adjoint buffers must be distinct from the state");
  if (grad->len < 3) return fail(ctx, LGD_ERR_SIZE_MISMATCH, "grad holds fewer than 3 values");
  if (ctx->adjoint_kernel == 1)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "the three-pass reverse step needs an lgd_adjoint workspace");
  AdjGrid G{};
  const int64_t ncell = c.n[0] * c.nb[0] * c.n[1] * c.nb[1];
  if (ncell > kNoAllocCells || !adj_grid_of(c, G))
    return fail(ctx, LGD_ERR_UNSUPPORTED,
                "grid above the allocation-free reverse step's 2^20 cells (use lgd_adjoint_create)");
  const int rows = (int)std::min<int64_t>(ctx->adjoint_rows, 1 << 20);  // 0: by occupancy
  const lgd_kr_consts k = lgd_make_kr_consts(*params, c);
  int64_t nw = 0;
  if (lgd_adjoint_stream_launch(ctx, c, k, dt, nullptr, nullptr, nullptr, nullptr, nullptr,
                                nullptr, nullptr, rows, &nw) != LGD_OK)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "streaming reverse step: unsupported block shape");
  if (4 * nw > kNoAllocAcc)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "too many reverse-step waves for the static scratch");
  double *base = nullptr;
  if ((rc = noalloc_scratch(ctx, &base))) return rc;
  double *AP = base, *AU = base + kNoAllocArray, *AP2 = base + 2 * kNoAllocArray,
         *AU2 = base + 3 * kNoAllocArray, *acc = base + 4 * kNoAllocArray,
         *sums = acc + kNoAllocAcc;
  // The previous call's partials are still there: clear them with a kernel (a
  // memset node misbehaves under graph capture).
  if ((rc = zero(ctx, acc, 4 * nw))) return rc;
  adj_flatten_kernel<<<blocks_for(ncell, kWG), kWG, 0, ctx->stream>>>(G, pphi->data, pu->data, AP, AU);
  if ((rc = launch_check(ctx, "adjoint flatten"))) return rc;
  rc = lgd_adjoint_stream_launch(ctx, c, k, dt, phi->data, u->data, AP, AU, AP2, AU2, acc, rows,
                                 nullptr);
  if (rc) return rc;
  adj_unflatten_kernel<<<blocks_for(ncell, kWG), kWG, 0, ctx->stream>>>(G, AP2, AU2, pphi->data,
                                                                       pu->data);
  if ((rc = launch_check(ctx, "adjoint unflatten"))) return rc;
  if ((rc = sum_strided(ctx, nw, acc, 4, sums))) return rc;
  adj_noalloc_grad_kernel<<<1, 64, 0, ctx->stream>>>(sums, grad->data);
  return launch_check(ctx, "adjoint gradient");
}

// ------------------------------------------------------------------ calibration
struct lgd_calibration {
  lgd_ctx *ctx = nullptr;
  const lgd_grid *grid = nullptr;
  lgd_adjoint *adj = nullptr;
  int64_t nsteps = 0;
  double dt = 0.0;
  int32_t snapshots = 1;
  lgd_modelc_params params{};
  std::vector<lgd_buf *> slot_phi, slot_u;
  std::vector<int64_t> slot_pos;  // state index held by each slot (-1 = none)
  lgd_buf *w_phi[2] = {nullptr, nullptr}, *w_u[2] = {nullptr, nullptr};
  double *r2 = nullptr;
  bool forward_done = false;
  int64_t fwd_steps = 0, rev_steps = 0;
};

namespace {

int cal_free(lgd_calibration *c) {
  for (lgd_buf *b : c->slot_phi) lgd_buf_free(b);
  for (lgd_buf *b : c->slot_u) lgd_buf_free(b);
  for (int i = 0; i < 2; ++i) {
    lgd_buf_free(c->w_phi[i]);
    lgd_buf_free(c->w_u[i]);
  }
  if (c->r2) (void)hipFree(c->r2);
  lgd_adjoint_destroy(c->adj);
  delete c;
  return LGD_OK;
}

// Advance the state (sp, su) by m >= 1 forward steps.  Intermediate states
// alternate between the two work pairs (never the source); the last lands in
// (dp, du) when given (a checkpoint slot), else in a work pair.  *op, *ou receive
// where the result is.  No state is ever copied.
int advance_from(lgd_calibration *c, const lgd_buf *sp, const lgd_buf *su, int64_t m, lgd_buf *dp,
                 lgd_buf *du, const lgd_buf **op, const lgd_buf **ou) {
  const lgd_buf *cp = sp, *cu = su;
  int wi = 0;
  for (int64_t s = 0; s < m; ++s) {
    lgd_buf *np, *nu;
    if (s == m - 1 && dp) {
      np = dp;
      nu = du;
    } else {
      if (c->w_phi[wi] == cp) wi ^= 1;
      np = c->w_phi[wi];
      nu = c->w_u[wi];
      wi ^= 1;
    }
    int rc = lgd_modelc_step(c->ctx, c->grid, &c->params, cp, cu, np, nu, c->dt, 0, 0, 0);
    if (rc) return rc;
    ++c->fwd_steps;
    cp = np;
    cu = nu;
  }
  *op = cp;
  *ou = cu;
  return LGD_OK;
}

int reverse_step(lgd_calibration *c, const lgd_buf *phi, const lgd_buf *u) {
  ++c->rev_steps;
  return adjoint_passes(c->adj, c->params, phi->data, u->data, c->dt, nullptr, nullptr, false);
}

// Reverse steps [i0, i0 + n) given the state at i0 in slot `s0`, with slots
// s0+1 .. s0+free available.
int reverse(lgd_calibration *c, int64_t i0, int64_t n, int s0, int free) {
  int rc;
  const lgd_buf *p, *u;
  if (n <= 0) return LGD_OK;
  if (n == 1) return reverse_step(c, c->slot_phi[s0], c->slot_u[s0]);
  if (free == 0) {
    for (int64_t m = n - 1; m >= 0; --m) {
      p = c->slot_phi[s0];
      u = c->slot_u[s0];
      if (m > 0 && (rc = advance_from(c, p, u, m, nullptr, nullptr, &p, &u))) return rc;
      if ((rc = reverse_step(c, p, u))) return rc;
    }
    return LGD_OK;
  }
  const int64_t m = best_split(n, free);
  const int s1 = s0 + 1;
  if (c->slot_pos[s1] != i0 + m) {
    if ((rc = advance_from(c, c->slot_phi[s0], c->slot_u[s0], m, c->slot_phi[s1], c->slot_u[s1],
                           &p, &u)))
      return rc;
    c->slot_pos[s1] = i0 + m;
  }
  if ((rc = reverse(c, i0 + m, n - m, s1, free - 1))) return rc;
  return reverse(c, i0, m, s0, free);
}

}  // namespace

extern "C" {

int64_t lgd_checkpoint_forward_steps(int64_t nsteps, int32_t snapshots) {
  if (nsteps <= 0 || snapshots < 1) return 0;
  // The forward sweep (nsteps) + the recursion's advances not already laid down by it.
  int64_t chain = 0, i0 = 0, n = nsteps;
  int free = snapshots - 1;
  while (n >= 2 && free >= 1) {
    int64_t m = best_split(n, free);
    chain += m;
    i0 += m;
    n -= m;
    --free;
  }
  (void)i0;
  return nsteps + reverse_cost(nsteps, snapshots - 1) - chain;
}

int lgd_calibration_create(lgd_ctx *ctx, const lgd_grid *grid, int64_t nsteps, double dt,
                           int32_t snapshots, lgd_calibration **out) {
  if (!ctx || !grid || !out) return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "null argument");
  *out = nullptr;
  if (nsteps < 1 || snapshots < 1)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "nsteps and snapshots must be at least 1");
  lgd_calibration *c = new (std::nothrow) lgd_calibration();
  if (!c) return fail(ctx, LGD_ERR_OUT_OF_MEMORY, "host allocation failed");
  c->ctx = ctx;
  c->grid = grid;
  c->nsteps = nsteps;
  c->dt = dt;
  c->snapshots = (int32_t)std::min<int64_t>(snapshots, nsteps);
  int rc = lgd_adjoint_create(ctx, grid, &c->adj);
  const size_t len = (size_t)(grid->cart.nblocks * slab_len(grid->cart, 1));
  for (int s = 0; !rc && s < c->snapshots; ++s) {
    lgd_buf *p = nullptr, *q = nullptr;
    rc = lgd_buf_alloc(ctx, len, &p);
    if (!rc) rc = lgd_buf_alloc(ctx, len, &q);
    c->slot_phi.push_back(p);
    c->slot_u.push_back(q);
    c->slot_pos.push_back(-1);
  }
  for (int i = 0; !rc && i < 2; ++i) {
    rc = lgd_buf_alloc(ctx, len, &c->w_phi[i]);
    if (!rc) rc = lgd_buf_alloc(ctx, len, &c->w_u[i]);
  }
  if (!rc) {
    hipError_t e = hipMalloc((void **)&c->r2, (size_t)(c->adj->ncell + 1) * sizeof(double));
    if (e != hipSuccess) {
      c->r2 = nullptr;
      rc = hip_fail(ctx, e, "calibration workspace");
    }
  }
  if (rc) {
    cal_free(c);
    return rc;
  }
  *out = c;
  return LGD_OK;
}

int lgd_calibration_destroy(lgd_calibration *c) {
  if (!c) return LGD_OK;
  return cal_free(c);
}

int lgd_calibration_forward(lgd_calibration *c, const lgd_modelc_params *params,
                            const lgd_buf *phi0, const lgd_buf *u0, const lgd_buf *target,
                            double *loss) {
  if (!c || !params || !loss) return LGD_ERR_INVALID_ARGUMENT;
  lgd_ctx *ctx = c->ctx;
  int rc;
  if ((rc = check_field(ctx, c->grid, phi0, 1, "phi0"))) return rc;
  if ((rc = check_field(ctx, c->grid, u0, 1, "u0"))) return rc;
  if ((rc = check_field(ctx, c->grid, target, 1, "target"))) return rc;
  if (params->noise_amplitude != 0.0)
    return fail(ctx, LGD_ERR_UNSUPPORTED, "the adjoint differentiates the b = 0 step");
  c->params = *params;
  c->fwd_steps = c->rev_steps = 0;
  c->forward_done = false;
  for (auto &p : c->slot_pos) p = -1;
  // Slot 0 holds the initial state; the sweep lays down the schedule's first chain,
  // each checkpoint written by the step that produces it.
  if ((rc = lgd_buf_copy(ctx, c->slot_phi[0], phi0))) return rc;
  if ((rc = lgd_buf_copy(ctx, c->slot_u[0], u0))) return rc;
  c->slot_pos[0] = 0;
  const lgd_buf *p = c->slot_phi[0], *u = c->slot_u[0];
  int64_t i0 = 0, n = c->nsteps;
  int free = c->snapshots - 1, s = 0;
  while (n >= 2 && free >= 1) {
    const int64_t m = best_split(n, free);
    ++s;
    if ((rc = advance_from(c, p, u, m, c->slot_phi[s], c->slot_u[s], &p, &u))) return rc;
    c->slot_pos[s] = i0 + m;
    i0 += m;
    n -= m;
    --free;
  }
  if ((rc = advance_from(c, p, u, n, nullptr, nullptr, &p, &u))) return rc;
  // Terminal adjoint (in the workspace's global arrays) and loss.
  const AdjGrid &G = c->adj->G;
  adj_loss_kernel<<<blocks_for(c->adj->ncell, kWG), kWG, 0, ctx->stream>>>(
      G, p->data, target->data, c->adj->AP, c->adj->AU, c->r2);
  if ((rc = launch_check(ctx, "calibration loss"))) return rc;
  if ((rc = sum_into(ctx, c->adj->ncell, c->r2, c->r2 + c->adj->ncell))) return rc;
  double s2 = 0.0;
  hipError_t e = hipMemcpyAsync(&s2, c->r2 + c->adj->ncell, sizeof s2, hipMemcpyDeviceToHost,
                                ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e != hipSuccess) return hip_fail(ctx, e, "calibration loss download");
  *loss = 0.5 * s2;
  c->forward_done = true;
  return LGD_OK;
}

int lgd_calibration_backward(lgd_calibration *c, double *grad, lgd_buf *gphi0, lgd_buf *gu0) {
  if (!c || !grad) return LGD_ERR_INVALID_ARGUMENT;
  lgd_ctx *ctx = c->ctx;
  if (!c->forward_done)
    return fail(ctx, LGD_ERR_INVALID_ARGUMENT, "backward needs a preceding forward");
  int rc;
  if (gphi0 && (rc = check_field(ctx, c->grid, gphi0, 1, "gphi0"))) return rc;
  if (gu0 && (rc = check_field(ctx, c->grid, gu0, 1, "gu0"))) return rc;
  if ((rc = lgd_adjoint_reset(c->adj))) return rc;
  if ((rc = reverse(c, 0, c->nsteps, 0, c->snapshots - 1))) return rc;
  if ((rc = lgd_adjoint_grad(c->adj, grad))) return rc;
  if (gphi0 || gu0) {
    adj_unflatten_kernel<<<blocks_for(c->adj->ncell, kWG), kWG, 0, ctx->stream>>>(
        c->adj->G, c->adj->AP, c->adj->AU, gphi0 ? gphi0->data : nullptr, gu0 ? gu0->data : nullptr);
    if ((rc = launch_check(ctx, "calibration gradient fields"))) return rc;
  }
  c->forward_done = false;
  return LGD_OK;
}

int lgd_calibration_counts(const lgd_calibration *c, int64_t *forward_steps,
                           int64_t *reverse_steps) {
  if (!c) return LGD_ERR_INVALID_ARGUMENT;
  if (forward_steps) *forward_steps = c->fwd_steps;
  if (reverse_steps) *reverse_steps = c->rev_steps;
  return LGD_OK;
}

}  // extern "C"
