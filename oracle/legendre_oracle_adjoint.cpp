// legendre CPU oracle, derivative part — TEST INFRASTRUCTURE ONLY (see legendre_oracle.h).
//
// The reference has no derivative code.  These are the checkers for the device
// adjoint of the explicit Model C step (Euler–Maruyama with b = 0, which is
// forward Euler bit for bit: src/integrators/euler_maruyama.rs:34-75,
// tests/convergence.rs:100-107), built as two independent derivations:
//
//  * a tangent-linear model: the forward step's exact operation sequence (that of
//    orc_modelc_run: KarmaRappelFlux::kernel, ModelC::center_anisotropy,
//    ModelC::vector_field_block, CentralLaplacian, apply_step — finite_volume.rs:113-142,
//    248-284; phasefield.rs:100-116, 159-175, 394-421; finite_difference.rs:49-67;
//    state.rs:656-699) run on dual numbers, so the value part equals the forward
//    bit for bit and the tangent part is the exact derivative of that sequence;
//  * a hand-written reverse-mode adjoint of one step (scatter form, global
//    coordinates), whose transpose identity against the tangent-linear model is
//    checked in tests/test_oracle_adjoint.py.
//
// Parameters θ = (λ, D, ε₄).  ε₄ enters through ā = fma(3, −ε₄, 1), ε′ = 3ε₄/ā and
// a₁₂ = 4āε′ (finite_volume.rs:113-122, phasefield.rs:100-116).
//
// Halos: with copy-type boundary rules (mirror, periodic) the dimension-sweep ghost
// fill (src/geometry/cartesian.rs:439-526) gives every ghost the value of the cell
// at the per-dimension reflected / wrapped global coordinate, so the step is written
// on the global grid with resolved neighbour coordinates.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "legendre_oracle.h"

namespace {

// Dual number with K tangent directions (forward-mode derivative).
template <int K>
struct DualN {
  double v;
  double d[K];
  DualN() = default;
  DualN(double x) : v(x) {  // NOLINT: constants enter as exact duals
    for (int k = 0; k < K; ++k) d[k] = 0.0;
  }
};
template <int K>
inline DualN<K> operator+(const DualN<K> &a, const DualN<K> &b) {
  DualN<K> r;
  r.v = a.v + b.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] + b.d[k];
  return r;
}
template <int K>
inline DualN<K> operator-(const DualN<K> &a, const DualN<K> &b) {
  DualN<K> r;
  r.v = a.v - b.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] - b.d[k];
  return r;
}
template <int K>
inline DualN<K> operator-(const DualN<K> &a) {
  DualN<K> r;
  r.v = -a.v;
  for (int k = 0; k < K; ++k) r.d[k] = -a.d[k];
  return r;
}
template <int K>
inline DualN<K> operator*(const DualN<K> &a, const DualN<K> &b) {
  DualN<K> r;
  r.v = a.v * b.v;
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k];
  return r;
}
template <int K>
inline DualN<K> operator/(const DualN<K> &a, const DualN<K> &b) {
  DualN<K> r;
  r.v = a.v / b.v;
  for (int k = 0; k < K; ++k) r.d[k] = (a.d[k] - r.v * b.d[k]) / b.v;
  return r;
}
// Mixed operands: a plain double is a constant.
template <int K> inline DualN<K> operator+(double a, const DualN<K> &b) { return DualN<K>(a) + b; }
template <int K> inline DualN<K> operator*(double a, const DualN<K> &b) { return DualN<K>(a) * b; }
inline double val(double x) { return x; }
template <int K>
inline double val(const DualN<K> &x) {
  return x.v;
}
inline double mfma(double a, double b, double c) { return std::fma(a, b, c); }
template <int K>
inline DualN<K> mfma(const DualN<K> &a, const DualN<K> &b, const DualN<K> &c) {
  DualN<K> r;
  r.v = std::fma(a.v, b.v, c.v);
  for (int k = 0; k < K; ++k) r.d[k] = a.d[k] * b.v + a.v * b.d[k] + c.d[k];
  return r;
}

// 2-D global grid with per-dimension mirror (reflect_coord, src/geometry/boundary.rs:82-90)
// or periodic wrap.
struct Grid2 {
  int64_t n0, n1;
  int per0, per1;
  double hx, hy;
  static int64_t resolve(int64_t i, int64_t n, int per) {
    if (i < 0) return per ? i + n : -1 - i;
    if (i >= n) return per ? i - n : 2 * n - 1 - i;
    return i;
  }
  int64_t rx(int64_t i) const { return resolve(i, n0, per0); }
  int64_t ry(int64_t j) const { return resolve(j, n1, per1); }
  int64_t at(int64_t i, int64_t j) const { return rx(i) + n0 * ry(j); }
};

Grid2 grid2(const orc_grid *g) {
  Grid2 G;
  G.n0 = g->cells[0];
  G.n1 = g->cells[1];
  G.per0 = g->periodic[0];
  G.per1 = g->periodic[1];
  G.hx = g->spacing[0];
  G.hy = g->spacing[1];
  return G;
}

// Field buffer (ghost-1 slabs, block-major) ↔ global interior array (i fastest).
void to_global(const orc_grid *g, const double *field, double *out) {
  int64_t n0 = g->block_cells[0], n1 = g->block_cells[1];
  int64_t nb0 = g->cells[0] / n0, nb1 = g->cells[1] / n1;
  int64_t sx = n0 + 2, len = sx * (n1 + 2);
  for (int64_t by = 0; by < nb1; ++by)
    for (int64_t bx = 0; bx < nb0; ++bx) {
      const double *s = field + (bx + nb0 * by) * len;
      for (int64_t j = 0; j < n1; ++j)
        for (int64_t i = 0; i < n0; ++i)
          out[(bx * n0 + i) + g->cells[0] * (by * n1 + j)] = s[(j + 1) * sx + (i + 1)];
    }
}

void from_global(const orc_grid *g, const double *in, double *field) {
  int64_t n0 = g->block_cells[0], n1 = g->block_cells[1];
  int64_t nb0 = g->cells[0] / n0, nb1 = g->cells[1] / n1;
  int64_t sx = n0 + 2, len = sx * (n1 + 2);
  for (int64_t by = 0; by < nb1; ++by)
    for (int64_t bx = 0; bx < nb0; ++bx) {
      double *s = field + (bx + nb0 * by) * len;
      for (int64_t j = 0; j < n1; ++j)
        for (int64_t i = 0; i < n0; ++i)
          s[(j + 1) * sx + (i + 1)] = in[(bx * n0 + i) + g->cells[0] * (by * n1 + j)];
    }
}

template <class T>
struct Consts {
  T lambda, d, a_bar, eps_prime, a12;
  double fv_tol, model_tol, inv_h2, ihx2, ihy2;
};

template <class T>
Consts<T> make_consts(const orc_modelc_params *p, const Grid2 &G, T lambda, T d, T eps4) {
  Consts<T> k;
  k.lambda = lambda;
  k.d = d;
  k.a_bar = mfma(T(3.0), -eps4, T(1.0));
  k.eps_prime = T(3.0) * eps4 / k.a_bar;
  k.a12 = T(4.0) * k.a_bar * k.eps_prime;
  k.fv_tol = p->fv_tol;
  k.model_tol = p->anisotropy_tol;
  k.inv_h2 = 1.0 / (G.hx * G.hy);
  k.ihx2 = 1.0 / (G.hx * G.hx);
  k.ihy2 = 1.0 / (G.hy * G.hy);
  return k;
}

// face(dx, dy) (finite_volume.rs:129-142).
template <class T>
inline void face(const Consts<T> &k, T dx, T dy, T &a, T &da) {
  T dx2 = dx * dx;
  T dy2 = dy * dy;
  T sum = dx2 + dy2;
  T mag2 = sum * sum;
  if (val(mag2) <= k.fv_tol) {
    a = k.a_bar;
    da = T(0.0);
  } else {
    a = k.a_bar * (T(1.0) + k.eps_prime * (dx2 * dx2 + dy2 * dy2) / mag2);
    da = -(k.a12) * dx * dy * (dx2 - dy2) / mag2;
  }
}

// One forward-Euler step on the global grid, in the forward's operation order.
template <class T>
void step_global(const Grid2 &G, const Consts<T> &k, double dt, const T *phi, const T *u,
                 T *phi_n, T *u_n) {
  const T q(0.25);
#pragma omp parallel for schedule(static)
  for (int64_t j = 0; j < G.n1; ++j)
    for (int64_t i = 0; i < G.n0; ++i) {
      auto p = [&](int64_t a, int64_t b) { return phi[G.at(a, b)]; };
      auto uu = [&](int64_t a, int64_t b) { return u[G.at(a, b)]; };
      T derx_r = p(i + 1, j) - p(i, j);
      T derx_l = p(i, j) - p(i - 1, j);
      T derx_t = q * (p(i + 1, j + 1) - p(i - 1, j + 1) + p(i + 1, j) - p(i - 1, j));
      T derx_b = q * (p(i + 1, j) - p(i - 1, j) + p(i + 1, j - 1) - p(i - 1, j - 1));
      T dery_t = p(i, j + 1) - p(i, j);
      T dery_b = p(i, j) - p(i, j - 1);
      T dery_r = q * (p(i + 1, j + 1) - p(i + 1, j - 1) + p(i, j + 1) - p(i, j - 1));
      T dery_l = q * (p(i, j + 1) - p(i, j - 1) + p(i - 1, j + 1) - p(i - 1, j - 1));
      T a_r, da_r, a_l, da_l, a_t, da_t, a_b, da_b;
      face(k, derx_r, dery_r, a_r, da_r);
      face(k, derx_l, dery_l, a_l, da_l);
      face(k, derx_t, dery_t, a_t, da_t);
      face(k, derx_b, dery_b, a_b, da_b);
      T j_r = a_r * (a_r * derx_r - da_r * dery_r);
      T j_l = a_l * (a_l * derx_l - da_l * dery_l);
      T j_t = a_t * (a_t * dery_t + da_t * derx_t);
      T j_b = a_b * (a_b * dery_b + da_b * derx_b);
      T flux = (j_r - j_l + j_t - j_b) * T(k.inv_h2);
      const int64_t c = i + G.n0 * j;
      T pv = phi[c];
      T g_prime = mfma(pv * pv, pv, -pv);
      T one_m_p2 = mfma(pv, -pv, T(1.0));
      T p_prime = one_m_p2 * one_m_p2;
      // center_anisotropy (phasefield.rs:159-175).
      T cdx = T(0.5) * (p(i + 1, j) - p(i - 1, j));
      T cdy = T(0.5) * (p(i, j + 1) - p(i, j - 1));
      T cdx2 = cdx * cdx, cdy2 = cdy * cdy;
      T csum = cdx2 + cdy2;
      T cmag2 = csum * csum;
      T a = val(cmag2) <= k.model_tol
                ? k.a_bar
                : k.a_bar * (T(1.0) + k.eps_prime * (cdx2 * cdx2 + cdy2 * cdy2) / cmag2);
      T kphi = mfma(k.lambda * u[c], -p_prime, flux - g_prime) / (a * a);
      T uc = u[c];
      T acc(0.0);
      acc = acc + (uu(i + 1, j) - (uc + uc) + uu(i - 1, j)) * T(k.ihx2);
      acc = acc + (uu(i, j + 1) - (uc + uc) + uu(i, j - 1)) * T(k.ihy2);
      T ku = mfma(acc, k.d, T(0.5) * kphi);
      phi_n[c] = pv + T(dt) * kphi;
      u_n[c] = uc + T(dt) * ku;
    }
}

// Partial derivatives of face(dx, dy) and of the centre anisotropy (same regularized
// function): a, da and their derivatives w.r.t. dx, dy and ε₄.
struct FaceD {
  double a, da, a_x, a_y, a_e, da_x, da_y, da_e;
};

inline FaceD face_deriv(const Consts<double> &k, double tol, double dx, double dy) {
  FaceD f;
  double dx2 = dx * dx, dy2 = dy * dy;
  double s = dx2 + dy2;
  double m = s * s;
  if (m <= tol) {
    f.a = k.a_bar;
    f.da = 0.0;
    f.a_x = f.a_y = f.da_x = f.da_y = f.da_e = 0.0;
    f.a_e = -3.0;  // dā/dε₄
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
  // dā/dε₄ = −3, dε′/dε₄ = 3/ā², da₁₂/dε₄ = 12.
  f.a_e = -3.0 * (1.0 + k.eps_prime * Q / m) + k.a_bar * (3.0 / (k.a_bar * k.a_bar)) * Q / m;
  f.da_e = -12.0 * H / m;
  return f;
}

}  // namespace

extern "C" {

}  // extern "C"

namespace {

template <int K>
void tlm_run_k(const orc_grid *g, const orc_modelc_params *p, double *phi, double *u,
               int32_t ndir, double *dphi, double *du, const double *dtheta, int64_t nsteps,
               double dt) {
  using T = DualN<K>;
  Grid2 G = grid2(g);
  const size_t n = (size_t)(G.n0 * G.n1);
  const size_t len = (size_t)(orc_num_blocks(g) * orc_block_len(g, 1));
  std::vector<double> a(n), b(n), c(n);
  to_global(g, phi, a.data());
  to_global(g, u, b.data());
  std::vector<T> P(n), U(n), Pn(n), Un(n);
  for (size_t i = 0; i < n; ++i) {
    P[i] = T(a[i]);
    U[i] = T(b[i]);
  }
  for (int k = 0; k < ndir; ++k) {
    to_global(g, dphi + k * len, c.data());
    for (size_t i = 0; i < n; ++i) P[i].d[k] = c[i];
    to_global(g, du + k * len, c.data());
    for (size_t i = 0; i < n; ++i) U[i].d[k] = c[i];
  }
  T th[3] = {T(p->lambda), T(p->d_thermal), T(p->eps4)};
  for (int k = 0; k < ndir; ++k)
    for (int q = 0; q < 3; ++q) th[q].d[k] = dtheta[3 * k + q];
  Consts<T> kc = make_consts<T>(p, G, th[0], th[1], th[2]);
  for (int64_t s = 0; s < nsteps; ++s) {
    step_global(G, kc, dt, P.data(), U.data(), Pn.data(), Un.data());
    P.swap(Pn);
    U.swap(Un);
  }
  for (size_t i = 0; i < n; ++i) {
    a[i] = P[i].v;
    b[i] = U[i].v;
  }
  from_global(g, a.data(), phi);
  from_global(g, b.data(), u);
  for (int k = 0; k < ndir; ++k) {
    for (size_t i = 0; i < n; ++i) c[i] = P[i].d[k];
    from_global(g, c.data(), dphi + k * len);
    for (size_t i = 0; i < n; ++i) c[i] = U[i].d[k];
    from_global(g, c.data(), du + k * len);
  }
}

}  // namespace

extern "C" {

int orc_modelc_tlm_run(const orc_grid *g, const orc_modelc_params *p, double *phi, double *u,
                       int32_t ndir, double *dphi, double *du, const double *dtheta,
                       int64_t nsteps, double dt) {
  if (g->dim != 2 || orc_grid_check(g) != 0 || ndir < 1 || ndir > 4) return -1;
  if (ndir == 1)
    tlm_run_k<1>(g, p, phi, u, ndir, dphi, du, dtheta, nsteps, dt);
  else if (ndir == 2)
    tlm_run_k<2>(g, p, phi, u, ndir, dphi, du, dtheta, nsteps, dt);
  else
    tlm_run_k<4>(g, p, phi, u, ndir, dphi, du, dtheta, nsteps, dt);
  return 0;
}

int orc_modelc_adjoint_step(const orc_grid *g, const orc_modelc_params *p, const double *phi,
                            const double *u, double dt, double *pphi, double *pu, double *grad) {
  if (g->dim != 2 || orc_grid_check(g) != 0) return -1;
  Grid2 G = grid2(g);
  const size_t n = (size_t)(G.n0 * G.n1);
  std::vector<double> P(n), U(n), AP(n), AU(n), OP(n, 0.0), OU(n, 0.0);
  to_global(g, phi, P.data());
  to_global(g, u, U.data());
  to_global(g, pphi, AP.data());
  to_global(g, pu, AU.data());
  Consts<double> k = make_consts<double>(p, G, p->lambda, p->d_thermal, p->eps4);
  const double q = 0.25;
  double g_lambda = 0.0, g_d = 0.0, g_eps = 0.0;
  auto ph = [&](int64_t a, int64_t b) { return P[G.at(a, b)]; };
  auto add = [&](std::vector<double> &v, int64_t a, int64_t b, double x) { v[G.at(a, b)] += x; };
  for (int64_t j = 0; j < G.n1; ++j)
    for (int64_t i = 0; i < G.n0; ++i) {
      const int64_t c = i + G.n0 * j;
      // Forward quantities at c (same formulas as step_global).
      double derx[4], dery[4];  // faces R, L, T, B
      derx[0] = ph(i + 1, j) - ph(i, j);
      derx[1] = ph(i, j) - ph(i - 1, j);
      derx[2] = q * (ph(i + 1, j + 1) - ph(i - 1, j + 1) + ph(i + 1, j) - ph(i - 1, j));
      derx[3] = q * (ph(i + 1, j) - ph(i - 1, j) + ph(i + 1, j - 1) - ph(i - 1, j - 1));
      dery[2] = ph(i, j + 1) - ph(i, j);
      dery[3] = ph(i, j) - ph(i, j - 1);
      dery[0] = q * (ph(i + 1, j + 1) - ph(i + 1, j - 1) + ph(i, j + 1) - ph(i, j - 1));
      dery[1] = q * (ph(i, j + 1) - ph(i, j - 1) + ph(i - 1, j + 1) - ph(i - 1, j - 1));
      FaceD f[4];
      double J[4];
      for (int s = 0; s < 4; ++s) {
        f[s] = face_deriv(k, k.fv_tol, derx[s], dery[s]);
        J[s] = s < 2 ? f[s].a * (f[s].a * derx[s] - f[s].da * dery[s])
                     : f[s].a * (f[s].a * dery[s] + f[s].da * derx[s]);
      }
      double flux = (J[0] - J[1] + J[2] - J[3]) * k.inv_h2;
      double pv = P[c], uc = U[c];
      double g_prime = std::fma(pv * pv, pv, -pv);
      double one_m_p2 = std::fma(pv, -pv, 1.0);
      double p_prime = one_m_p2 * one_m_p2;
      double cdx = 0.5 * (ph(i + 1, j) - ph(i - 1, j));
      double cdy = 0.5 * (ph(i, j + 1) - ph(i, j - 1));
      FaceD ca = face_deriv(k, k.model_tol, cdx, cdy);
      double A = ca.a;
      double N = std::fma(k.lambda * uc, -p_prime, flux - g_prime);
      double F = N / (A * A);
      double acc = 0.0;
      acc = acc + (U[G.at(i + 1, j)] - (uc + uc) + U[G.at(i - 1, j)]) * k.ihx2;
      acc = acc + (U[G.at(i, j + 1)] - (uc + uc) + U[G.at(i, j - 1)]) * k.ihy2;
      // Reverse sweep.  φ' = φ + dt·F, u' = u + dt·(D·acc + ½F).
      const double W = dt * (AP[c] + 0.5 * AU[c]);  // ∂L/∂F
      const double V = dt * k.d * AU[c];             // ∂L/∂acc
      g_d += dt * AU[c] * acc;
      const double R = W / (A * A);  // ∂L/∂N
      const double S = -2.0 * W * F / A;  // ∂L/∂A
      g_lambda += R * (-uc * p_prime);
      OP[c] += AP[c] + R * (-(3.0 * pv * pv - 1.0) - k.lambda * uc * (2.0 * one_m_p2 * (-2.0 * pv)));
      OU[c] += AU[c] + R * (-k.lambda * p_prime);
      // Centre anisotropy.
      g_eps += S * ca.a_e;
      add(OP, i + 1, j, S * ca.a_x * 0.5);
      add(OP, i - 1, j, -S * ca.a_x * 0.5);
      add(OP, i, j + 1, S * ca.a_y * 0.5);
      add(OP, i, j - 1, -S * ca.a_y * 0.5);
      // Faces: ∂flux/∂J = ±1/(hx·hy).
      const double sign[4] = {1.0, -1.0, 1.0, -1.0};
      for (int s = 0; s < 4; ++s) {
        const double GJ = R * sign[s] * k.inv_h2;
        const FaceD &h = f[s];
        double j_x, j_y, j_e;
        if (s < 2) {  // J = a(a·derx − da·dery)
          double inner = h.a * derx[s] - h.da * dery[s];
          j_x = h.a_x * inner + h.a * (h.a_x * derx[s] + h.a - h.da_x * dery[s]);
          j_y = h.a_y * inner + h.a * (h.a_y * derx[s] - h.da_y * dery[s] - h.da);
          j_e = h.a_e * inner + h.a * (h.a_e * derx[s] - h.da_e * dery[s]);
        } else {  // J = a(a·dery + da·derx)
          double inner = h.a * dery[s] + h.da * derx[s];
          j_x = h.a_x * inner + h.a * (h.a_x * dery[s] + h.da_x * derx[s] + h.da);
          j_y = h.a_y * inner + h.a * (h.a_y * dery[s] + h.a + h.da_y * derx[s]);
          j_e = h.a_e * inner + h.a * (h.a_e * dery[s] + h.da_e * derx[s]);
        }
        g_eps += GJ * j_e;
        const double gx = GJ * j_x, gy = GJ * j_y;
        switch (s) {
          case 0:  // x-face (i|i+1): derx normal, dery transverse over columns i, i+1
            add(OP, i + 1, j, gx);
            add(OP, i, j, -gx);
            add(OP, i + 1, j + 1, q * gy);
            add(OP, i + 1, j - 1, -q * gy);
            add(OP, i, j + 1, q * gy);
            add(OP, i, j - 1, -q * gy);
            break;
          case 1:  // x-face (i-1|i)
            add(OP, i, j, gx);
            add(OP, i - 1, j, -gx);
            add(OP, i, j + 1, q * gy);
            add(OP, i, j - 1, -q * gy);
            add(OP, i - 1, j + 1, q * gy);
            add(OP, i - 1, j - 1, -q * gy);
            break;
          case 2:  // y-face (j|j+1): dery normal, derx transverse over rows j, j+1
            add(OP, i, j + 1, gy);
            add(OP, i, j, -gy);
            add(OP, i + 1, j + 1, q * gx);
            add(OP, i - 1, j + 1, -q * gx);
            add(OP, i + 1, j, q * gx);
            add(OP, i - 1, j, -q * gx);
            break;
          default:  // y-face (j-1|j)
            add(OP, i, j, gy);
            add(OP, i, j - 1, -gy);
            add(OP, i + 1, j, q * gx);
            add(OP, i - 1, j, -q * gx);
            add(OP, i + 1, j - 1, q * gx);
            add(OP, i - 1, j - 1, -q * gx);
            break;
        }
      }
      // Laplacian of u (symmetric stencil with the resolved neighbours).
      add(OU, i + 1, j, V * k.ihx2);
      add(OU, i - 1, j, V * k.ihx2);
      add(OU, i, j + 1, V * k.ihy2);
      add(OU, i, j - 1, V * k.ihy2);
      OU[c] += -2.0 * V * (k.ihx2 + k.ihy2);
    }
  from_global(g, OP.data(), pphi);
  from_global(g, OU.data(), pu);
  grad[0] += g_lambda;
  grad[1] += g_d;
  grad[2] += g_eps;
  return 0;
}

int orc_modelc_gradient(const orc_grid *g, const orc_modelc_params *p, const double *phi0,
                        const double *u0, const double *target, int64_t nsteps, double dt,
                        double *loss, double *grad, double *gphi0, double *gu0) {
  if (g->dim != 2 || orc_grid_check(g) != 0 || nsteps < 0) return -1;
  const size_t len = (size_t)(orc_num_blocks(g) * orc_block_len(g, 1));
  std::vector<std::vector<double>> ph(nsteps + 1), uu(nsteps + 1);
  ph[0].assign(phi0, phi0 + len);
  uu[0].assign(u0, u0 + len);
  for (int64_t s = 0; s < nsteps; ++s) {
    ph[s + 1] = ph[s];
    uu[s + 1] = uu[s];
    double t = 0.0;
    if (orc_modelc_run(g, p, ph[s + 1].data(), uu[s + 1].data(), 1, dt, &t, 0, 0, 1) != 0)
      return -1;
  }
  Grid2 G = grid2(g);
  const size_t n = (size_t)(G.n0 * G.n1);
  std::vector<double> fin(n), tgt(n), r(n);
  to_global(g, ph[nsteps].data(), fin.data());
  to_global(g, target, tgt.data());
  double l = 0.0;
  for (size_t i = 0; i < n; ++i) {
    r[i] = fin[i] - tgt[i];
    l += r[i] * r[i];
  }
  *loss = 0.5 * l;
  std::vector<double> ap(len, 0.0), au(len, 0.0);
  from_global(g, r.data(), ap.data());
  grad[0] = grad[1] = grad[2] = 0.0;
  for (int64_t s = nsteps - 1; s >= 0; --s)
    if (orc_modelc_adjoint_step(g, p, ph[s].data(), uu[s].data(), dt, ap.data(), au.data(), grad) != 0)
      return -1;
  std::memcpy(gphi0, ap.data(), len * sizeof(double));
  std::memcpy(gu0, au.data(), len * sizeof(double));
  return 0;
}

}  // extern "C"
