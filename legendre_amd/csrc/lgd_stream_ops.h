// Row-streaming building blocks shared by the Model C streaming kernels (the
// Cartesian step in lgd_modelc_stream.hip, the AMR level stage in lgd_amr.hip):
// one-lane DPP shifts, the Karma–Rappel face fluxes, the pointwise cell update and
// scalar loads of wave-uniform table entries.
#pragma once

#include "lgd_internal.h"

namespace lgd_stream {

// Cross-lane moves by one lane (DPP wavefront shifts: no LDS round trip).  The
// lane with no source reads 0; it is never an interior column's neighbour.
__device__ __forceinline__ double dpp_shift(double v, int ctrl_is_right) {
  int lo = __double2loint(v), hi = __double2hiint(v);
  if (ctrl_is_right) {
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x138, 0xf, 0xf, true);  // wave_shr:1
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x138, 0xf, 0xf, true);
  } else {
    lo = __builtin_amdgcn_update_dpp(0, lo, 0x130, 0xf, 0xf, true);  // wave_shl:1
    hi = __builtin_amdgcn_update_dpp(0, hi, 0x130, 0xf, 0xf, true);
  }
  return __hiloint2double(hi, lo);
}
// Value of the lane to the left (lane - 1) / right (lane + 1).
__device__ __forceinline__ double from_left(double v) { return dpp_shift(v, 1); }
__device__ __forceinline__ double from_right(double v) { return dpp_shift(v, 0); }

// Karma–Rappel flux through an x-face: J = A·(A·derx − A′·dery).
__device__ __forceinline__ double flux_x(const lgd_kr_consts &k, double derx, double dery) {
  double a, da;
  lgd_kr_face(k, derx, dery, a, da);
  return a * ((a * derx) - (da * dery));
}

// Karma–Rappel flux through a y-face: J = A·(A·dery + A′·derx).
__device__ __forceinline__ double flux_y(const lgd_kr_consts &k, double derx, double dery) {
  double a, da;
  lgd_kr_face(k, derx, dery, a, da);
  return a * ((a * dery) + (da * derx));
}

// Pointwise Model C update of one cell from its flux divergence
// (src/physics/phasefield.rs:395-421) and the stage update (src/core/driver.rs:131-138).
__device__ __forceinline__ void cell_update(const lgd_kr_consts &k, double flux, double pc,
                                            double pe, double pw, double pn, double ps, double uc,
                                            double ue, double uw, double un, double us,
                                            double dt, double &phi_new, double &u_new,
                                            double &dphi_out, double &du_out) {
  double g_prime = fma(pc * pc, pc, -pc);
  double one_m_p2 = fma(pc, -pc, 1.0);
  double p_prime = one_m_p2 * one_m_p2;
  double a = lgd_center_anisotropy(k, pe, pw, pn, ps);
  double dphi = fma(k.lambda * uc, -p_prime, flux - g_prime) / (a * a);
  double acc = 0.0;
  acc = acc + (((ue - (uc + uc)) + uw) * k.inv_hx2);
  acc = acc + (((un - (uc + uc)) + us) * k.inv_hy2);
  double du = fma(acc, k.d_thermal, 0.5 * dphi);
  phi_new = pc + dt * dphi;
  u_new = uc + dt * du;
  dphi_out = dphi;
  du_out = du;
}

// Read a table entry at a wave-uniform address into scalar registers (the
// compiler cannot prove the table is not written by the kernel, so a plain load
// would land in vector registers).
template <class T>
__device__ __forceinline__ T load_uniform(const T *p) {
  static_assert(sizeof(T) % 4 == 0, "word-sized table entries");
  T out;
  const int *src = reinterpret_cast<const int *>(p);
  int *dst = reinterpret_cast<int *>(&out);
#pragma unroll
  for (int i = 0; i < (int)(sizeof(T) / 4); ++i) dst[i] = __builtin_amdgcn_readfirstlane(src[i]);
  return out;
}

}  // namespace lgd_stream
