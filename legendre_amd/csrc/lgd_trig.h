// cos(x) for the Box–Muller angle x = fl(2π·v), v in (0, 1], so 0 < x <= 2π.
//
// Method (no large-argument reduction is needed in this range):
// - Cody & Waite's additive reduction by π/2 with a three-part constant: with
//   n = rint(2x/π) <= 4, x − n·P1 is exact (both are multiples of x's ulp or finer and
//   the difference is below one), and the P2, P3 terms follow by fused multiply-adds;
// - the minimax polynomials of FDLIBM's __kernel_cos and __kernel_sin (Sun
//   Microsystems, 1993, as in FreeBSD's msun k_cos.c and k_sin.c) on [−π/4, π/4],
//   evaluated by Horner's rule with fused multiply-adds;
// - the quadrant n mod 4 picks cos r, −sin r, −cos r or sin r.
//
// The result is within about one ulp of the exact cosine. It is tested against the
// host libm's cos over the whole range of angles, including every zero crossing.
// Host and device compile the same code.
#pragma once

#include <cmath>
#include <cstdint>

#if defined(__HIPCC__)
#define LGD_TRIG_HD __host__ __device__
#else
#define LGD_TRIG_HD
#endif

LGD_TRIG_HD inline double lgd_bits_to_double(uint64_t b) {
  double d;
  __builtin_memcpy(&d, &b, 8);
  return d;
}

LGD_TRIG_HD inline double lgd_cos_angle(double x) {
  const double two_over_pi = lgd_bits_to_double(0x3fe45f306dc9c883ull);
  const double p1 = lgd_bits_to_double(0x3ff921fb54442d18ull);  // fl(π/2)
  const double p2 = lgd_bits_to_double(0x3c91a62633145c07ull);  // fl(π/2 − p1)
  const double p3 = lgd_bits_to_double(0xb91f1976b7ed8fbcull);  // fl(π/2 − p1 − p2)
  const double n = rint(x * two_over_pi);
  double r = fma(-n, p1, x);
  r = fma(-n, p2, r);
  r = fma(-n, p3, r);
  const double z = r * r;
  // __kernel_cos: 1 − z/2 + z·R(z), with the 1 − z/2 split to keep its rounding.
  const double c1 = lgd_bits_to_double(0x3fa555555555554cull);
  const double c2 = lgd_bits_to_double(0xbf56c16c16c15177ull);
  const double c3 = lgd_bits_to_double(0x3efa01a019cb1590ull);
  const double c4 = lgd_bits_to_double(0xbe927e4f809c52adull);
  const double c5 = lgd_bits_to_double(0x3e21ee9ebdb4b1c4ull);
  const double c6 = lgd_bits_to_double(0xbda8fae9be8838d4ull);
  const double rc = z * fma(z, fma(z, fma(z, fma(z, fma(z, c6, c5), c4), c3), c2), c1);
  const double hz = 0.5 * z;
  const double w = 1.0 - hz;
  const double cosr = w + (((1.0 - w) - hz) + z * rc);
  // __kernel_sin: r + r³·S(z).
  const double s1 = lgd_bits_to_double(0xbfc5555555555549ull);
  const double s2 = lgd_bits_to_double(0x3f8111111110f8a6ull);
  const double s3 = lgd_bits_to_double(0xbf2a01a019c161d5ull);
  const double s4 = lgd_bits_to_double(0x3ec71de357b1fe7dull);
  const double s5 = lgd_bits_to_double(0xbe5ae5e68a2b9cebull);
  const double s6 = lgd_bits_to_double(0x3de5d93a5acfd57cull);
  const double rs = fma(z, fma(z, fma(z, fma(z, s6, s5), s4), s3), s2);
  const double sinr = fma(z * r, fma(z, rs, s1), r);
  const int q = (int)n & 3;
  const double v = (q & 1) ? sinr : cosr;
  return (q == 1 || q == 2) ? -v : v;
}
