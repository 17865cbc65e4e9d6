// The two transcendental functions of the Box–Muller draw, written for the
// arguments the draw forms (host and device compile the same code):
// - cos(x) for the angle x = fl(2π·v), v in (0, 1], so 0 < x <= 2π;
// - log(u) for the uniform u = k·2^-53, k in [1, 2^53], so 2^-53 <= u <= 1.
//
// cos:
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
//
// log: FDLIBM's __ieee754_log (Sun Microsystems, 1993, as in FreeBSD's msun
// e_log.c), specialised to positive normal arguments:
// - x = 2^k·(1 + f) with √2/2 <= 1 + f < √2, taken from x's bits;
// - s = f/(2 + f), so that log(1 + f) = 2s + s·R(s²), with FDLIBM's degree-7
//   minimax R (Horner's rule with fused multiply-adds);
// - log x = k·ln2_hi + (f − (hf² − s·(hf² + R) + k·ln2_lo)) for f near the ends of
//   the reduced range (hf² = f²/2), else k·ln2_hi + (f − (s·(f − R) − k·ln2_lo)),
//   with ln2_hi's trailing zeros making k·ln2_hi exact.
// FDLIBM's shortcuts for k = 0 and |f| < 2^-20 give the same value or one within
// its error bound, and are left out. Zero, negative, subnormal and non-finite
// arguments are not handled: the draw never forms them. The result is within
// one ulp of the exact logarithm; it is tested against the host libm's log over
// random uniforms and every uniform near 1, near 2^-53 and near the reduction's
// boundaries.
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

LGD_TRIG_HD inline double lgd_log_unit(double x) {
  uint64_t bits;
  __builtin_memcpy(&bits, &x, 8);
  const int32_t hi = (int32_t)(bits >> 32);
  const int32_t hx = hi & 0x000fffff;             // top 20 mantissa bits
  const int32_t half = (hx + 0x95f64) & 0x100000;  // set when 1.m >= √2: halve
  const int k = ((hi >> 20) - 1023) + (half >> 20);
  const uint64_t mbits = ((uint64_t)(uint32_t)(hx | (half ^ 0x3ff00000)) << 32) |
                         (bits & 0xffffffffull);
  const double f = lgd_bits_to_double(mbits) - 1.0;
  const double lg1 = lgd_bits_to_double(0x3fe5555555555593ull);
  const double lg2 = lgd_bits_to_double(0x3fd999999997fa04ull);
  const double lg3 = lgd_bits_to_double(0x3fd2492494229359ull);
  const double lg4 = lgd_bits_to_double(0x3fcc71c51d8e78afull);
  const double lg5 = lgd_bits_to_double(0x3fc7466496cb03deull);
  const double lg6 = lgd_bits_to_double(0x3fc39a09d078c69full);
  const double lg7 = lgd_bits_to_double(0x3fc2f112df3e5244ull);
  const double ln2_hi = lgd_bits_to_double(0x3fe62e42fee00000ull);
  const double ln2_lo = lgd_bits_to_double(0x3dea39ef35793c76ull);
  const double s = f / (2.0 + f);
  const double dk = (double)k;
  const double z = s * s;
  const double w = z * z;
  const double t1 = w * fma(w, fma(w, lg6, lg4), lg2);
  const double t2 = z * fma(w, fma(w, fma(w, lg7, lg5), lg3), lg1);
  const double r = t2 + t1;
  // FDLIBM's two final forms on one evaluation path. Near the ends of the reduced
  // range: q = p = hf², giving hf² − (s·(hf² + R) + k·ln2_lo). Elsewhere q = 0,
  // p = −f, giving −(s·(R − f) + k·ln2_lo), which is s·(f − R) − k·ln2_lo exactly
  // (rounding is symmetric under negation; a zero's sign is lost in the − f).
  const bool ends = ((hx - 0x6147a) | (0x6b851 - hx)) > 0;
  const double hfsq = 0.5 * f * f;
  const double q = ends ? hfsq : 0.0;
  const double p = ends ? hfsq : -f;
  return dk * ln2_hi - ((q - (s * (p + r) + dk * ln2_lo)) - f);
}
