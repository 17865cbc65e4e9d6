"""The device Box–Muller cosine and logarithm (legendre_amd/csrc/lgd_trig.h, host and
device share the code) against the host libm's cos and log, within one ulp:
- cos over the angles fl(2π·v) the noise generator forms: random 53-bit uniforms
  plus every uniform within 2^17 steps of the zero crossings and the ends of the
  range;
- log over the uniforms u = k·2^-53 it forms: random k, every k within 10^6 of 1
  and of 2^53, and every k within 2·10^4 of the reduction's boundaries (√2/2, √2)
  and of FDLIBM's switch between its two final forms, at every binade."""
from __future__ import annotations

import pathlib
import shutil
import subprocess

import pytest

HEADER = pathlib.Path(__file__).resolve().parents[1] / "legendre_amd" / "csrc" / "lgd_trig.h"

HARNESS = r"""
#include "%s"
#include <cstdio>
#include <random>
static double ulp_err(double got, double want) {
  if (got == want) return 0;
  double u = std::nextafter(std::fabs(want), INFINITY) - std::fabs(want);
  return std::fabs(got - want) / u;
}
int main() {
  std::mt19937_64 g(7);
  double worst = 0;
  auto check = [&](unsigned long long k) {
    double v = (double)((k >> 11) + 1) * (1.0 / 9007199254740992.0);
    double x = 6.283185307179586 * v;
    double e = ulp_err(lgd_cos_angle(x), std::cos(x));
    if (e > worst) worst = e;
  };
  for (long i = 0; i < 2000000; ++i) check(g());
  for (double c : {0.25, 0.75, 0.5, 1.0 - 1e-16, 0.0}) {
    unsigned long long k0 = (unsigned long long)(c * 9007199254740992.0) << 11;
    for (long d = -131072; d <= 131072; ++d) check(k0 + ((unsigned long long)d << 11));
  }
  printf("%%.6f\n", worst);
  return 0;
}
"""


LOG_HARNESS = r"""
#include "%s"
#include <cstdio>
#include <random>
static double ulp_err(double got, double want) {
  if (got == want) return 0;
  double u = std::nextafter(std::fabs(want), INFINITY) - std::fabs(want);
  return std::fabs(got - want) / u;
}
int main() {
  std::mt19937_64 g(7);
  const unsigned long long top = 1ull << 53;
  double worst = 0;
  auto check = [&](unsigned long long k) {
    if (k < 1 || k > top) return;
    double u = (double)k * (1.0 / 9007199254740992.0);
    double e = ulp_err(lgd_log_unit(u), std::log(u));
    if (e > worst) worst = e;
  };
  for (long i = 0; i < 4000000; ++i) check((g() >> 11) + 1);
  for (unsigned long long d = 1; d <= 1000000; ++d) { check(d); check(top + 1 - d); }
  for (int e = 0; e < 53; ++e)
    for (double c : {0.70710678118654752, 1.0, 1.3803, 1.4142135623730951, 1.4200, 1.999999}) {
      unsigned long long k0 = (unsigned long long)(std::ldexp(c, -e - 1) * 9007199254740992.0);
      for (long d = -20000; d <= 20000; ++d) check(k0 + d);
    }
  printf("%%.6f\n", worst);
  return 0;
}
"""


def run_harness(tmp_path, name: str, text: str) -> float:
    src = tmp_path / f"{name}.cpp"
    src.write_text(text % HEADER)
    exe = tmp_path / name
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", str(src), "-o", str(exe)], check=True)
    return float(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_box_muller_cosine_is_within_one_ulp_of_libm(tmp_path):
    worst = run_harness(tmp_path, "trig", HARNESS)
    assert worst <= 1.0, worst


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_box_muller_log_is_within_one_ulp_of_libm(tmp_path):
    worst = run_harness(tmp_path, "log", LOG_HARNESS)
    assert worst <= 1.0, worst
