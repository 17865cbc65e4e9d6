"""The device Box–Muller cosine (legendre_amd/csrc/lgd_trig.h, host and device share
the code) against the host libm's cos over the angles fl(2π·v) the noise generator
forms: random 53-bit uniforms plus every uniform within 2^17 steps of the zero
crossings and the ends of the range. Within one ulp."""
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


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs a host C++ compiler")
def test_box_muller_cosine_is_within_one_ulp_of_libm(tmp_path):
    src = tmp_path / "trig.cpp"
    src.write_text(HARNESS % HEADER)
    exe = tmp_path / "trig"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off", str(src), "-o", str(exe)], check=True)
    worst = float(subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout)
    assert worst <= 1.0, worst
