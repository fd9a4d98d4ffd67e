"""The phase-field tanh of the HIP kernels (bgk_pcm_amd/csrc/bgk_tanh.h),
compiled for the host, equals glibc's tanh bit for bit.

The reference calls glibc tanh (STEP_7/phase_field.f90:24-25, amdflang
emits a libm call).  The header reimplements the published fdlibm method
with glibc's evaluation of the expm1 correction polynomial; this test runs
it on the host over the phase field's argument range ((T-273)/T_width for
T from 0 to 600, T_width 10), across the reduction-branch boundaries of
expm1 and tanh, and over random doubles, and requires zero mismatches.
"""
import pathlib
import subprocess

ROOT = pathlib.Path(__file__).resolve().parents[1]

PROGRAM = r"""
#include <cmath>
#include <cstdio>
#include <cstdint>
#include <cstring>
#include "bgk_tanh.h"
static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t next() { s ^= s << 13; s ^= s >> 7; s ^= s << 17; return s; }
static long bad = 0, total = 0;
static void check(double x) {
  double a = bgk::tanh_fdlibm(x), b = std::tanh(x);
  uint64_t ua, ub; std::memcpy(&ua, &a, 8); std::memcpy(&ub, &b, 8);
  ++total;
  if (ua != ub) { if (bad < 5) std::printf("mismatch x=%a mine=%a glibc=%a\n", x, a, b); ++bad; }
}
int main() {
  // the phase field's arguments: ((T - 273) - 0)/T_width, T in [0, 600]
  for (long k = 0; k < 3000000; ++k) {
    double T = 600.0 * (double)(next() >> 11) / 9007199254740992.0;
    check(((T - 273.0) - 0.0) / 10.0);
  }
  // dense sweep around the branch points of tanh and expm1
  const double pts[] = {1.0, 0.5 * 0.6931471805599453, 1.5 * 0.6931471805599453, 22.0,
                        0.25, 0.5, 2.0, 10.0, 1e-8, 2.7755575615628914e-17};
  for (double c : pts)
    for (int d = -20000; d <= 20000; ++d) {
      check(std::nextafter(c, d < 0 ? 0.0 : 1e9) * (1.0 + d * 1e-12));
      check(-(c * (1.0 + d * 1e-9)));
    }
  // random bit patterns
  for (long k = 0; k < 3000000; ++k) {
    uint64_t u = next() & 0x7fffffffffffffffull; double x; std::memcpy(&x, &u, 8);
    if (!std::isfinite(x)) continue;
    check(std::fmod(x, 40.0) * ((k & 1) ? -1.0 : 1.0));
  }
  check(0.0); check(-0.0); check(INFINITY); check(-INFINITY); check(1e-320);
  std::printf("total=%ld bad=%ld\n", total, bad);
  return bad != 0;
}
"""


def test_device_tanh_equals_glibc_tanh(tmp_path):
    src = tmp_path / "t.cpp"
    src.write_text(PROGRAM)
    exe = tmp_path / "t"
    subprocess.run(["g++", "-O2", "-std=c++17", "-ffp-contract=off",
                    f"-I{ROOT / 'bgk_pcm_amd' / 'csrc'}", str(src), "-o", str(exe), "-lm"],
                   check=True)
    r = subprocess.run([str(exe)], capture_output=True, text=True)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "bad=0" in r.stdout
