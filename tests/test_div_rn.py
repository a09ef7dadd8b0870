"""The division-free FP64 relaxation quotient (div_rn in csrc/pmg_ops.h):
RN(a y) followed by two FMA corrections with the exact remainder returns the
IEEE quotient a / b when y = RN(1/b).  The same five operations are run on
the host (C99 fma is correctly rounded, like the GPU's v_fma_f64) for the
Jacobians of every dyadic grid size the solver uses, random divisors and
tens of millions of random dividends across many binades, and compared with
the division bit for bit."""
import os
import subprocess
import tempfile

import pytest

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
static uint64_t s = 0x9E3779B97F4A7C15ull;
static uint64_t next(void) {
  uint64_t z = (s += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double div_rn(double a, double b, double y) {
  double q0 = a * y;
  double r0 = fma(-b, q0, a);
  double q1 = fma(r0, y, q0);
  double r1 = fma(-b, q1, a);
  return fma(r1, y, q1);
}
static double rnd(int lo, int hi) {  /* random sign, significand and exponent in [lo, hi) */
  uint64_t u = next();
  double m = 1.0 + (double)(u >> 12) * 0x1p-52;
  int e = lo + (int)(next() % (uint64_t)(hi - lo));
  return ((u & 1) ? -1.0 : 1.0) * ldexp(m, e);
}
int main(void) {
  long bad = 0, n = 0;
  double bs[64];
  int nb = 0;
  for (int k = 1; k <= 12; k++) {        /* -6/h^2 (cube), and anisotropic sums, h = 2^-k */
    double h2 = ldexp(1.0, -2 * k);
    bs[nb++] = -2 / h2 - 2 / h2 - 2 / h2;
    bs[nb++] = -2 / h2 - 2 / (4 * h2) - 2 / h2;
  }
  for (int q = 0; q < 24; q++) bs[nb++] = rnd(-20, 40);
  for (int q = 0; q < nb; q++) {
    double b = bs[q], y = 1.0 / b;
    for (long t = 0; t < 1500000; t++) {
      double a = (t % 7 == 0) ? (double)(int64_t)(next() >> 40) : rnd(-60, 60);
      double x = a / b, z = div_rn(a, b, y);
      n++;
      if (memcmp(&x, &z, sizeof x) != 0) bad++;
    }
  }
  printf("%ld %ld\n", n, bad);
  return 0;
}
"""


def test_div_rn_equals_ieee_division():
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        exe = os.path.join(d, "t")
        with open(c, "w") as f:
            f.write(SRC)
        r = subprocess.run(["gcc", "-O2", "-std=c99", "-ffp-contract=off", "-mfma", c, "-o", exe, "-lm"],
                           capture_output=True, text=True)
        if r.returncode != 0:
            r = subprocess.run(["gcc", "-O2", "-std=c99", "-ffp-contract=off", c, "-o", exe, "-lm"],
                               capture_output=True, text=True)
        if r.returncode != 0:
            pytest.skip("no C compiler: " + r.stderr[-300:])
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    n, bad = int(out[0]), int(out[1])
    print(f"{n} quotients, {bad} differ")
    assert n > 50_000_000 and bad == 0
