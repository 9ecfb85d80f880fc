"""div_by (sheet_device.hpp) divides by a float d without a division
instruction: q = RN32(x * y) with y = RN64(1 / d), the product taken in
double.  The double product is within 2^-52 (relative) of x / d, and a
quotient of two floats never comes within about 2^-49 of a float rounding
boundary, so the one rounding to float gives the IEEE quotient for every x:
zeros, subnormals, overflow to infinity and NaNs included.  The Gauss-Seidel
sweep, the exact fused sheet kernel and the viscosity laws (mobility,
damping, overload) divide this way.  This checks the identity on the host's
IEEE arithmetic over every 64th float bit pattern (67 M values per divisor),
with no range excluded, for the divisors those callers use: 2 cell, cell and
cell^2 of the sheets' cell sizes, 12, 231 and 0.42."""
import os
import subprocess
import tempfile

SRC = r"""
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>
int main(void) {
  const float cells[] = {0.2f, 0.02f, 0.05f, 1.0f / 3.0f};
  float divisors[32];
  int n = 0;
  for (int k = 0; k < 4; k++) {
    divisors[n++] = 2.0f * cells[k];
    divisors[n++] = cells[k];
    divisors[n++] = cells[k] * cells[k];
  }
  divisors[n++] = 12.0f;
  divisors[n++] = 231.0f;
  divisors[n++] = 0.42f;
  unsigned long long bad = 0, checked = 0;
  for (int k = 0; k < n; k++) {
    const float d = divisors[k];
    const double y = 1.0 / (double)d;
    for (uint64_t u = 0; u < (1ull << 32); u += 64) {
      uint32_t w = (uint32_t)u ^ (uint32_t)(k * 2654435761u & 63u);
      float x;
      memcpy(&x, &w, 4);
      float q = (float)((double)x * y);
      float ref = x / d;
      checked++;
      if (isnan(ref) ? !isnan(q) : memcmp(&q, &ref, 4) != 0) bad++;
    }
  }
  printf("%llu %llu\n", checked, bad);
  return 0;
}
"""


def test_double_product_quotient_is_the_ieee_quotient():
    with tempfile.TemporaryDirectory() as tmp:
        src, exe = os.path.join(tmp, "d.c"), os.path.join(tmp, "d")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout.split()
    checked, bad = int(out[0]), int(out[1])
    assert checked > 15 * 67_000_000 and bad == 0
