"""The Gauss-Seidel sweep (sheet_exact.hip, k_force_gauss_seidel_lds)
divides by 2 cell without a division instruction: q = x * y with
y = RN(1 / d), r = fma(-d, q, x), q' = fma(r, y, q).  By Markstein's theorem
q' is the correctly rounded x / d whenever nothing underflows or overflows;
the sweep keeps the division for |q| outside [2^-100, 2^100] and for x = 0.
This checks the identity on the host's IEEE arithmetic for the sheets' cell
sizes over every 64th float bit pattern (67 M values per divisor)."""
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
  unsigned long long bad = 0, checked = 0;
  for (int k = 0; k < 4; k++) {
    const float d = 2.0f * cells[k];
    const float y = 1.0f / d;
    for (uint64_t u = 0; u < (1ull << 32); u += 64) {
      uint32_t w = (uint32_t)u ^ (uint32_t)(k * 2654435761u & 63u);
      float x;
      memcpy(&x, &w, 4);
      if (!isfinite(x)) continue;
      float q = x * y;
      float aq = fabsf(q);
      if (!(aq >= 0x1p-100f && aq <= 0x1p100f)) continue;
      float q2 = fmaf(fmaf(-d, q, x), y, q);
      float ref = x / d;
      checked++;
      if (memcmp(&q2, &ref, 4) != 0) bad++;
    }
  }
  printf("%llu %llu\n", checked, bad);
  return 0;
}
"""


def test_fma_corrected_quotient_is_the_ieee_quotient():
    with tempfile.TemporaryDirectory() as tmp:
        src, exe = os.path.join(tmp, "d.c"), os.path.join(tmp, "d")
        with open(src, "w") as f:
            f.write(SRC)
        subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, src, "-lm"], check=True)
        out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout.split()
    checked, bad = int(out[0]), int(out[1])
    assert checked > 200_000_000 and bad == 0
