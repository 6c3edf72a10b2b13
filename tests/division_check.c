/* CPU check of the exact-division form used by the step kernels
 * (seiswg_amd/csrc/kernels.hip, div_exact):
 *
 *   r = RN(1/y);  q0 = RN(x*r);  e = fma(q0, y, -x);  q = fma(-e, r, q0)
 *
 * Claim: for y in [2^-20, 2^20] and x zero or 2^-100 <= |x| <= 2^100, q is
 * the IEEE single-precision quotient x/y bit for bit (signed zeros
 * included).  Mode "guarded" checks the claim on whole binades of x (every
 * significand, both signs) at both ends of the range and inside it, for fixed
 * and random divisors; exit status = number of mismatches (capped).  Mode
 * "unguarded" counts mismatches of the same form for quotients in the
 * subnormal range, which the kernels route to the IEEE division instead: a
 * non-zero count shows that the range test is needed and that this checker
 * can see a one-unit difference.
 *
 *   division_check guarded STRIDE      division_check unguarded
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float mk(float x, float y, float r) {
  const float q0 = x * r;
  const float e = fmaf(q0, y, -x);
  return fmaf(-e, r, q0);
}
static float f_of(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t u_of(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }
static uint64_t rs = 88172645463325252ull;
static uint64_t rnd(void) { rs ^= rs << 13; rs ^= rs >> 7; rs ^= rs << 17; return rs; }

int main(int argc, char** argv) {
  const int guarded = argc > 1 && strcmp(argv[1], "guarded") == 0;
  const uint32_t stride = argc > 2 ? (uint32_t)strtoul(argv[2], 0, 10) : 1;
  const float fixed[] = {80.0f, 8.0f * 1.7f, 8.0f * 2.3f, 2700.0f, 1.0f, 3.0f, 0.5f, 1.2345678f,
                         0x1p-20f, 0x1.fffffep19f, 0x1p20f, 0x1.000002p-20f};
  const int nfixed = (int)(sizeof fixed / sizeof fixed[0]);
  long bad = 0, tot = 0;
  volatile float one = 1.0f;
  if (guarded) {
    const int binades[] = {-100, -99, -37, 0, 31, 99};
    for (int t = 0; t < nfixed + 8; ++t) {
      float y = t < nfixed ? fixed[t] : f_of((uint32_t)(rnd() & 0x007FFFFFu) | ((uint32_t)(107 + rnd() % 40) << 23));
      const float r = one / y;
      for (int b = 0; b < 6; ++b) {
        for (uint32_t m = (uint32_t)(rnd() % stride); m < (1u << 23); m += stride) {
          for (int sg = 0; sg < 2; ++sg) {
            const float x = f_of(((uint32_t)(127 + binades[b]) << 23) | m | (sg ? 0x80000000u : 0u));
            const float q = mk(x, y, r), d = x / y;
            ++tot;
            if (u_of(q) != u_of(d)) {
              if (bad < 5) printf("mismatch y=%a x=%a q=%a ieee=%a\n", y, x, q, d);
              ++bad;
            }
          }
        }
      }
      /* the largest magnitude in range and signed zeros */
      const float xs[] = {0x1p100f, -0x1p100f, 0.0f, -0.0f};
      for (int k = 0; k < 4; ++k) {
        const float q = mk(xs[k], y, r), d = xs[k] / y;
        ++tot;
        if (u_of(q) != u_of(d)) { printf("mismatch y=%a x=%a\n", y, xs[k]); ++bad; }
      }
    }
  } else {
    for (int t = 0; t < nfixed; ++t) {
      const float y = fixed[t], r = one / y;
      for (uint32_t u = 0; u < (1u << 22); ++u) {
        /* numerators whose quotient by y lies in the subnormal range */
        const int ex = ilogbf(y) - 140 + (int)(u % 14);
        if (ex < -126) continue;
        const float x = f_of(((uint32_t)(127 + ex) << 23) | (uint32_t)(rnd() & 0x007FFFFFu));
        if (!(fabsf(x / y) < 0x1p-126f)) continue;
        ++tot;
        if (u_of(mk(x, y, r)) != u_of(x / y)) ++bad;
      }
    }
  }
  printf("%s: checked %ld, mismatches %ld\n", guarded ? "guarded" : "unguarded", tot, bad);
  return bad > 100 ? 100 : (int)bad;
}
