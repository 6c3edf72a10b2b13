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
 *
 * Mode "bmr" checks the form the step kernels use for the constant divisors
 * y = 8h (CDiv in kernels.hip, const_div in capi.cpp):
 *
 *   zh = RN(1/y) (RD(1/y) where RN(1/y) > 1/y);  zl = RN(1/y - zh) >= 0;
 *   q = fma(x, zh, RN(x*zl))
 *
 * For each divisor it first repeats the host's admission test (q equals x/y
 * for every significand x in [1, 2)) and the host's threshold T (the
 * smallest numerator magnitude the fast form takes), then compares q with
 * x/y on whole binades of x from T up to 2^100 (both signs, every STRIDE-th
 * significand), and on signed zeros.  Mode "bmr_below" counts mismatches for
 * numerators a few binades below T, where x*zl leaves the normal range: a
 * non-zero count shows the threshold is needed.
 *
 *   division_check bmr STRIDE          division_check bmr_below
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

static float bmr(float x, float zh, float zl) { return fmaf(x, zh, x * zl); }

/* the host's admission test and threshold for divisor y (const_div) */
static int bmr_admit(float y, float* zh, float* zl, float* thr) {
  *zh = 1.0f / y;
  if ((double)*zh > 1.0 / (double)y) *zh = nextafterf(*zh, 0.0f);
  *zl = (float)(1.0 / (double)y - (double)*zh);
  int ok = y >= 0x1p-20f && y <= 0x1p20f && *zl >= 0.0f;
  for (uint32_t m = 0; ok && m < (1u << 23); ++m) {
    const float x = f_of(0x3f800000u | m);
    if (bmr(x, *zh, *zl) != x / y) ok = 0;
  }
  double t = 0x1p-100;
  if (*zl != 0.0f && 0x1p-125 / fabs((double)*zl) > t) t = 0x1p-125 / fabs((double)*zl);
  if (0x1p-125 * (double)y > t) t = 0x1p-125 * (double)y;
  *thr = (float)ldexp(1.0, (int)ceil(log2(t)));
  return ok;
}

static int bmr_main(int below, uint32_t stride) {
  const float fixed[] = {80.0f, 8.0f * 1.7f, 8.0f * 2.3f, 8.0f * 12.5f, 40.0f, 8.0f, 1.0f, 3.0f, 0.5f,
                         1.2345678f, 0x1p-20f, 0x1p20f};
  const int nfixed = (int)(sizeof fixed / sizeof fixed[0]);
  long bad = 0, tot = 0, admitted = 0;
  for (int t = 0; t < nfixed + 6; ++t) {
    const float y = t < nfixed ? fixed[t] : 8.0f * (float)(0.25 + (double)(rnd() % 100000) / 1000.0);
    float zh, zl, thr;
    if (!bmr_admit(y, &zh, &zl, &thr)) {
      printf("divisor %a not admitted\n", y);
      if (t < nfixed) ++bad;   /* every fixed divisor is expected to be admitted */
      continue;
    }
    ++admitted;
    const int lo = ilogbf(thr);
    if (below) {
      /* numerators 1..12 binades below the threshold (nonzero, normal or not) */
      for (int b = lo - 12; b < lo; ++b)
        for (uint32_t m = 0; m < (1u << 23); m += 97) {
          const float x = b >= -126 ? f_of(((uint32_t)(127 + b) << 23) | m) : ldexpf(1.0f + (float)m * 0x1p-23f, b);
          if (x == 0.0f) continue;
          ++tot;
          if (u_of(bmr(x, zh, zl)) != u_of(x / y)) ++bad;
        }
      continue;
    }
    const int binades[] = {lo, lo + 1, lo + 2, -37, 0, 31, 98, 99};
    for (int b = 0; b < 8; ++b)
      for (uint32_t m = (uint32_t)(rnd() % stride); m < (1u << 23); m += stride)
        for (int sg = 0; sg < 2; ++sg) {
          const float x = f_of(((uint32_t)(127 + binades[b]) << 23) | m | (sg ? 0x80000000u : 0u));
          const float q = bmr(x, zh, zl), d = x / y;
          ++tot;
          if (u_of(q) != u_of(d)) {
            if (bad < 5) printf("mismatch y=%a x=%a q=%a ieee=%a\n", y, x, q, d);
            ++bad;
          }
        }
    const float xs[] = {0x1p100f, -0x1p100f, 0.0f, -0.0f, thr, -thr};
    for (int k = 0; k < 6; ++k) {
      ++tot;
      if (u_of(bmr(xs[k], zh, zl)) != u_of(xs[k] / y)) { printf("mismatch y=%a x=%a\n", y, xs[k]); ++bad; }
    }
  }
  printf("%s: %ld divisors admitted, checked %ld, mismatches %ld\n", below ? "bmr_below" : "bmr", admitted, tot, bad);
  return bad > 100 ? 100 : (int)bad;
}

int main(int argc, char** argv) {
  if (argc > 1 && strcmp(argv[1], "bmr") == 0) return bmr_main(0, argc > 2 ? (uint32_t)strtoul(argv[2], 0, 10) : 1);
  if (argc > 1 && strcmp(argv[1], "bmr_below") == 0) return bmr_main(1, 1);
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
