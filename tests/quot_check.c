/* Checks the device's quotient routine (quot3 in jp-solver_amd/csrc/lattice.hpp) against IEEE binary32
 * division on the host: Y0 = an approximate reciprocal of b, two binary64 Newton steps
 * Y <- fma(Y, fma(-B, Y, 1), Y), then RN32(RN64(a Y)).
 *
 * The device's v_rcp_f32 is modelled by RN32(1/b) moved by -2..+2 ulps and by a relative error of
 * +-2^-14 (the routine's argument holds for any start within 2^-13). Cases: every float b in [0.9, 1.1]
 * (densities) with eight numerators each; random b in [0.5, 2) with numerators down to 2^-40; and
 * numerators chosen so that a / b falls next to a binary32 rounding midpoint (the cases a weak reciprocal
 * gets wrong).
 *
 * usage: quot_check <newton_steps>; prints "cases <n> mismatches <m>". */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static uint64_t st = 0x9e3779b97f4a7c15ull;
static uint64_t rnd(void) {
  st ^= st << 13;
  st ^= st >> 7;
  st ^= st << 17;
  return st;
}
static float f_of(uint32_t u) {
  float f;
  memcpy(&f, &u, 4);
  return f;
}
static uint32_t u_of(float f) {
  uint32_t u;
  memcpy(&u, &f, 4);
  return u;
}

static int newton = 2;
static long cases = 0, bad = 0;

static float quot(float a, float b, double y0) {
  const double B = (double)b;
  double Y = y0;
  for (int i = 0; i < newton; ++i) {
    const double E = fma(-B, Y, 1.0);
    Y = fma(Y, E, Y);
  }
  return (float)((double)a * Y);
}

static void check(float a, float b) {
  const float want = a / b;
  const float r = 1.0f / b;
  for (int k = -2; k <= 2; ++k) {
    const double y0s[2] = {(double)f_of(u_of(r) + (uint32_t)k), (1.0 + k * ldexp(1.0, -15)) / (double)b};
    for (int j = 0; j < 2; ++j) {
      const float got = quot(a, b, y0s[j]);
      ++cases;
      if (u_of(got) != u_of(want)) {
        if (bad < 5) fprintf(stderr, "mismatch a=%a b=%a y0=%a got=%a want=%a\n", a, b, y0s[j], got, want);
        ++bad;
      }
    }
  }
}

int main(int argc, char** argv) {
  if (argc > 1) newton = atoi(argv[1]);
  /* every density in [0.9, 1.1], eight numerators each */
  for (uint32_t u = u_of(0.9f); u <= u_of(1.1f); u += 1) {
    const float b = f_of(u);
    for (int i = 0; i < 8; ++i) {
      const float a = (float)(((double)(rnd() >> 11) * 0x1p-53 - 0.5) * ldexp(1.0, -(int)(rnd() % 20)));
      check(a, b);
    }
  }
  /* random divisors in [0.5, 2), numerators down to 2^-40, both signs, and zeros */
  for (int i = 0; i < 2000000; ++i) {
    const float b = (float)ldexp(1.0 + (double)(rnd() >> 40) * 0x1p-24, (int)(rnd() % 2) - 1);
    float a = (float)ldexp(1.0 + (double)(rnd() >> 40) * 0x1p-24, -(int)(rnd() % 41));
    if (rnd() & 1) a = -a;
    if (i % 100000 == 0) a = (i & 1) ? 0.0f : -0.0f;
    check(a, b);
  }
  /* quotients next to a rounding midpoint: for an odd divisor significand B and M = +-1/B mod 2^25 taken in
     [2^24, 2^25) (an odd 25-bit significand, i.e. a binary32 midpoint), B M = j 2^25 -+ 1 with j < 2^24, so
     a = j 2^-48 and b = B 2^-23 give a / b = (M -+ 1/B) 2^-50, within 2^-49 relative of the midpoint M 2^-50:
     the closest a float quotient comes to one */
  for (int i = 0; i < 4000000; ++i) {
    const uint64_t B = (1ull << 23) | (rnd() >> 41) | 1ull;
    uint64_t x = B;
    for (int it = 0; it < 5; ++it) x *= 2 - B * x; /* inverse of B mod 2^64 */
    uint64_t M = x & ((1ull << 25) - 1);
    if (M < (1ull << 24)) M = (1ull << 25) - M;  /* -1/B: B M = -1 mod 2^25 */
    const uint64_t P = B * M;
    const uint64_t j = (P + ((P & 1ull) ? 1ull : 0ull)) >> 25; /* P = j 2^25 -+ 1 */
    if (j == 0 || j >= (1ull << 24)) continue;
    const float b = (float)ldexp((double)B, -23), a = (float)ldexp((double)j, -48);
    check(a, b);
    check(-a, b);
  }
  printf("cases %ld mismatches %ld\n", cases, bad);
  return 0;
}
