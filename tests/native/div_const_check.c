/* Checks the constant-divisor quotient of the sh2peaks kernel (div_const in
 * mrtrix3_amd/csrc/kernels/peaks.hip): q = fma(-r, y, q0) with y = 1/f, q0 = a*y and
 * r = fma(f, q0, -a), against the correctly rounded a / f, for every divisor f of the associated
 * Legendre recursion up to lmax 12 (core/math/legendre.h:129-154 in the reference).
 *
 *   div_const_check            every significand at exponents -100, -60, -20, 0, 20, 60, 100,
 *                              and +-0 (what the unit test runs: a few seconds)
 *   div_const_check all        every float bit pattern (2^32 per divisor; minutes on 8 threads)
 *   div_const_check double     the double-precision form (mrtrix3_amd/csrc/kernels/resp.hip) for the
 *                              m = 0 divisors up to lmax 12: 4M random significands per divisor at
 *                              exponents from -300 to 300, and +-0
 *
 * Prints one line per divisor and "FAIL" lines for mismatches with finite |a| >= 2^-100; exits 1
 * on any such mismatch.  Build with -ffp-contract=off so that only the explicit fmaf calls fuse. */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <string.h>

static float from_bits(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static uint32_t to_bits(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

typedef struct {
  float f;
  int all;
  uint64_t checked, bad, bad_tiny;
} job;

static int check_one(float a, float f, float y) {
  const float ref = a / f;
  const float q0 = a * y;
  const float r = fmaf(f, q0, -a);
  const float q = fmaf(-r, y, q0);
  return to_bits(q) == to_bits(ref) || (isnan(q) && isnan(ref));
}

static void* run(void* p) {
  job* j = (job*)p;
  const float f = j->f, y = 1.0f / f;
  const float tiny = ldexpf(1.0f, -100);
  if (j->all) {
    for (uint64_t u = 0; u < (1ull << 32); ++u) {
      const float a = from_bits((uint32_t)u);
      if (!isfinite(a)) continue;
      ++j->checked;
      if (!check_one(a, f, y)) {
        if (fabsf(a) >= tiny) ++j->bad; else ++j->bad_tiny;
      }
    }
  } else {
    static const int ex[] = {-100, -60, -20, 0, 20, 60, 100};
    for (int e = 0; e < 7; ++e)
      for (uint32_t m = 0; m < (1u << 23); ++m)
        for (int sign = 0; sign < 2; ++sign) {
          const uint32_t bits = ((uint32_t)sign << 31) | ((uint32_t)(ex[e] + 127) << 23) | m;
          ++j->checked;
          if (!check_one(from_bits(bits), f, y)) ++j->bad;
        }
    for (int sign = 0; sign < 2; ++sign) {
      ++j->checked;
      if (!check_one(sign ? -0.0f : 0.0f, f, y)) ++j->bad;
    }
  }
  return 0;
}

static int check_double(void) {
  int failed = 0;
  uint64_t state = 0x9E3779B97F4A7C15ull;
  for (int l = 2; l <= 12; ++l) {
    /* the divisor at degree l is the factor of degree l - 1 (legendre.h:118-121 with m = 0) */
    const double f = l == 2 ? sqrt(3.0) : sqrt((double)(4 * (l - 1) * (l - 1) - 1) / (double)((l - 1) * (l - 1)));
    const double y = 1.0 / f;
    uint64_t bad = 0, checked = 0;
    for (int i = 0; i < 4000000 + 2; ++i) {
      double a;
      if (i < 2) {
        a = i ? -0.0 : 0.0;
      } else {
        state ^= state << 13; state ^= state >> 7; state ^= state << 17;
        const uint64_t mant = state & ((1ull << 52) - 1);
        const int ex = (int)((state >> 52) % 601) - 300;
        const uint64_t bits = ((uint64_t)(state >> 63) << 63) | ((uint64_t)(ex + 1023) << 52) | mant;
        memcpy(&a, &bits, 8);
      }
      const double ref = a / f;
      const double q0 = a * y;
      const double r = fma(f, q0, -a);
      const double q = fma(-r, y, q0);
      ++checked;
      if (memcmp(&q, &ref, 8) != 0) ++bad;
    }
    printf("%s double f=%.17g checked %llu mismatches %llu\n", bad ? "FAIL" : "ok", f, (unsigned long long)checked,
           (unsigned long long)bad);
    failed |= bad != 0;
  }
  printf("11 double divisors\n");
  return failed;
}

int main(int argc, char** argv) {
  if (argc > 1 && strcmp(argv[1], "double") == 0) return check_double();
  const int all = argc > 1 && strcmp(argv[1], "all") == 0;
  job jobs[128];
  int nj = 0;
  for (int M = 0; M + 2 <= 12; ++M) {
    float f = sqrtf((float)(2 * M + 3));   /* the divisor at n = M + 2 */
    for (int n = M + 2; n <= 12; ++n) {
      jobs[nj].f = f;
      jobs[nj].all = all;
      jobs[nj].checked = jobs[nj].bad = jobs[nj].bad_tiny = 0;
      ++nj;
      f = sqrtf((float)(4 * n * n - 1) / (float)(n * n - M * M));
    }
  }
  int failed = 0;
  for (int b = 0; b < nj; b += 8) {
    pthread_t t[8];
    int k = 0;
    for (; k < 8 && b + k < nj; ++k) pthread_create(&t[k], 0, run, &jobs[b + k]);
    for (int i = 0; i < k; ++i) pthread_join(t[i], 0);
    for (int i = 0; i < k; ++i) {
      const job* j = &jobs[b + i];
      printf("%s f=%.9g (0x%08x) checked %llu mismatches %llu (|a| < 2^-100: %llu)\n", j->bad ? "FAIL" : "ok",
             j->f, to_bits(j->f), (unsigned long long)j->checked, (unsigned long long)j->bad,
             (unsigned long long)j->bad_tiny);
      failed |= j->bad != 0;
    }
  }
  printf("%d divisors\n", nj);
  return failed;
}
