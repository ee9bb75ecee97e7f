/* Exhaustive check of the constant division used by the flux5 / flux3 face
 * values: for every non-negative float x (the form is odd in x, as is RN(x/d),
 * so the negative half mirrors it) and the divisor d given on the command
 * line, the product form q = RN(x r), e = fma(-q, d, x), y = copysign(fma(e, r,
 * q), x) with r = RN(1/d) must equal RN(x / d) bit for bit unless x is
 * subnormal or infinite or y is subnormal or NaN (those cases take a true
 * division).  Prints the number of mismatches outside those cases; exits 1 on
 * any. */
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static float D, R;
enum { NT = 8 };
typedef struct { uint32_t lo, hi; uint64_t bad, fast; } Part;

static int slow_case(float x, float y) {
  int cx = fpclassify(x), cy = fpclassify(y);
  return cx == FP_SUBNORMAL || cx == FP_INFINITE || cy == FP_SUBNORMAL || cy == FP_NAN;
}

static void* run(void* p) {
  Part* a = (Part*)p;
  for (uint64_t u = a->lo; u < a->hi; ++u) {
    uint32_t b = (uint32_t)u, rb, yb;
    float x;
    memcpy(&x, &b, 4);
    if (isnan(x)) continue;
    volatile float ref = x / D;
    float q = x * R;
    float e = fmaf(-q, D, x);
    float y = copysignf(fmaf(e, R, q), x);
    if (slow_case(x, y)) continue;
    a->fast++;
    float rr = ref;
    memcpy(&rb, &rr, 4);
    memcpy(&yb, &y, 4);
    if (rb != yb) a->bad++;
  }
  return 0;
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  D = (float)atof(argv[1]);
  R = 1.0f / D;
  Part P[NT];
  pthread_t t[NT];
  const uint64_t top = 0x80000000ull;
  for (int i = 0; i < NT; ++i) {
    P[i].lo = (uint32_t)(top * i / NT);
    P[i].hi = (uint32_t)(top * (i + 1) / NT);
    P[i].bad = P[i].fast = 0;
    pthread_create(&t[i], 0, run, &P[i]);
  }
  uint64_t bad = 0, fast = 0;
  for (int i = 0; i < NT; ++i) {
    pthread_join(t[i], 0);
    bad += P[i].bad;
    fast += P[i].fast;
  }
  printf("d=%g r=%a fast-path inputs %llu mismatches %llu\n", D, R, (unsigned long long)fast,
         (unsigned long long)bad);
  return bad ? 1 : 0;
}
