/* Checks of the shared-reciprocal division used by the GPU element kernel:
 *   y = RN(1/b), q0 = RN(a*y), e = RN(b*q0 - a), q1 = RN(q0 - e*y)
 * against the correctly rounded quotient RN(a/b).
 *
 *   division_check exhaustive P   all pairs of P-bit significands a, b in [1,2),
 *                                 in exact integer arithmetic (P <= 20)
 *   division_check binary64 N S   N binary64 pairs (seed S), half of them drawn
 *                                 near the hard cases (b close to 2, a close to b),
 *                                 with the C library's fma and IEEE division
 * Prints the number of wrong quotients; exits 1 if any.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

typedef __int128 i128;
typedef struct { i128 m; int e; } val; /* m * 2^e */

static int bitlen(i128 x) { int n = 0; if (x < 0) x = -x; while (x) { n++; x >>= 1; } return n; }

static val round_to(val v, int p) {
  if (v.m == 0) return v;
  int neg = v.m < 0;
  i128 m = neg ? -v.m : v.m;
  int L = bitlen(m);
  if (L <= p) return v;
  int sh = L - p;
  i128 q = m >> sh, rem = m - (q << sh), half = (i128)1 << (sh - 1);
  if (rem > half || (rem == half && (q & 1))) q++;
  val r = {neg ? -q : q, v.e + sh};
  return r;
}
static val mul(val a, val b) { val r = {a.m * b.m, a.e + b.e}; return r; }
static val sub(val a, val b) {
  if (b.m == 0) return a;
  if (a.m == 0) { b.m = -b.m; return b; }
  int e = a.e < b.e ? a.e : b.e;
  val r = {(a.m << (a.e - e)) - (b.m << (b.e - e)), e};
  return r;
}
/* RN_p(N/D), N, D > 0 */
static val div_round(i128 N, i128 D, int p) {
  int k = 0;
  while ((N << k) / D < ((i128)1 << (p + 1))) k++;
  i128 q = (N << k) / D, rem = (N << k) - q * D;
  int sh = bitlen(q) - p;
  i128 hi = q >> sh, low = q - (hi << sh), half = (i128)1 << (sh - 1);
  if (low > half || (low == half && (rem != 0 || (hi & 1)))) hi++;
  val r = {hi, sh - k};
  return r;
}

static long long exhaustive(int p) {
  long long wrong = 0;
  const long long lo = 1LL << (p - 1), hi = 1LL << p;
  for (long long A = lo; A < hi; A++) {
    val a = {A, 1 - p};
    for (long long B = lo; B < hi; B++) {
      val b = {B, 1 - p};
      val y = div_round((i128)1 << (p - 1), B, p); /* 1/b */
      val q0 = round_to(mul(a, y), p);
      val e = round_to(sub(mul(b, q0), a), p);
      val q1 = round_to(sub(q0, mul(e, y)), p);
      val Q = div_round(A, B, p);
      if (sub(q1, Q).m != 0) wrong++;
    }
  }
  return wrong;
}

static uint64_t rng_state;
static uint64_t next_u64(void) { /* splitmix64 */
  uint64_t z = (rng_state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
static double from_bits(uint64_t u) { double d; memcpy(&d, &u, 8); return d; }

static long long binary64(long long n) {
  long long wrong = 0;
  for (long long i = 0; i < n; i++) {
    uint64_t r = next_u64(), s = next_u64();
    double a, b;
    if (i & 1) {
      /* significands near 2 for b and near b for a */
      b = from_bits(0x3FFFFFFFFFFFFFFFull - (r & 0xFFFFFFull));
      a = from_bits(0x3FF0000000000000ull | ((s & 0x1) ? (r & 0xFFFFFFFFFFFFFull) : ((r & 0xFFFFFFFFFFFFFull) ^ (s & 0xFFFFFull))));
      if (s & 2) a = b * (1.0 + ldexp((double)(s >> 40), -60));
    } else {
      b = from_bits(0x3FF0000000000000ull | (r & 0xFFFFFFFFFFFFFull));
      a = from_bits(0x3FF0000000000000ull | (s & 0xFFFFFFFFFFFFFull));
    }
    /* scale into the range the kernel accepts: d in [2^-60, 2^40], |q0| in [2^-901, 2^900) */
    b = ldexp(b, (int)((r >> 52) % 100) - 60);
    a = ldexp(a, (int)((s >> 52) % 1600) - 800);
    if (s & (1ull << 63)) a = -a;
    const double y = 1.0 / b;
    const double q0 = a * y;
    const double e = fma(b, q0, -a);
    const double q1 = fma(-e, y, q0);
    const double q = a / b;
    if (memcmp(&q1, &q, 8) != 0) wrong++;
  }
  return wrong;
}

int main(int argc, char** argv) {
  if (argc >= 3 && !strcmp(argv[1], "exhaustive")) {
    const int p = atoi(argv[2]);
    if (p < 3 || p > 20) return 2;
    const long long w = exhaustive(p);
    printf("exhaustive p=%d wrong=%lld\n", p, w);
    return w != 0;
  }
  if (argc >= 4 && !strcmp(argv[1], "binary64")) {
    rng_state = strtoull(argv[3], NULL, 10);
    const long long w = binary64(atoll(argv[2]));
    printf("binary64 n=%s wrong=%lld\n", argv[2], w);
    return w != 0;
  }
  fprintf(stderr, "usage: division_check exhaustive P | binary64 N SEED\n");
  return 2;
}
