// montgomery_replay.cpp -- test-only CPU check of the device's 64-bit Montgomery product
// (erdos-strauss_amd/csrc/es_montgomery.h, the same source the kernels compile) against
// 128-bit integer arithmetic: mont::mul(a, b, n) * 2^64 == a * b (mod n), with the result
// in [0, n), for edge moduli (3, 2^32 +- small, 2^62 +- small, 2^63 +- small, the largest
// odd values below 2^64) and random odd moduli of every bit length, with edge operands
// (0, 1, n - 1, n - 2, 2^32 - 1, 2^32) and random ones. Also mont::mul_small (x a mod n for
// multipliers a < 2^32, with its FP64 quotient estimate) against 128-bit arithmetic, and
// mont::strong_test (the windowed left-to-right strong probable-prime test) against a
// plain 128-bit right-to-left test for every base 2..37 on random odd moduli, primes and
// known strong pseudoprimes.
//
// usage: montgomery_replay TRIALS SEED     prints "ok TRIALS" or the first mismatch.
#include <cinttypes>
#include <cstdio>
#include <cstdlib>
#include <vector>

#include "es_montgomery.h"

typedef uint64_t u64;
typedef unsigned __int128 u128;

static u64 rng_state;
static u64 next() {
  rng_state ^= rng_state << 13;
  rng_state ^= rng_state >> 7;
  rng_state ^= rng_state << 17;
  return rng_state;
}

static bool check(u64 a, u64 b, u64 n) {
  const u64 ni = es::mont::neg_inv(n);
  if (n * (0 - ni) != 1) { printf("neg_inv wrong n=%" PRIu64 "\n", n); return false; }
  const u64 r = es::mont::mul(a, b, n, (uint32_t)ni);
  const bool ok = r < n && (((u128)r << 64) % n) == ((u128)a * b % n);
  if (!ok) printf("mismatch a=%" PRIu64 " b=%" PRIu64 " n=%" PRIu64 " got=%" PRIu64 "\n", a, b, n, r);
  return ok;
}

static bool check_small(u64 x, uint32_t a, u64 n) {
  const u64 r = es::mont::mul_small(x, a, n, 1.0 / (double)n);
  const bool ok = r == (u64)((u128)x * a % n);
  if (!ok) printf("mul_small mismatch x=%" PRIu64 " a=%u n=%" PRIu64 " got=%" PRIu64 "\n", x, a, n, r);
  return ok;
}

static u64 mulmod(u64 a, u64 b, u64 n) { return (u64)((u128)a * b % n); }

// Plain strong test: a^d by right-to-left squaring in 128-bit arithmetic.
static bool strong_ref(u64 n, u64 a, u64 d, uint32_t s) {
  u64 x = 1, b = a % n;
  for (u64 e = d; e; e >>= 1) {
    if (e & 1) x = mulmod(x, b, n);
    b = mulmod(b, b, n);
  }
  if (x == 1 || x == n - 1) return true;
  for (uint32_t i = 1; i < s; ++i) {
    x = mulmod(x, x, n);
    if (x == n - 1) return true;
  }
  return false;
}

static bool check_strong(u64 n) {
  u64 d = n - 1;
  const uint32_t s = (uint32_t)__builtin_ctzll(d);
  d >>= s;
  const u64 ni = es::mont::neg_inv(n), one = (u64)(((u128)1 << 64) % n);
  for (uint32_t a = 2; a <= 37; ++a) {
    if (a >= n) break;
    const bool got = es::mont::strong_test(n, a, d, s, (uint32_t)ni, one, 1.0 / (double)n);
    if (got != strong_ref(n, a, d, s)) {
      printf("strong_test mismatch n=%" PRIu64 " a=%u got=%d\n", n, a, (int)got);
      return false;
    }
  }
  return true;
}

int main(int argc, char **argv) {
  const long trials = argc > 1 ? atol(argv[1]) : 1000000;
  rng_state = argc > 2 ? strtoull(argv[2], 0, 10) * 2654435761ull + 1 : 1;
  std::vector<u64> mods = {3, 5, 7, 0xFFFFFFFFull, 0x100000001ull, 0x1FFFFFFFFull, (1ull << 62) - 57,
                           (1ull << 62) + 1, (1ull << 63) - 25, (1ull << 63) + 1, 18446744065119614953ull,
                           18446744073709551557ull, 0xFFFFFFFFFFFFFFFFull, 0xFFFFFFFFFFFFFFFDull};
  for (int bits = 2; bits <= 64; ++bits)
    for (int k = 0; k < 8; ++k) {
      u64 n = next() >> (64 - bits);
      n |= (bits == 64 ? 0 : 0) | 1 | (bits > 1 ? 1ull << (bits - 1) : 0);
      if (n >= 3) mods.push_back(n);
    }
  long done = 0;
  for (u64 n : mods) {
    const u64 edge[] = {0, 1, n - 1, n - 2, 0xFFFFFFFFull % n, 0x100000000ull % n, n / 2, n / 3};
    for (u64 a : edge)
      for (u64 b : edge) {
        if (!check(a, b, n)) return 1;
        ++done;
      }
  }
  for (long t = 0; t < trials; ++t) {
    const int bits = 2 + (int)(next() % 63);
    u64 n = (next() >> (64 - bits)) | 1 | (1ull << (bits - 1));
    if (n < 3) n = 3;
    const u64 a = next() % n, b = next() % n;
    if (!check(a, b, n)) return 1;
  }
  // mul_small: edge multipliers and operands on every modulus, then random ones.
  for (u64 n : mods) {
    if (!(n & 1)) continue;
    const u64 xs[] = {0, 1, n - 1, n - 2, n / 2, n / 3, 0xFFFFFFFFull % n};
    const uint32_t as[] = {1, 2, 3, 37, 1369, 50653, 1u << 31, 0xFFFFFFFFu, 0x80000001u};
    for (u64 x : xs)
      for (uint32_t a : as)
        if (!check_small(x, a, n)) return 1;
  }
  for (long t = 0; t < trials; ++t) {
    const int bits = 2 + (int)(next() % 63);
    u64 n = (next() >> (64 - bits)) | 1 | (1ull << (bits - 1));
    if (n < 3) n = 3;
    const uint32_t a = (uint32_t)(next() >> (32 + next() % 32)) | 1u;
    if (!check_small(next() % n, a, n)) return 1;
  }
  // strong_test: known strong pseudoprimes (to base 2; to bases 2..23 for the last),
  // primes near the top, and random odd moduli above 37.
  const u64 spsp[] = {2047ull, 3277ull, 4033ull, 3215031751ull, 2152302898747ull, 3474749660383ull,
                      341550071728321ull, 3825123056546413051ull, 18446744073709551557ull,
                      18446744073709551533ull, 18446744065119617025ull, (1ull << 61) - 1, (1ull << 50) + 1,
                      (1ull << 50) - 27, 1000000007ull};
  for (u64 n : spsp)
    if (!check_strong(n)) return 1;
  for (long t = 0; t < trials / 20; ++t) {
    const int bits = 7 + (int)(next() % 58);
    const u64 n = (next() >> (64 - bits)) | 1 | (1ull << (bits - 1));
    if (!check_strong(n)) return 1;
  }
  printf("ok %ld\n", trials);
  (void)done;
  return 0;
}
