// mulmod_replay.cpp -- test-only CPU replay of the device's FP64 modular arithmetic
// (moduli below 2^50) against 128-bit integer arithmetic:
//   mulmod_sgn   the signed-residue multiply Miller-Rabin and rho use: a * b with
//                q = rint(a*b/n) gives
//                r == a*b (mod n) with |r| <= (1/2 + 3*2^-53*|a*b|/n) * n, exactly, for
//                |a*b| < 2^102 -- so |r| < n for |a|, |b| < n + 64 (the walk), |r| < 1.25 n
//                for |a| < n, |b| < 2n + 128, and |r| < 2n for |a| <= 2n, |b| < 2n (the
//                accumulated product of differences, which therefore needs no correction);
//   strong_probable_sgn  the strong probable-prime test run in those residues with no
//                correction anywhere in the chain, whose verdict must equal the test in
//                exact integers for every base 2 .. 37.
// IEEE-754 binary64 with a fused multiply-add is what the device executes, so std::fma
// and std::rint (round to nearest even) give the same results bit for bit.
//
// usage: mulmod_replay TRIALS SEED     prints "ok TRIALS" or the first mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>

typedef uint64_t u64;
typedef unsigned __int128 u128;

static double mulmod_sgn(double a, double b, double nd, double ninv) {
  const double h = a * b;
  const double l = std::fma(a, b, -h);
  const double q = std::rint(h * ninv);
  return std::fma(-q, nd, h) + l;
}

typedef __int128 i128;

static bool strong_probable_sgn(u64 n, unsigned a, u64 d, unsigned s) {
  const double nd = (double)n, ninv = 1.0 / nd, m1 = nd - 1.0, p1 = 1.0 - nd;
  double x = 1.0, b = (double)a;
  for (u64 e = d; e; e >>= 1) {
    if (e & 1) x = mulmod_sgn(x, b, nd, ninv);
    b = mulmod_sgn(b, b, nd, ninv);
  }
  if (x == 1.0 || x == p1 || x == -1.0 || x == m1) return true;
  for (unsigned i = 1; i < s; ++i) {
    x = mulmod_sgn(x, x, nd, ninv);
    if (x == -1.0 || x == m1) return true;
  }
  return false;
}

static bool strong_probable_int(u64 n, unsigned a, u64 d, unsigned s) {
  u64 x = 1, b = a % n;
  for (u64 e = d; e; e >>= 1) {
    if (e & 1) x = (u64)((u128)x * b % n);
    b = (u64)((u128)b * b % n);
  }
  if (x == 1 || x == n - 1) return true;
  for (unsigned i = 1; i < s; ++i) {
    x = (u64)((u128)x * x % n);
    if (x == n - 1) return true;
  }
  return false;
}

// r must be an integer-valued double with r == a*b (mod n) and |r| < bound * n.
static int check_sgn(long long a, long long b, u64 n, double bound) {
  const double nd = (double)n, ninv = 1.0 / nd;
  const double r = mulmod_sgn((double)a, (double)b, nd, ninv);
  const i128 ri = (i128)r;
  const bool integral = (double)ri == r;
  const i128 diff = (i128)a * b - ri;
  const bool congruent = diff % (i128)n == 0;
  const bool bounded = std::fabs(r) < bound * nd;
  if (!integral || !congruent || !bounded) {
    printf("signed mismatch a=%lld b=%lld n=%llu r=%.1f\n", a, b, (unsigned long long)n, r);
    return 1;
  }
  return 0;
}

static u64 state;
static u64 next() {   // splitmix64
  u64 z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

// The signed-residue test and the exact one agree on n for every base 2 .. 37.
static int check_mr(u64 n) {
  u64 d = n - 1;
  unsigned s = 0;
  while (!(d & 1)) { d >>= 1; ++s; }
  for (unsigned a : {2u, 3u, 5u, 7u, 11u, 13u, 17u, 19u, 23u, 29u, 31u, 37u})
    if (strong_probable_sgn(n, a, d, s) != strong_probable_int(n, a, d, s)) {
      printf("MR mismatch n=%llu a=%u\n", (unsigned long long)n, a);
      return 1;
    }
  return 0;
}

int main(int argc, char **argv) {
  const u64 trials = argc > 1 ? strtoull(argv[1], 0, 10) : 1000000;
  state = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
  const u64 top = 1ull << 50;
  const u64 ns[] = {(1ull << 32) + 15, top - 1, top - 27, top - 2, (1ull << 49) + 1, 4294967311ull * 65537ull};
  // Miller-Rabin chains: primes, strong pseudoprimes to several bases, products of two
  // primes above 2^16, and random odd moduli (most composite, some prime).
  const u64 mr_edges[] = {(1ull << 50) - 27, (1ull << 49) - 81, 4294967311ull, 3215031751ull,
                          2152302898747ull, 3474749660383ull, 341550071728321ull, 65537ull * 65539ull,
                          4294967291ull * 262139ull, 1125899906842597ull, 999999999999989ull};
  for (u64 n : mr_edges)
    if (check_mr(n)) return 1;
  for (u64 t = 0; t < trials / 8; ++t) {
    const int bits = 33 + (int)(next() % 18);
    u64 n = (next() >> (64 - bits)) | (1ull << (bits - 1)) | 1;
    if (n >= top) n = top - 1;
    if (check_mr(n)) return 1;
  }
  // Signed residues: the walk (|a|, |b| < n + 64) and the difference products.
  for (u64 n : ns) {
    const long long m = (long long)n;
    for (long long a : {0LL, 1LL, -1LL, m - 1, 1 - m, m + 63, -m - 63, m / 2, -m / 2})
      for (long long b : {0LL, 1LL, -1LL, m - 1, 1 - m, m + 63, -m - 63, m / 2 + 1, -m / 2 - 1})
        if (check_sgn(a, b, n, 1.0)) return 1;
    for (long long a : {1LL, m - 1, 1 - m, m / 3})
      for (long long b : {2 * m + 127, -2 * m - 127, 2 * m - 1, m + 1})
        if (check_sgn(a, b, n, 1.25)) return 1;
    for (long long a : {2 * m, -2 * m, 2 * m - 1, 1 - 2 * m, m, 1LL})
      for (long long b : {2 * m - 1, 1 - 2 * m, 2 * m - 3, m + 1, -m - 1})
        if (check_sgn(a, b, n, 2.0)) return 1;
  }
  for (u64 t = 0; t < trials; ++t) {
    const int bits = 33 + (int)(next() % 18);
    u64 n = (next() >> (64 - bits)) | (1ull << (bits - 1)) | 1;
    if (n >= top) n = top - 1;
    const long long m = (long long)n;
    const long long a = (long long)(next() % (2 * n + 127)) - m - 63;
    if ((t & 1) == 0) {
      const long long b = (long long)(next() % (2 * n + 127)) - m - 63;
      if (check_sgn(a, b, n, 1.0)) return 1;
    } else {
      const long long x = (long long)(next() % (2 * n - 1)) - m + 1;
      const long long d = (long long)(next() % (4 * n + 255)) - 2 * m - 127;
      if (check_sgn(x, d, n, 1.25)) return 1;
      const long long acc = (long long)(next() % (4 * n + 1)) - 2 * m;       // |acc| <= 2n
      const long long e = (long long)(next() % (4 * n - 1)) - 2 * m + 1;     // |e| < 2n
      if (check_sgn(acc, e, n, 2.0)) return 1;
    }
  }
  printf("ok %llu\n", (unsigned long long)trials);
  return 0;
}
