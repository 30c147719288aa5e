// mulmod_replay.cpp -- test-only CPU replay of the stage-D FP64 modular multiply
// (x * y mod n for x, y < n < 2^50: product and FMA residual, floor quotient estimate,
// one exact FMA for h - q*n, one correction each way) against 128-bit integer
// arithmetic. IEEE-754 binary64 with a fused multiply-add is what the device executes,
// so std::fma gives the same results bit for bit.
//
// usage: mulmod_replay TRIALS SEED     prints "ok TRIALS" or the first mismatch.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <initializer_list>

typedef uint64_t u64;
typedef unsigned __int128 u128;

static u64 mulmod_f64(u64 x, u64 y, double nd, double ninv) {
  const double a = (double)x, b = (double)y;
  const double h = a * b;
  const double l = std::fma(a, b, -h);
  const double q = std::floor(h * ninv);
  double r = std::fma(-q, nd, h) + l;
  if (r < 0) r += nd;
  if (r >= nd) r -= nd;
  return (u64)r;
}

static u64 state;
static u64 next() {   // splitmix64
  u64 z = (state += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}

static int check(u64 x, u64 y, u64 n) {
  const double nd = (double)n, ninv = 1.0 / nd;
  const u64 got = mulmod_f64(x, y, nd, ninv), want = (u64)((u128)x * y % n);
  if (got != want) {
    printf("mismatch x=%llu y=%llu n=%llu got=%llu want=%llu\n", (unsigned long long)x, (unsigned long long)y,
           (unsigned long long)n, (unsigned long long)got, (unsigned long long)want);
    return 1;
  }
  return 0;
}

int main(int argc, char **argv) {
  const u64 trials = argc > 1 ? strtoull(argv[1], 0, 10) : 1000000;
  state = argc > 2 ? strtoull(argv[2], 0, 10) : 1;
  const u64 top = 1ull << 50;
  // Edges: the largest operands and moduli next to 2^32 and 2^50.
  const u64 ns[] = {(1ull << 32) + 15, top - 1, top - 27, top - 2, (1ull << 49) + 1, 4294967311ull * 65537ull};
  for (u64 n : ns)
    for (u64 x : {(u64)0, (u64)1, n - 1, n - 2, n / 2, n / 2 + 1})
      for (u64 y : {(u64)0, (u64)1, n - 1, n - 2, n / 2, n / 2 + 1})
        if (check(x, y, n)) return 1;
  for (u64 t = 0; t < trials; ++t) {
    const int bits = 33 + (int)(next() % 18);           // n in [2^32, 2^50)
    u64 n = (next() >> (64 - bits)) | (1ull << (bits - 1)) | 1;
    if (n >= top) n = top - 1;
    const u64 x = next() % n, y = (t & 7) == 0 ? n - 1 - (next() & 255) % n : next() % n;
    if (check(x, y, n)) return 1;
  }
  printf("ok %llu\n", (unsigned long long)trials);
  return 0;
}
