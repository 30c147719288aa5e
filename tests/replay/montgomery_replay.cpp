// montgomery_replay.cpp -- test-only CPU check of the device's 64-bit Montgomery product
// (erdos-strauss_amd/csrc/es_montgomery.h, the same source the kernels compile) against
// 128-bit integer arithmetic: mont::mul(a, b, n) * 2^64 == a * b (mod n), with the result
// in [0, n), for edge moduli (3, 2^32 +- small, 2^62 +- small, 2^63 +- small, the largest
// odd values below 2^64) and random odd moduli of every bit length, with edge operands
// (0, 1, n - 1, n - 2, 2^32 - 1, 2^32) and random ones.
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
  printf("ok %ld\n", trials);
  (void)done;
  return 0;
}
